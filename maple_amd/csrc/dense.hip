// Coarsest multigrid level: exact solve with a dense SPD inverse.
//
// The periodic operator is singular (3 rigid translations for elasticity, the
// constant for heat).  The coarsest matrix is made SPD with the translation
// constraint, A' = A + gamma Z Z^T (Z the normalised translations, gamma the
// mean diagonal); for a right-hand side orthogonal to Z the solution of
// A' x = b is the minimum-norm solution of A x = b.  A'^-1 is formed once per
// density field by an in-place blocked Gauss-Jordan sweep (64 x 64 tiles, no
// pivoting, valid for SPD; the rank-64 updates of every step, 2 n^3 of its flops,
// on FP64 MFMA) and each V-cycle applies it as a dense mat-vec over
// all load cases, which keeps the coarsest solve a bandwidth-bound GEMV
// instead of two latency-bound triangular solves.
#include "kernels.h"

#define TB 64

// gamma[cell] = trace(A)/n from the centre blocks of the coarsest stencil.
template <typename T>
__global__ __launch_bounds__(256) void k_dense_gamma(const LevelCell* __restrict__ cells,
                                                     const DenseCell* __restrict__ dcells, long long Ntot,
                                                     int nd, const T* __restrict__ S, double* __restrict__ gamma) {
  const int ci = blockIdx.x;
  const LevelCell c = cells[ci];
  const long long nn = (long long)c.nx * c.ny * c.nz;
  double v[1] = {0.0};
  for (long long n = threadIdx.x; n < nn; n += blockDim.x)
    for (int i = 0; i < nd; ++i) v[0] += (double)S[(long long)(i * nd + i) * Ntot + c.off + n];   // centre block
  __shared__ double tot;
  block_sum<1>(v, &tot);
  if (threadIdx.x == 0) gamma[ci] = tot / (double)dcells[ci].n;
}

// A[(n,i),(m,j)] = sum_{d: n+d == m} S[d](i,j)(n) + gamma/N delta_ij; S in the half storage
// (blocks d >= 13 at the node, d < 13 transposed at the neighbour m)
template <typename T>
__global__ __launch_bounds__(256) void k_dense_assemble(const LevelCell* __restrict__ cells,
                                                        const DenseCell* __restrict__ dcells, long long Ntot,
                                                        int nd, const T* __restrict__ S,
                                                        const double* __restrict__ gamma, double* __restrict__ A) {
  const int ci = blockIdx.y;
  const LevelCell c = cells[ci];
  const DenseCell dc = dcells[ci];
  const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
  const long long n2 = (long long)dc.n * dc.n;
  if (e >= n2) return;
  const int row = (int)(e / dc.n), col = (int)(e - (long long)row * dc.n);
  const int nr = row / nd, ir = row - nr * nd;
  const int mc = col / nd, jc = col - mc * nd;
  const int nz = c.nz, ny = c.ny;
  const int zr = nr % nz, yr = (nr / nz) % ny, xr = nr / (nz * ny);
  const int zc = mc % nz, yc = (mc / nz) % ny, xc = mc / (nz * ny);
  double acc = 0.0;
  for (int dx = -1; dx <= 1; ++dx) {
    if (wrap(xr + dx, c.nx) != xc) continue;
    for (int dy = -1; dy <= 1; ++dy) {
      if (wrap(yr + dy, c.ny) != yc) continue;
      for (int dz = -1; dz <= 1; ++dz) {
        if (wrap(zr + dz, c.nz) != zc) continue;
        const int d = (dx + 1) * 9 + (dy + 1) * 3 + (dz + 1);
        acc += d >= 13 ? (double)S[(long long)((d - 13) * nd * nd + ir * nd + jc) * Ntot + c.off + nr]
                       : (double)S[(long long)((13 - d) * nd * nd + jc * nd + ir) * Ntot + c.off + mc];
      }
    }
  }
  const long long nn = (long long)c.nx * c.ny * c.nz;
  if (ir == jc) acc += gamma[ci] / (double)nn;
  A[dc.aoff + e] = acc;
}

template <typename T>
void dense_assemble_g(const Launch& Lc, const DenseCell* dcells, int maxn, int nd, const T* S, double* A,
                      double* gamma, hipStream_t s) {
  k_dense_gamma<T><<<Lc.ncells, 256, 0, s>>>(Lc.cells, dcells, Lc.Ntot, nd, S, gamma);
  dim3 grid((unsigned)(((long long)maxn * maxn + 255) / 256), Lc.ncells);
  k_dense_assemble<T><<<grid, 256, 0, s>>>(Lc.cells, dcells, Lc.Ntot, nd, S, gamma, A);
}
template void dense_assemble_g<double>(const Launch&, const DenseCell*, int, int, const double*, double*, double*, hipStream_t);
template void dense_assemble_g<float>(const Launch&, const DenseCell*, int, int, const float*, double*, double*, hipStream_t);

// ---- blocked Gauss-Jordan ------------------------------------------------------

__device__ __forceinline__ int tile_rows(const DenseCell& dc, int t) { return min(TB, dc.n - t * TB); }

// load tile (ti,tj) of cell matrix into LDS (zero padded)
__device__ __forceinline__ void load_tile(const DenseCell& dc, const double* __restrict__ A, int ti, int tj,
                                          double (*L)[TB]) {
  const int nr = tile_rows(dc, ti), nc = tile_rows(dc, tj);
  for (int e = threadIdx.x; e < TB * TB; e += blockDim.x) {
    const int r = e / TB, cc = e - r * TB;
    L[r][cc] = (r < nr && cc < nc) ? A[dc.aoff + (long long)(ti * TB + r) * dc.n + tj * TB + cc] : 0.0;
  }
}

// K1: invert the pivot tile (scalar Gauss-Jordan, no pivoting: SPD), write it to
// work[cell].  Each of the 256 threads keeps its 4 x 4 sub-block in registers;
// per pivot only row p and column p go through LDS (double-buffered, one
// barrier per pivot).
__global__ __launch_bounds__(256) void k_gj_diag(const DenseCell* __restrict__ dcells, int k,
                                                 const double* __restrict__ A, double* __restrict__ work) {
  const int ci = blockIdx.x;
  const DenseCell dc = dcells[ci];
  if (k >= dc.nt) return;
  __shared__ double rowb[2][TB], colb[2][TB];
  const int tr = threadIdx.x >> 4, tc = threadIdx.x & 15;   // 16 x 16 threads, 4 x 4 entries each
  const int nb = tile_rows(dc, k);
  double M[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int i = tr * 4 + a, j = tc * 4 + b;
      // identity in the padding so the padded pivots are 1
      M[a][b] = (i < nb && j < nb) ? A[dc.aoff + (long long)(k * TB + i) * dc.n + k * TB + j] : (i == j ? 1.0 : 0.0);
    }
  for (int p = 0; p < TB; ++p) {
    const int buf = p & 1;
    const int pa = p & 3, pt = p >> 2;
    if (tr == pt) {
#pragma unroll
      for (int b = 0; b < 4; ++b)
        if (pa == 0) rowb[buf][tc * 4 + b] = M[0][b];
        else if (pa == 1) rowb[buf][tc * 4 + b] = M[1][b];
        else if (pa == 2) rowb[buf][tc * 4 + b] = M[2][b];
        else rowb[buf][tc * 4 + b] = M[3][b];
    }
    if (tc == pt) {
#pragma unroll
      for (int a = 0; a < 4; ++a)
        if (pa == 0) colb[buf][tr * 4 + a] = M[a][0];
        else if (pa == 1) colb[buf][tr * 4 + a] = M[a][1];
        else if (pa == 2) colb[buf][tr * 4 + a] = M[a][2];
        else colb[buf][tr * 4 + a] = M[a][3];
    }
    __syncthreads();
    const double pv = 1.0 / rowb[buf][p];
    double rj[4], ci4[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) rj[b] = rowb[buf][tc * 4 + b];
#pragma unroll
    for (int a = 0; a < 4; ++a) ci4[a] = colb[buf][tr * 4 + a];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const int i = tr * 4 + a, j = tc * 4 + b;
        if (i == p && j == p) M[a][b] = pv;
        else if (i == p) M[a][b] = rj[b] * pv;
        else if (j == p) M[a][b] = -ci4[a] * pv;
        else M[a][b] = fma(-ci4[a] * pv, rj[b], M[a][b]);
      }
  }
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) work[(long long)ci * TB * TB + (tr * 4 + a) * TB + tc * 4 + b] = M[a][b];
}

// C(64x64) = X * Y with X, Y in LDS; each thread 4x4 outputs.
__device__ __forceinline__ void tile_mm(double (*X)[TB], double (*Y)[TB], double (&acc)[4][4]) {
  const int tr = threadIdx.x >> 4, tc = threadIdx.x & 15;
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
  for (int kk = 0; kk < TB; ++kk) {
    double xa[4], yb[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) xa[a] = X[tr * 4 + a][kk];
#pragma unroll
    for (int b = 0; b < 4; ++b) yb[b] = Y[kk][tc * 4 + b];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[a][b] = fma(xa[a], yb[b], acc[a][b]);
  }
}

__device__ __forceinline__ void store_tile(const DenseCell& dc, double* __restrict__ A, int ti, int tj,
                                           const double (&v)[4][4], double scale, bool accumulate) {
  const int tr = threadIdx.x >> 4, tc = threadIdx.x & 15;
  const int nr = tile_rows(dc, ti), nc = tile_rows(dc, tj);
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int r = tr * 4 + a, cc = tc * 4 + b;
      if (r < nr && cc < nc) {
        double* p = A + dc.aoff + (long long)(ti * TB + r) * dc.n + tj * TB + cc;
        *p = accumulate ? *p + scale * v[a][b] : scale * v[a][b];
      }
    }
}

__device__ __forceinline__ void load_work(const double* __restrict__ work, int ci, double (*L)[TB]) {
  for (int e = threadIdx.x; e < TB * TB; e += blockDim.x) L[e / TB][e % TB] = work[(long long)ci * TB * TB + e];
}

// K2: row panel A_kj <- P A_kj (j != k)
__global__ __launch_bounds__(256) void k_gj_row(const DenseCell* __restrict__ dcells, int k, double* __restrict__ A,
                                                const double* __restrict__ work) {
  const int ci = blockIdx.y, j = blockIdx.x;
  const DenseCell dc = dcells[ci];
  if (k >= dc.nt || j >= dc.nt || j == k) return;
  __shared__ double P[TB][TB];
  __shared__ double B[TB][TB];
  load_work(work, ci, P);
  load_tile(dc, A, k, j, B);
  __syncthreads();
  double acc[4][4];
  tile_mm(P, B, acc);
  store_tile(dc, A, k, j, acc, 1.0, false);
}

// K3 on the matrix cores: A_ij -= A_ik A_kj (i != k, j != k; A_kj already updated),
// a 64 x 64 x 64 tile product with v_mfma_f64_16x16x4 (FP64 throughout; the sums
// differ from scalar FMAs only in their order).  The operand tiles are staged in LDS
// with rows padded by one double (the A fragments read 16 rows of one column);
// wave w computes the 32 x 32 quarter (w / 2, w % 2) as 2 x 2 blocks of 16 x 16.
// Fragments of the 16x16x4 product: A (16 x 4), lane l holds row l % 16, column
// l / 16; B (4 x 16), lane l holds row l / 16, column l % 16; D (16 x 16), lane l
// holds column l % 16, rows l / 16 + 4 r in its four values (checked on gfx950).
typedef double dbl4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void load_tile_pad(const DenseCell& dc, const double* __restrict__ A, int ti, int tj,
                                              double (*L)[TB + 1]) {
  const int nr = tile_rows(dc, ti), nc = tile_rows(dc, tj);
  for (int e = threadIdx.x; e < TB * TB; e += blockDim.x) {
    const int r = e / TB, cc = e - r * TB;
    L[r][cc] = (r < nr && cc < nc) ? A[dc.aoff + (long long)(ti * TB + r) * dc.n + tj * TB + cc] : 0.0;
  }
}

__global__ __launch_bounds__(256) void k_gj_update(const DenseCell* __restrict__ dcells, int k,
                                                   double* __restrict__ A) {
  const int ci = blockIdx.z, i = blockIdx.y, j = blockIdx.x;
  const DenseCell dc = dcells[ci];
  if (k >= dc.nt || i >= dc.nt || j >= dc.nt || i == k || j == k) return;
  __shared__ double X[TB][TB + 1];
  __shared__ double Y[TB][TB + 1];
  load_tile_pad(dc, A, i, k, X);
  load_tile_pad(dc, A, k, j, Y);
  __syncthreads();
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int r0 = (w >> 1) * 32, c0 = (w & 1) * 32;
  const int li = l & 15, lk = l >> 4;
  dbl4 acc[2][2];
#pragma unroll
  for (int rb = 0; rb < 2; ++rb)
#pragma unroll
    for (int cb = 0; cb < 2; ++cb) acc[rb][cb] = dbl4{0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
  for (int kk = 0; kk < TB; kk += 4) {
    double a[2], b[2];
#pragma unroll
    for (int rb = 0; rb < 2; ++rb) a[rb] = X[r0 + rb * 16 + li][kk + lk];
#pragma unroll
    for (int cb = 0; cb < 2; ++cb) b[cb] = Y[kk + lk][c0 + cb * 16 + li];
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
      for (int cb = 0; cb < 2; ++cb) acc[rb][cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[rb], b[cb], acc[rb][cb], 0, 0, 0);
  }
  const int nr = tile_rows(dc, i), nc = tile_rows(dc, j);
#pragma unroll
  for (int rb = 0; rb < 2; ++rb)
#pragma unroll
    for (int cb = 0; cb < 2; ++cb)
#pragma unroll
      for (int rr = 0; rr < 4; ++rr) {
        const int row = r0 + rb * 16 + lk + 4 * rr, col = c0 + cb * 16 + li;
        if (row < nr && col < nc) {
          double* p = A + dc.aoff + (long long)(i * TB + row) * dc.n + j * TB + col;
          *p -= acc[rb][cb][rr];
        }
      }
}

// K4: A_ik <- -A_ik P (i != k), A_kk <- P
__global__ __launch_bounds__(256) void k_gj_col(const DenseCell* __restrict__ dcells, int k, double* __restrict__ A,
                                                const double* __restrict__ work) {
  const int ci = blockIdx.y, i = blockIdx.x;
  const DenseCell dc = dcells[ci];
  if (k >= dc.nt || i >= dc.nt) return;
  __shared__ double P[TB][TB];
  __shared__ double X[TB][TB];
  load_work(work, ci, P);
  if (i == k) {
    __syncthreads();
    const int tr = threadIdx.x >> 4, tc = threadIdx.x & 15;
    double v[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) v[a][b] = P[tr * 4 + a][tc * 4 + b];
    store_tile(dc, A, k, k, v, 1.0, false);
    return;
  }
  load_tile(dc, A, i, k, X);
  __syncthreads();
  double acc[4][4];
  tile_mm(X, P, acc);
  store_tile(dc, A, i, k, acc, -1.0, false);
}

void dense_inverse(const DenseCell* dcells, int ncells, int max_nt, double* A, double* work, hipStream_t s) {
  for (int k = 0; k < max_nt; ++k) {
    k_gj_diag<<<ncells, 256, 0, s>>>(dcells, k, A, work);
    k_gj_row<<<dim3(max_nt, ncells), 256, 0, s>>>(dcells, k, A, work);
    k_gj_update<<<dim3(max_nt, max_nt, ncells), 256, 0, s>>>(dcells, k, A);
    k_gj_col<<<dim3(max_nt, ncells), 256, 0, s>>>(dcells, k, A, work);
  }
}

// x_j = Ainv b_j for every case; one block of 256 threads per matrix row.
template <int ND, typename T>
__global__ __launch_bounds__(256) void k_dense_apply(const LevelCell* __restrict__ cells,
                                                     const DenseCell* __restrict__ dcells, long long Ntot,
                                                     int ncase, const double* __restrict__ Ainv,
                                                     const T* __restrict__ b, T* __restrict__ x,
                                                     const int* __restrict__ done) {
  const int ci = cells[blockIdx.y].id;
  const DenseCell dc = dcells[ci];
  const int row = blockIdx.x;
  if (row >= dc.n) return;
  if (cell_done(done, ncase, ci)) return;
  double acc[MAPLE_MAX_CASES];
#pragma unroll
  for (int j = 0; j < MAPLE_MAX_CASES; ++j) acc[j] = 0.0;
  const double* Ar = Ainv + dc.aoff + (long long)row * dc.n;
  for (int col = threadIdx.x; col < dc.n; col += 256) {
    const double a = Ar[col];
    const int m = col / ND, i = col - m * ND;
#pragma unroll
    for (int j = 0; j < MAPLE_MAX_CASES; ++j)
      if (j < ncase) acc[j] = fma(a, (double)b[(long long)(j * ND + i) * Ntot + dc.voff + m], acc[j]);
  }
  __shared__ double red[MAPLE_MAX_CASES];
  block_sum<MAPLE_MAX_CASES>(acc, red);
  const int m = row / ND, i = row - m * ND;
  if (threadIdx.x < ncase) x[(long long)(threadIdx.x * ND + i) * Ntot + dc.voff + m] = (T)red[threadIdx.x];
}

// The same product with RB rows per block: each thread reads a right-hand-side
// column entry once for RB rows (the one-row form reads the whole right-hand side
// once per row, several times the bytes of the matrix itself).
template <int ND, int RB, typename T>
__global__ __launch_bounds__(256) void k_dense_apply_rb(const LevelCell* __restrict__ cells,
                                                        const DenseCell* __restrict__ dcells, long long Ntot,
                                                        int ncase, const double* __restrict__ Ainv,
                                                        const T* __restrict__ b, T* __restrict__ x,
                                                        const int* __restrict__ done) {
  const int ci = cells[blockIdx.y].id;
  const DenseCell dc = dcells[ci];
  const int row0 = blockIdx.x * RB;
  if (row0 >= dc.n) return;
  if (cell_done(done, ncase, ci)) return;
  const int nr = min(RB, dc.n - row0);
  double acc[RB * MAPLE_MAX_CASES];
#pragma unroll
  for (int q = 0; q < RB * MAPLE_MAX_CASES; ++q) acc[q] = 0.0;
  const double* __restrict__ A0 = Ainv + dc.aoff + (long long)row0 * dc.n;
  for (int col = threadIdx.x; col < dc.n; col += 256) {
    const int m = col / ND, i = col - m * ND;
    double bv[MAPLE_MAX_CASES];
#pragma unroll
    for (int j = 0; j < MAPLE_MAX_CASES; ++j)
      bv[j] = j < ncase ? (double)b[(long long)(j * ND + i) * Ntot + dc.voff + m] : 0.0;
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      if (r >= nr) break;   // uniform over the block
      const double a = A0[(long long)r * dc.n + col];
#pragma unroll
      for (int j = 0; j < MAPLE_MAX_CASES; ++j) acc[r * MAPLE_MAX_CASES + j] = fma(a, bv[j], acc[r * MAPLE_MAX_CASES + j]);
    }
  }
  __shared__ double red[RB * MAPLE_MAX_CASES];
  block_sum<RB * MAPLE_MAX_CASES>(acc, red);
  if ((int)threadIdx.x < nr * ncase) {
    const int r = threadIdx.x / ncase, j = threadIdx.x - r * ncase;
    const int row = row0 + r;
    const int m = row / ND, i = row - m * ND;
    x[(long long)(j * ND + i) * Ntot + dc.voff + m] = (T)red[r * MAPLE_MAX_CASES + j];
  }
}

template <typename T>
void dense_apply(const Launch& Lc, const DenseCell* dcells, int max_n, int nd, int ncase, const double* Ainv,
                 const T* b, T* x, hipStream_t s) {
  constexpr int RB = 8;
  dim3 grid((max_n + RB - 1) / RB, Lc.ncells);
  if (nd == 3) k_dense_apply_rb<3, RB, T><<<grid, 256, 0, s>>>(Lc.cells, dcells, Lc.Ntot, ncase, Ainv, b, x, Lc.done);
  else k_dense_apply_rb<1, RB, T><<<grid, 256, 0, s>>>(Lc.cells, dcells, Lc.Ntot, ncase, Ainv, b, x, Lc.done);
}
template void dense_apply<double>(const Launch&, const DenseCell*, int, int, int, const double*, const double*, double*, hipStream_t);
template void dense_apply<float>(const Launch&, const DenseCell*, int, int, int, const double*, const float*, float*, hipStream_t);
