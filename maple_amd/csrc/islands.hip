// Island correction of the multigrid preconditioner.
//
// OC designs grow small solid islands: clusters of a few voxels that touch the
// rest of the cell only through elements at the void modulus (1e-9 of solid).
// Their rigid and internal modes have energies the geometric V-cycle cannot
// see (its coarse grids average an island with the void around it), and every
// such mode costs CG iterations.  Here each small island's block of the level-0
// operator (its nodes, every element touching them, void ones included) is
// inverted once per density field, and every preconditioner application adds
//   z += A_i^{-1} r|_i
// on the island's nodes (an additive Schwarz term beside the V-cycle: the
// preconditioner stays symmetric positive definite, so CG still converges to
// the same tolerance; only its iteration count changes).  No counterpart in the
// reference, whose V-cycle (EBE_GMG.py:886-984) has the same blind spot.
//
// Detection (per density field, every cell at once):
//   solid element: rho > thr;
//   labels: each solid element starts with its own global id and takes the
//     smallest id among its solid 26-neighbours, MAPLE_ISL_ROUNDS times;
//   a label region is an island when no solid neighbour of any of its elements
//     carries another label (then the region is a whole connected component:
//     a component still being relabelled always has such a neighbour) and it
//     has at most MAPLE_ISL_MAXE elements and MAPLE_ISL_MAXN nodes.
// Islands never share nodes (a shared node would connect them), so their
// corrections touch disjoint entries of z.  Element and node lists are sorted
// before use: the result does not depend on the order atomics hand out slots.
#include "kernels.h"

namespace {

// element / node (x, y, z) of this thread in a level-0 node launch
__device__ __forceinline__ bool elem_of(const Launch& L, LevelCell& c, int& x, int& y, int& z, long long& gid) {
  const int k = find_cell(L.blk_start, L.ncells, lbid());
  c = L.cells[k];
  const long long n = (long long)(lbid() - L.blk_start[k]) * 256 + threadIdx.x;
  if (n >= (long long)c.nx * c.ny * c.nz) return false;
  z = (int)(n % c.nz);
  const long long t = n / c.nz;
  y = (int)(t % c.ny);
  x = (int)(t / c.ny);
  gid = c.off + n;
  return true;
}

__device__ __forceinline__ long long gidx(const LevelCell& c, int x, int y, int z) {
  return c.off + ((long long)wrap(x, c.nx) * c.ny + wrap(y, c.ny)) * c.nz + wrap(z, c.nz);
}

// periodic offset of b from a along one axis, in {-1, 0, 1}, or 2 if farther
__device__ __forceinline__ int pdiff(int a, int b, int n) {
  int d = b - a;
  if (d > n / 2) d -= n;
  if (d < -(n / 2)) d += n;
  return (d >= -1 && d <= 1) ? d : 2;
}

}  // namespace

__global__ __launch_bounds__(256) void k_isl_init(Launch L, const double* __restrict__ rho, double thr,
                                                  int* __restrict__ lab, int* __restrict__ cnt,
                                                  int* __restrict__ bad, int* __restrict__ slot) {
  LevelCell c; int x, y, z; long long g;
  if (!elem_of(L, c, x, y, z, g)) return;
  lab[g] = (rho[g] > thr && c.nx >= 4 && c.ny >= 4 && c.nz >= 4) ? (int)g : -1;
  cnt[g] = 0;
  bad[g] = 0;
  slot[g] = -1;
}

__global__ __launch_bounds__(256) void k_isl_prop(Launch L, const int* __restrict__ lin, int* __restrict__ lout) {
  LevelCell c; int x, y, z; long long g;
  if (!elem_of(L, c, x, y, z, g)) return;
  int m = lin[g];
  if (m >= 0) {
#pragma unroll
    for (int dx = -1; dx <= 1; ++dx)
#pragma unroll
      for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
        for (int dz = -1; dz <= 1; ++dz) {
          const int v = lin[gidx(c, x + dx, y + dy, z + dz)];
          if (v >= 0 && v < m) m = v;
        }
  }
  lout[g] = m;
}

__global__ __launch_bounds__(256) void k_isl_check(Launch L, const int* __restrict__ lab, int* __restrict__ bad,
                                                   int* __restrict__ cnt) {
  LevelCell c; int x, y, z; long long g;
  if (!elem_of(L, c, x, y, z, g)) return;
  const int m = lab[g];
  if (m < 0) return;
  bool open = false;
#pragma unroll
  for (int dx = -1; dx <= 1; ++dx)
#pragma unroll
    for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
      for (int dz = -1; dz <= 1; ++dz) {
        const int v = lab[gidx(c, x + dx, y + dy, z + dz)];
        open = open || (v >= 0 && v != m);
      }
  if (open) bad[m] = 1;
  atomicAdd(&cnt[m], 1);
}

__global__ __launch_bounds__(256) void k_isl_roots(Launch L, const int* __restrict__ lab, const int* __restrict__ bad,
                                                   const int* __restrict__ cnt, int* __restrict__ slot,
                                                   Island* __restrict__ isl, int* __restrict__ nisl, int maxi) {
  LevelCell c; int x, y, z; long long g;
  if (!elem_of(L, c, x, y, z, g)) return;
  if (lab[g] != (int)g || bad[g] || cnt[g] > MAPLE_ISL_MAXE) return;
  const int s = atomicAdd(nisl, 1);
  if (s >= maxi) return;
  slot[g] = s;
  isl[s].cell = c.id;
  isl[s].ne = 0;
  isl[s].nn = 0;
}

__global__ __launch_bounds__(256) void k_isl_members(Launch L, const int* __restrict__ lab,
                                                     const int* __restrict__ slot, Island* __restrict__ isl) {
  LevelCell c; int x, y, z; long long g;
  if (!elem_of(L, c, x, y, z, g)) return;
  const int m = lab[g];
  if (m < 0) return;
  const int s = slot[m];
  if (s < 0) return;
  const int k = atomicAdd(&isl[s].ne, 1);
  if (k < MAPLE_ISL_MAXE) isl[s].elems[k] = (int)g;
}

// One thread per island: sorted element list, sorted node list (nn = 0 when the
// island has more than MAPLE_ISL_MAXN nodes: it is left to the V-cycle).
__global__ __launch_bounds__(64) void k_isl_nodes(const LevelCell* __restrict__ cells, Island* __restrict__ isl,
                                                  int n) {
  const int s = blockIdx.x * 64 + threadIdx.x;
  if (s >= n) return;
  Island& I = isl[s];
  const int ne = min(I.ne, MAPLE_ISL_MAXE);
  int el[MAPLE_ISL_MAXE];
  for (int k = 0; k < ne; ++k) el[k] = I.elems[k];
  for (int a = 1; a < ne; ++a) {
    const int v = el[a];
    int b = a - 1;
    while (b >= 0 && el[b] > v) { el[b + 1] = el[b]; --b; }
    el[b + 1] = v;
  }
  const LevelCell c = cells[I.cell];
  int nd[MAPLE_ISL_MAXN];
  int nn = 0;
  bool ok = ne > 0 && ne == I.ne;
  for (int k = 0; k < ne && ok; ++k) {
    const long long e = el[k] - c.off;
    const int ez = (int)(e % c.nz), ey = (int)((e / c.nz) % c.ny), ex = (int)(e / ((long long)c.nz * c.ny));
    for (int q = 0; q < 8 && ok; ++q) {
      const int v = (int)gidx(c, ex + hex_dx(q), ey + hex_dy(q), ez + hex_dz(q));
      int p = 0;
      while (p < nn && nd[p] < v) ++p;
      if (p < nn && nd[p] == v) continue;
      if (nn == MAPLE_ISL_MAXN) { ok = false; break; }
      for (int t = nn; t > p; --t) nd[t] = nd[t - 1];
      nd[p] = v;
      ++nn;
    }
  }
  for (int k = 0; k < ne; ++k) I.elems[k] = el[k];
  for (int k = 0; k < nn; ++k) I.nodes[k] = nd[k];
  I.nn = ok ? nn : 0;
}

// One block per island: assemble the island's block of K (FP64, every element
// touching its nodes) in LDS and invert it by Gauss-Jordan (the block is SPD:
// no pivoting).  inv: per island (ND*MAXN)^2 doubles, the n x n inverse at the
// start (n = ND*nn).  A non-positive pivot (not expected) drops the island.
template <int ND>
__global__ __launch_bounds__(256) void k_isl_setup(const LevelCell* __restrict__ cells, Island* __restrict__ isl,
                                                   const double* __restrict__ Ke, const double* __restrict__ rho,
                                                   double* __restrict__ inv, int* __restrict__ valid) {
  constexpr int NM = ND * MAPLE_ISL_MAXN;
  constexpr int NE = 8 * ND;
  __shared__ double A[NM * NM];
  __shared__ double col[NM];
  __shared__ int nodes[MAPLE_ISL_MAXN];
  __shared__ int fail;
  const int s = blockIdx.x;
  const Island& I = isl[s];
  const int nn = I.nn;
  if (threadIdx.x == 0) valid[s] = 0;
  if (nn == 0) return;
  const LevelCell c = cells[I.cell];
  const double* __restrict__ K = Ke + (long long)c.id * NE * NE;
  const int n = ND * nn;
  if (threadIdx.x < nn) nodes[threadIdx.x] = I.nodes[threadIdx.x];
  if (threadIdx.x == 0) fail = 0;
  __syncthreads();
  // A[(ND a + i), (ND b + k)] = sum over elements e holding nodes a and b of rho_e Ke[ND ca + i][ND cb + k]
  for (int t = threadIdx.x; t < nn * nn; t += blockDim.x) {
    const int a = t / nn, b = t - (t / nn) * nn;
    const long long la = nodes[a] - c.off, lb = nodes[b] - c.off;
    const int az = (int)(la % c.nz), ay = (int)((la / c.nz) % c.ny), ax = (int)(la / ((long long)c.nz * c.ny));
    const int bz = (int)(lb % c.nz), by = (int)((lb / c.nz) % c.ny), bx = (int)(lb / ((long long)c.nz * c.ny));
    const int dx = pdiff(ax, bx, c.nx), dy = pdiff(ay, by, c.ny), dz = pdiff(az, bz, c.nz);
    double blk[ND][ND];
#pragma unroll
    for (int i = 0; i < ND; ++i)
#pragma unroll
      for (int k = 0; k < ND; ++k) blk[i][k] = 0.0;
    if (dx != 2 && dy != 2 && dz != 2) {
      for (int p = 0; p < 2; ++p)
        for (int q = 0; q < 2; ++q)
          for (int r = 0; r < 2; ++r) {
            const int mx = 1 - p + dx, my = 1 - q + dy, mz = 1 - r + dz;
            if (mx < 0 || mx > 1 || my < 0 || my > 1 || mz < 0 || mz > 1) continue;
            const int ca = hex_corner(1 - p, 1 - q, 1 - r), cb = hex_corner(mx, my, mz);
            const double re = rho[gidx(c, ax - 1 + p, ay - 1 + q, az - 1 + r)];
#pragma unroll
            for (int i = 0; i < ND; ++i)
#pragma unroll
              for (int k = 0; k < ND; ++k) blk[i][k] = fma(re, K[(ND * ca + i) * NE + ND * cb + k], blk[i][k]);
          }
    }
#pragma unroll
    for (int i = 0; i < ND; ++i)
#pragma unroll
      for (int k = 0; k < ND; ++k) A[(ND * a + i) * n + ND * b + k] = blk[i][k];
  }
  __syncthreads();
  // in-place Gauss-Jordan inversion
  for (int p = 0; p < n; ++p) {
    const double piv = A[p * n + p];
    if (!(piv > 0.0) || !isfinite(piv)) {
      if (threadIdx.x == 0) fail = 1;
      break;   // uniform: every thread reads the same pivot
    }
    for (int i = threadIdx.x; i < n; i += blockDim.x) col[i] = A[i * n + p];
    __syncthreads();
    const double d = 1.0 / piv;
    for (int j = threadIdx.x; j < n; j += blockDim.x) A[p * n + j] = (j == p ? 1.0 : A[p * n + j]) * d;
    __syncthreads();
    for (int t = threadIdx.x; t < n * n; t += blockDim.x) {
      const int i = t / n, j = t - (t / n) * n;
      if (i == p) continue;
      const double base = (j == p) ? 0.0 : A[i * n + j];
      A[i * n + j] = fma(-col[i], A[p * n + j], base);
    }
    __syncthreads();
  }
  __syncthreads();
  if (fail) return;
  double* __restrict__ out = inv + (long long)s * NM * NM;
  for (int t = threadIdx.x; t < n * n; t += blockDim.x) out[t] = A[t];
  if (threadIdx.x == 0) valid[s] = 1;
}

// z += A_i^{-1} r on island i for load case blockIdx.y (skipped once converged)
template <int ND, typename Tz>
__global__ __launch_bounds__(128) void k_isl_apply(const Island* __restrict__ isl, const int* __restrict__ valid,
                                                   const double* __restrict__ inv, long long Ntot,
                                                   const int* __restrict__ done, int ndc,
                                                   const double* __restrict__ r, Tz* __restrict__ z) {
  constexpr int NM = ND * MAPLE_ISL_MAXN;
  __shared__ double rv[NM];
  const int s = blockIdx.x, j = blockIdx.y;
  if (!valid[s]) return;
  const Island& I = isl[s];
  if (done != nullptr && done[I.cell * ndc + j]) return;
  const int n = ND * I.nn;
  for (int t = threadIdx.x; t < n; t += blockDim.x)
    rv[t] = r[(long long)(j * ND + t % ND) * Ntot + I.nodes[t / ND]];
  __syncthreads();
  const double* __restrict__ A = inv + (long long)s * NM * NM;
  for (int t = threadIdx.x; t < n; t += blockDim.x) {
    double acc = 0.0;
    for (int k = 0; k < n; ++k) acc = fma(A[t * n + k], rv[k], acc);
    Tz& zt = z[(long long)(j * ND + t % ND) * Ntot + I.nodes[t / ND]];
    zt = (Tz)((double)zt + acc);
  }
}

void island_detect(const Launch& L0, const double* rho, double thr, int* lab, int* lab2, int* cnt, int* bad,
                   int* slot, Island* isl, int* nisl, int maxi, hipStream_t s) {
  k_isl_init<<<L0.nblocks, 256, 0, s>>>(L0, rho, thr, lab, cnt, bad, slot);
  for (int k = 0; k < MAPLE_ISL_ROUNDS; k += 2) {
    k_isl_prop<<<L0.nblocks, 256, 0, s>>>(L0, lab, lab2);
    k_isl_prop<<<L0.nblocks, 256, 0, s>>>(L0, lab2, lab);
  }
  k_isl_check<<<L0.nblocks, 256, 0, s>>>(L0, lab, bad, cnt);
  k_isl_roots<<<L0.nblocks, 256, 0, s>>>(L0, lab, bad, cnt, slot, isl, nisl, maxi);
  k_isl_members<<<L0.nblocks, 256, 0, s>>>(L0, lab, slot, isl);
}

void island_nodes(const LevelCell* cells, Island* isl, int n, hipStream_t s) {
  if (n > 0) k_isl_nodes<<<(n + 63) / 64, 64, 0, s>>>(cells, isl, n);
}

void island_setup(int nd, const LevelCell* cells, Island* isl, int n, const double* Ke, const double* rho,
                  double* inv, int* valid, hipStream_t s) {
  if (n <= 0) return;
  if (nd == 3) k_isl_setup<3><<<n, 256, 0, s>>>(cells, isl, Ke, rho, inv, valid);
  else k_isl_setup<1><<<n, 256, 0, s>>>(cells, isl, Ke, rho, inv, valid);
}

template <typename Tz>
void island_apply_t(int nd, int ncase, const Island* isl, const int* valid, int n, const double* inv, long long Ntot,
                    const int* done, int ndc, const double* r, Tz* z, hipStream_t s) {
  if (n <= 0) return;
  dim3 grid(n, ncase);
  if (nd == 3) k_isl_apply<3, Tz><<<grid, 128, 0, s>>>(isl, valid, inv, Ntot, done, ndc, r, z);
  else k_isl_apply<1, Tz><<<grid, 128, 0, s>>>(isl, valid, inv, Ntot, done, ndc, r, z);
}

void island_apply(int nd, int ncase, const Island* isl, const int* valid, int n, const double* inv, long long Ntot,
                  const int* done, int ndc, const double* r, double* z, hipStream_t s) {
  island_apply_t<double>(nd, ncase, isl, valid, n, inv, Ntot, done, ndc, r, z, s);
}

// the mixed form's FP32 preconditioned residual
void island_apply(int nd, int ncase, const Island* isl, const int* valid, int n, const double* inv, long long Ntot,
                  const int* done, int ndc, const double* r, float* z, hipStream_t s) {
  island_apply_t<float>(nd, ncase, isl, valid, n, inv, Ntot, done, ndc, r, z, s);
}
