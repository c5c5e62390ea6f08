// Homogenized tensors, element sensitivities, the periodic box filter and the
// OC update.
//
// * element_energy: S_e = L_e^T Ke L_e, L_e = X0 - u_e (model.py:39-43, 77-93),
//   stored as the packed upper triangle per element, and C_H = sum rho_e S_e
//   (volume 1).  At convergence dC_H/drho_e = S_e (the energy form is the
//   minimum over periodic fields), which is what the reference's autograd
//   through the solve approaches (main_IHD_v3.py:688).
// * sensitivity: dJ/drho_e = <G, S_e> with G = dJ/dC_H (and dJ/dkappa).
// * box_filter: periodic k^3 mean (main_IHD_v3.py:342-359).
// * OC: the reference's 60-step bisection (main_IHD_v3.py:362-380).  Every
//   pass evaluates the complete depth-6 subtree of midpoints (63 values of
//   lambda, each computed exactly as the sequential bisection forms it) in one
//   sweep over the elements, then a single wavefront walks the 6 predicates;
//   10 passes reproduce the 60 sequential steps with 20 launches.
#include "kernels.h"

__device__ __forceinline__ int packed(int i, int j) { return j * (j + 1) / 2 + i; }   // i <= j

template <int ND>
__global__ __launch_bounds__(256) void k_element_energy(Launch L, int ncase, const double* __restrict__ Ke,
                                                        const double* __restrict__ X0, const double* __restrict__ rho,
                                                        const double* __restrict__ u, double* __restrict__ Se,
                                                        double* __restrict__ partial) {
  constexpr int NE = 8 * ND;
  constexpr int NU = MAPLE_MAX_CASES * (MAPLE_MAX_CASES + 1) / 2;
  const int ci = find_cell(L.blk_start, L.ncells, lbid());
  const LevelCell c = L.cells[ci];
  const long long n = (long long)(lbid() - L.blk_start[ci]) * 256 + threadIdx.x;
  const bool ok = n < (long long)c.nx * c.ny * c.nz;
  Ke += (long long)ci * NE * NE;
  X0 += (long long)ci * NE * ncase;
  double acc[NU];
#pragma unroll
  for (int p = 0; p < NU; ++p) acc[p] = 0.0;
  if (ok) {
    const int z = (int)(n % c.nz), y = (int)((n / c.nz) % c.ny), x = (int)(n / ((long long)c.nz * c.ny));
    long long nodes[8];
#pragma unroll
    for (int k = 0; k < 8; ++k)
      nodes[k] = c.off + ((long long)wrap(x + hex_dx(k), c.nx) * c.ny + wrap(y + hex_dy(k), c.ny)) * c.nz +
                 wrap(z + hex_dz(k), c.nz);
    const double r = rho[c.off + n];
    const long long e = c.off + n;
    for (int j = 0; j < ncase; ++j) {
      double Lj[NE];
#pragma unroll
      for (int k = 0; k < 8; ++k)
#pragma unroll
        for (int d = 0; d < ND; ++d)
          Lj[ND * k + d] = X0[(ND * k + d) * ncase + j] - u[(long long)(j * ND + d) * L.Ntot + nodes[k]];
      double W[NE];
#pragma unroll
      for (int a = 0; a < NE; ++a) {
        double t = 0.0;
#pragma unroll
        for (int b = 0; b < NE; ++b) t = fma(Ke[a * NE + b], Lj[b], t);
        W[a] = t;
      }
      for (int i = 0; i <= j; ++i) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < 8; ++k)
#pragma unroll
          for (int d = 0; d < ND; ++d) {
            const double Li = X0[(ND * k + d) * ncase + i] - u[(long long)(i * ND + d) * L.Ntot + nodes[k]];
            s = fma(Li, W[ND * k + d], s);
          }
        const int p = packed(i, j);
        Se[(long long)p * L.Ntot + e] = s;
        // p < NU always; the switch keeps acc[] in registers
#pragma unroll
        for (int q = 0; q < NU; ++q)
          if (q == p) acc[q] = r * s;
      }
    }
  }
  const int nu = ncase * (ncase + 1) / 2;
  __shared__ double red[4][NU];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NU; ++q) {
    const double v = wave_sum(acc[q]);
    if (lane == 0) red[w][q] = v;
  }
  __syncthreads();
  if (threadIdx.x < nu)
    partial[(long long)lbid() * NU + threadIdx.x] =
        red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
}

void element_energy(const Launch& L, int nd, int ncase, const double* Ke, const double* X0, const double* rho,
                    const double* u, double* Se, double* partial, hipStream_t s) {
  if (nd == 3) k_element_energy<3><<<L.nblocks, 256, 0, s>>>(L, ncase, Ke, X0, rho, u, Se, partial);
  else k_element_energy<1><<<L.nblocks, 256, 0, s>>>(L, ncase, Ke, X0, rho, u, Se, partial);
}

// one block of 64 per cell: CH[cell][nc][nc] (full symmetric)
__global__ __launch_bounds__(64) void k_reduce_energy(const int* __restrict__ blk_start, const double* __restrict__ partial,
                                                      int nc, double* __restrict__ CH) {
  constexpr int NU = MAPLE_MAX_CASES * (MAPLE_MAX_CASES + 1) / 2;
  const int ci = blockIdx.x;
  const int b0 = blk_start[ci], b1 = blk_start[ci + 1];
  const int nu = nc * (nc + 1) / 2;
  for (int p = 0; p < nu; ++p) {
    double v = 0.0;
    for (int b = b0 + (int)threadIdx.x; b < b1; b += 64) v += partial[(long long)b * NU + p];
    v = wave_sum(v);
    if (threadIdx.x == 0) {
      int j = 0;
      while ((j + 1) * (j + 2) / 2 <= p) ++j;
      const int i = p - j * (j + 1) / 2;
      CH[(long long)ci * nc * nc + i * nc + j] = v;
      CH[(long long)ci * nc * nc + j * nc + i] = v;
    }
  }
}

void reduce_energy(const int* blk_start, int ncells, int nu, const double* partial, int nc, double* CH,
                   hipStream_t s) {
  (void)nu;
  k_reduce_energy<<<ncells, 64, 0, s>>>(blk_start, partial, nc, CH);
}

// dJ/drho_e = sum_{i<=j} g_ij S_e(i,j), g_ii = G_ii, g_ij = G_ij + G_ji
__global__ __launch_bounds__(256) void k_sensitivity(Launch L, int nc3, const double* __restrict__ Se3,
                                                     const double* __restrict__ G6, int nc1,
                                                     const double* __restrict__ Se1, const double* __restrict__ G3,
                                                     double* __restrict__ out) {
  const int ci = find_cell(L.blk_start, L.ncells, lbid());
  const LevelCell c = L.cells[ci];
  const long long n = (long long)(lbid() - L.blk_start[ci]) * 256 + threadIdx.x;
  if (n >= (long long)c.nx * c.ny * c.nz) return;
  const long long e = c.off + n;
  double v = 0.0;
  for (int j = 0; j < nc3; ++j)
    for (int i = 0; i <= j; ++i) {
      const double* G = G6 + (long long)ci * nc3 * nc3;
      const double g = i == j ? G[i * nc3 + i] : G[i * nc3 + j] + G[j * nc3 + i];
      v = fma(g, Se3[(long long)packed(i, j) * L.Ntot + e], v);
    }
  for (int j = 0; j < nc1; ++j)
    for (int i = 0; i <= j; ++i) {
      const double* G = G3 + (long long)ci * nc1 * nc1;
      const double g = i == j ? G[i * nc1 + i] : G[i * nc1 + j] + G[j * nc1 + i];
      v = fma(g, Se1[(long long)packed(i, j) * L.Ntot + e], v);
    }
  out[e] = v;
}

void sensitivity(const Launch& L, int nc3, const double* Se3, const double* G6, int nc1, const double* Se1,
                 const double* G3, double* dJdrho, hipStream_t s) {
  k_sensitivity<<<L.nblocks, 256, 0, s>>>(L, nc3, Se3, G6, nc1, Se1, G3, dJdrho);
}

__global__ __launch_bounds__(256) void k_box_filter(Launch L, int k, const double* __restrict__ in,
                                                    double* __restrict__ out) {
  const int ci = find_cell(L.blk_start, L.ncells, lbid());
  const LevelCell c = L.cells[ci];
  const long long n = (long long)(lbid() - L.blk_start[ci]) * 256 + threadIdx.x;
  if (n >= (long long)c.nx * c.ny * c.nz) return;
  const int z = (int)(n % c.nz), y = (int)((n / c.nz) % c.ny), x = (int)(n / ((long long)c.nz * c.ny));
  const int h = k / 2;
  double s = 0.0;
  for (int dx = -h; dx <= h; ++dx) {
    const int X = ((x + dx) % c.nx + c.nx) % c.nx;
    for (int dy = -h; dy <= h; ++dy) {
      const int Y = ((y + dy) % c.ny + c.ny) % c.ny;
      for (int dz = -h; dz <= h; ++dz) {
        const int Z = ((z + dz) % c.nz + c.nz) % c.nz;
        s += in[c.off + ((long long)X * c.ny + Y) * c.nz + Z];
      }
    }
  }
  out[c.off + n] = s / (double)(k * k * k);
}

void box_filter(const Launch& L, int k, const double* in, double* out, hipStream_t s) {
  k_box_filter<<<L.nblocks, 256, 0, s>>>(L, k, in, out);
}

// Symmetry projection of a design field (parse_symmetry / symmetrize_vec,
// main_IHD_v3.py:221-337): the mean of the field over a group of index maps of
// the cell.  kind 1: reflections about the centre planes of the axes in `axes`
// (bit 0 x, bit 1 y, bit 2 z; 2^k images, x -> n-1-x); kind 2: the 48 symmetries
// of the cube (6 axis permutations x 8 reflections; cubic cells).  Summing over
// the whole group makes the result independent of how its elements are listed.
__global__ __launch_bounds__(256) void k_symmetrize(Launch L, int kind, int axes, const double* __restrict__ in,
                                                    double* __restrict__ out) {
  const int ci = find_cell(L.blk_start, L.ncells, lbid());
  const LevelCell c = L.cells[ci];
  const long long n = (long long)(lbid() - L.blk_start[ci]) * 256 + threadIdx.x;
  if (n >= (long long)c.nx * c.ny * c.nz) return;
  const int z = (int)(n % c.nz), y = (int)((n / c.nz) % c.ny), x = (int)(n / ((long long)c.nz * c.ny));
  const int dims[3] = {c.nx, c.ny, c.nz};
  const int q[3] = {x, y, z};
  double s = 0.0;
  int cnt = 0;
  if (kind == 1) {
    for (int m = 0; m < 8; ++m) {
      if (m & ~axes) continue;
      int v[3];
      for (int a = 0; a < 3; ++a) v[a] = (m >> a & 1) ? dims[a] - 1 - q[a] : q[a];
      s += in[c.off + ((long long)v[0] * c.ny + v[1]) * c.nz + v[2]];
      ++cnt;
    }
  } else {
    const int perm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    for (int p = 0; p < 6; ++p)
      for (int m = 0; m < 8; ++m) {
        int v[3];
        for (int a = 0; a < 3; ++a) {
          const int t = q[perm[p][a]];
          v[a] = (m >> a & 1) ? c.nx - 1 - t : t;
        }
        s += in[c.off + ((long long)v[0] * c.nx + v[1]) * c.nx + v[2]];
        ++cnt;
      }
  }
  out[c.off + n] = s / (double)cnt;
}

void symmetrize(const Launch& L, int kind, int axes, const double* in, double* out, hipStream_t s) {
  k_symmetrize<<<L.nblocks, 256, 0, s>>>(L, kind, axes, in, out);
}

// ---- OC ----------------------------------------------------------------------

// Midpoint of heap node t (1..63) of the depth-6 subtree rooted at bracket (l1, l2):
// node t has children 2t (l2 <- lam_t) and 2t+1 (l1 <- lam_t); walking the bits of
// t from the top reproduces the sequential bisection's arithmetic exactly.
__device__ __forceinline__ double oc_node_lam(double l1, double l2, int t) {
  int depth = 31 - __clz(t);
  double lo = l1, hi = l2;
  for (int lev = depth - 1; lev >= 0; --lev) {
    const double lam = 0.5 * (lo + hi);
    if ((t >> lev) & 1) lo = lam; else hi = lam;
  }
  return 0.5 * (lo + hi);
}

// one wavefront per cell: lane t fills midpoint t
__global__ __launch_bounds__(64) void k_oc_start(OCState* __restrict__ st) {
  OCState& s = st[blockIdx.x];
  const int t = threadIdx.x;
  const double l1 = 1e-12, l2 = 1e6;
  const double lam = t == 0 ? 0.0 : oc_node_lam(l1, l2, t);
  s.lam[t] = lam;
  s.c[t] = t == 0 ? 0.0 : 1.0 / sqrt(lam);
  if (t == 0) {
    s.l1 = l1;
    s.l2 = l2;
    s.last_lam = 0.0;
  }
}

void oc_start(int ncells, OCState* st, hipStream_t s) { k_oc_start<<<ncells, 64, 0, s>>>(st); }

// a = x sqrt(max(dJ, 1e-12)) (the lambda-free factor of x sqrt(dJ/lambda))
__global__ __launch_bounds__(256) void k_oc_prepare(Launch L, const double* __restrict__ x,
                                                    const double* __restrict__ dJ, double* __restrict__ a) {
  const int ci = find_cell(L.blk_start, L.ncells, lbid());
  const LevelCell c = L.cells[ci];
  const long long n = (long long)(lbid() - L.blk_start[ci]) * 256 + threadIdx.x;
  if (n >= (long long)c.nx * c.ny * c.nz) return;
  const long long e = c.off + n;
  a[e] = x[e] * sqrt(fmax(dJ[e], 1e-12));
}

void oc_prepare(const Launch& L, const double* x, const double* dJ, double* a, hipStream_t s) {
  k_oc_prepare<<<L.nblocks, 256, 0, s>>>(L, x, dJ, a);
}

__device__ __forceinline__ double oc_value(double a, double x, double c, double move, double xmin) {
  double v = a * c;
  v = fmin(fmax(v, x - move), x + move);   // torch.clamp(v, x - move, x + move)
  return fmin(fmax(v, xmin), 1.0);         // torch.clamp(v, xmin, 1)
}

// a = x sqrt(max(dJ, 1e-12)) of one element: read from the prepared array, or (FROM_DJ)
// formed from dJ on the fly, to the same bits
template <bool FROM_DJ>
__device__ __forceinline__ double oc_a(const double* __restrict__ a_or_dJ, double xv, long long e) {
  if constexpr (FROM_DJ) return xv * sqrt(fmax(a_or_dJ[e], 1e-12));
  else return a_or_dJ[e];
}

// per block: partial[blk*64 + t] = sum over the block's elements of x'(lam_t)
template <bool FROM_DJ>
__global__ __launch_bounds__(256) void k_oc_pass(Launch L, const OCState* __restrict__ st,
                                                 const double* __restrict__ x, const double* __restrict__ a,
                                                 double* __restrict__ partial) {
  const int ci = find_cell(L.blk_start, L.ncells, lbid());
  const LevelCell c = L.cells[ci];
  const OCState& s = st[ci];
  const long long nn = (long long)c.nx * c.ny * c.nz;
  const long long n0 = (long long)(lbid() - L.blk_start[ci]) * 1024;
  double acc[64];
#pragma unroll
  for (int t = 0; t < 64; ++t) acc[t] = 0.0;
  const double move = s.move, xmin = s.xmin;
  for (int q = 0; q < 4; ++q) {
    const long long n = n0 + q * 256 + threadIdx.x;
    if (n < nn) {
      const double xv = x[c.off + n], av = oc_a<FROM_DJ>(a, xv, c.off + n);
#pragma unroll
      for (int t = 1; t < 64; ++t) acc[t] += oc_value(av, xv, s.c[t], move, xmin);
    }
  }
  __shared__ double red[4][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int t = 1; t < 64; ++t) {
    const double v = wave_sum(acc[t]);
    if (lane == 0) red[w][t] = v;
  }
  __syncthreads();
  if (threadIdx.x < 64)
    partial[(long long)lbid() * 64 + threadIdx.x] =
        threadIdx.x == 0 ? 0.0 : red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
}

void oc_pass(const Launch& L, const OCState* st, const double* x, const double* a, double* partial, hipStream_t s) {
  k_oc_pass<false><<<L.nblocks, 256, 0, s>>>(L, st, x, a, partial);
}

void oc_pass_dj(const Launch& L, const OCState* st, const double* x, const double* dJ, double* partial,
                hipStream_t s) {
  k_oc_pass<true><<<L.nblocks, 256, 0, s>>>(L, st, x, dJ, partial);
}

// one block of 1024 per cell: the 63 candidate means (16 slices of blocks per
// candidate, reduced in a fixed order), one lane walks `depth` levels, then the
// 64 lanes fill the next subtree.
__global__ __launch_bounds__(1024) void k_oc_walk(const int* __restrict__ blk_start, const LevelCell* __restrict__ cells,
                                                  int depth, const double* __restrict__ partial,
                                                  OCState* __restrict__ st) {
  const int ci = blockIdx.x;
  const LevelCell c = cells[ci];
  const int b0 = blk_start[ci], b1 = blk_start[ci + 1];
  const int t = threadIdx.x & 63, sl = threadIdx.x >> 6;
  double sum = 0.0;
  for (int b = b0 + sl; b < b1; b += 16) sum += partial[(long long)b * 64 + t];
  __shared__ double part[16][64];
  __shared__ double bracket[2];
  part[sl][t] = sum;
  __syncthreads();
  OCState& s = st[ci];
  if (threadIdx.x == 0) {
    const double nn = (double)((long long)c.nx * c.ny * c.nz);
    int node = 1;
    double l1 = s.l1, l2 = s.l2, lam = s.last_lam;
    for (int lev = 0; lev < depth; ++lev) {
      double tot = 0.0;
      for (int k = 0; k < 16; ++k) tot += part[k][node];
      lam = s.lam[node];
      const bool above = tot / nn > s.volfrac;
      if (above) l1 = lam; else l2 = lam;
      node = 2 * node + (above ? 1 : 0);
    }
    s.l1 = l1;
    s.l2 = l2;
    s.last_lam = lam;
    bracket[0] = l1;
    bracket[1] = l2;
  }
  __syncthreads();
  if (threadIdx.x < 64) {
    const double lm = t == 0 ? 0.0 : oc_node_lam(bracket[0], bracket[1], t);
    s.lam[t] = lm;
    s.c[t] = t == 0 ? 0.0 : 1.0 / sqrt(lm);
  }
}

void oc_walk(const int* blk_start, const LevelCell* cells, int ncells, int depth, const double* partial, OCState* st,
             hipStream_t s) {
  k_oc_walk<<<ncells, 1024, 0, s>>>(blk_start, cells, depth, partial, st);
}

template <bool FROM_DJ>
__global__ __launch_bounds__(256) void k_oc_final(Launch L, const OCState* __restrict__ st,
                                                  const double* __restrict__ x, const double* __restrict__ a,
                                                  double* __restrict__ xnew) {
  const int ci = find_cell(L.blk_start, L.ncells, lbid());
  const LevelCell c = L.cells[ci];
  const long long n = (long long)(lbid() - L.blk_start[ci]) * 256 + threadIdx.x;
  if (n >= (long long)c.nx * c.ny * c.nz) return;
  const OCState& s = st[ci];
  const long long e = c.off + n;
  const double xv = x[e];
  xnew[e] = oc_value(oc_a<FROM_DJ>(a, xv, e), xv, 1.0 / sqrt(s.last_lam), s.move, s.xmin);
}

void oc_final(const Launch& L, const OCState* st, const double* x, const double* a, double* xnew, hipStream_t s) {
  k_oc_final<false><<<L.nblocks, 256, 0, s>>>(L, st, x, a, xnew);
}

void oc_final_dj(const Launch& L, const OCState* st, const double* x, const double* dJ, double* xnew, hipStream_t s) {
  k_oc_final<true><<<L.nblocks, 256, 0, s>>>(L, st, x, dJ, xnew);
}
