// Level-0 (voxel) operator kernels: the matrix-free SIMP-scaled element-by-element
// product  v = sum_e A_e^T (rho_e Ke) A_e u  of the reference's EBEOperator.matvec
// (EBE_GMG.py:412-430, scalar twin EBE_GMG_scalar.py:292-306), the RHS assembly
// (_utils.py:453-507) and the assembled diagonal (EBE_GMG.py:375-393).
//
// Work decomposition (node-centric, atomic-free, deterministic): one wavefront
// owns z-pencils of output nodes.  Lane (pencil (x,y), z) computes the four
// elements (x-1+a, y-1+b, z), a,b in {0,1}; for each it forms only the 2*ND rows
// of Ke that belong to the nodes (x,y,z) ("lower") and (x,y,z+1) ("upper") --
// 576 FMA per node for elasticity, the minimum for this operator.  The upper
// contributions move to the lane that owns node z+1 with one cross-lane shuffle.
// Every lane uses the same Ke rows in the same order, so Ke is read through
// the scalar unit and each FMA takes its Ke operand from an SGPR pair.
//
//   nz <= 64: a wave holds floor(64/nz) complete periodic pencils (rotate).
//   nz  > 64: a wave holds a 64-element segment and writes 63 nodes.
#include "kernels.h"

namespace {

struct LaneMap {
  int x, y, ze;        // pencil and element z of this lane
  int src;             // lane holding element ze-1 (whose upper node is ze)
  bool comp, out;      // computes an element / writes node ze
};

__device__ __forceinline__ LaneMap lane_map(const LevelCell& c, int wave, int lane) {
  LaneMap m;
  const int np = c.nx * c.ny;
  if (c.nz <= 64) {
    const int W = c.nz, G = 64 / c.nz;
    const int g = lane / W, zl = lane - g * W;
    const int p = wave * G + g;
    m.comp = (g < G) && (p < np);
    m.out = m.comp;
    m.ze = zl;
    m.src = m.comp ? g * W + (zl == 0 ? W - 1 : zl - 1) : lane;
    m.x = m.comp ? p / c.ny : 0;
    m.y = m.comp ? p - (p / c.ny) * c.ny : 0;
  } else {
    const int nseg = (c.nz + 62) / 63;
    const int p = wave / nseg, seg = wave - p * nseg;
    const int z0 = seg * 63;
    const int S = min(63, c.nz - z0);
    m.comp = (p < np) && (lane <= S);
    m.out = m.comp && lane >= 1;
    m.ze = wrap(z0 - 1 + lane, c.nz);
    m.src = lane >= 1 ? lane - 1 : 0;
    m.x = p < np ? p / c.ny : 0;
    m.y = p < np ? p - (p / c.ny) * c.ny : 0;
  }
  return m;
}

template <typename T>
__device__ __forceinline__ T shfl_from(T v, int src) { return __shfl(v, src, 64); }

}  // namespace

int fine_waves(const LevelCell& c) {
  const int np = c.nx * c.ny;
  if (c.nz <= 64) {
    const int G = 64 / c.nz;
    return (np + G - 1) / G;
  }
  return np * ((c.nz + 62) / 63);
}

// MODE: 0 = v = K u, 1 = v = K u and partial (u, K u), 2 = v = f - K u,
//       3 = v = u + omega Dinv (f - K u)  (weighted Jacobi sweep)
template <int ND, typename T, int MODE>
__global__ __launch_bounds__(256) void k_fine_apply(
    const LevelCell* __restrict__ cells, const int* __restrict__ blk_start, int ncells,
    long long Ntot, int ncase, const T* __restrict__ Ke, const T* __restrict__ rho,
    const T* __restrict__ u, const T* __restrict__ f, const T* __restrict__ dinv, T omega,
    T* __restrict__ out, double* __restrict__ partial) {
  constexpr int NE = 8 * ND;
  const int ci = find_cell(blk_start, ncells, blockIdx.x);
  const LevelCell c = cells[ci];
  const int wave = (blockIdx.x - blk_start[ci]) * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const LaneMap m = lane_map(c, wave, lane);
  Ke += (long long)ci * NE * NE;   // per-cell element matrix (h = 1/r scaling)

  const int xs[3] = {wrap(m.x - 1, c.nx), m.x, wrap(m.x + 1, c.nx)};
  const int ys[3] = {wrap(m.y - 1, c.ny), m.y, wrap(m.y + 1, c.ny)};
  const int zs[2] = {m.ze, wrap(m.ze + 1, c.nz)};
  const long long base = c.off;
  auto nid = [&](int X, int Y, int Z) -> long long {
    return base + ((long long)xs[X] * c.ny + ys[Y]) * c.nz + zs[Z];
  };

  T r4[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) r4[a][b] = m.comp ? rho[nid(a, b, 0)] : T(0);

  const long long own = nid(1, 1, 0);

  for (int j = 0; j < ncase; ++j) {
    T uu[3][3][2][ND];
#pragma unroll
    for (int X = 0; X < 3; ++X)
#pragma unroll
      for (int Y = 0; Y < 3; ++Y)
#pragma unroll
        for (int Z = 0; Z < 2; ++Z)
#pragma unroll
          for (int d = 0; d < ND; ++d)
            uu[X][Y][Z][d] = m.comp ? u[(long long)(j * ND + d) * Ntot + nid(X, Y, Z)] : T(0);

    T lo[ND], up[ND];
#pragma unroll
    for (int d = 0; d < ND; ++d) lo[d] = up[d] = T(0);
#pragma unroll
    for (int a = 0; a < 2; ++a) {
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        const int kL = hex_corner(1 - a, 1 - b, 0), kU = hex_corner(1 - a, 1 - b, 1);
#pragma unroll
        for (int d = 0; d < ND; ++d) {
          T tL = T(0), tU = T(0);
#pragma unroll
          for (int mm = 0; mm < 8; ++mm) {
#pragma unroll
            for (int e = 0; e < ND; ++e) {
              const T um = uu[a + hex_dx(mm)][b + hex_dy(mm)][hex_dz(mm)][e];
              tL = fma(Ke[(ND * kL + d) * NE + ND * mm + e], um, tL);
              tU = fma(Ke[(ND * kU + d) * NE + ND * mm + e], um, tU);
            }
          }
          lo[d] = fma(r4[a][b], tL, lo[d]);
          up[d] = fma(r4[a][b], tU, up[d]);
        }
      }
    }
    double dj = 0.0;
#pragma unroll
    for (int d = 0; d < ND; ++d) {
      const T kv = lo[d] + shfl_from(up[d], m.src);
      if (m.out) {
        const long long idx = (long long)(j * ND + d) * Ntot + own;
        if constexpr (MODE == 0 || MODE == 1) {
          out[idx] = kv;
          if constexpr (MODE == 1) dj += (double)uu[1][1][0][d] * (double)kv;
        } else if constexpr (MODE == 2) {
          out[idx] = f[idx] - kv;
        } else {
          out[idx] = uu[1][1][0][d] + omega * dinv[(long long)d * Ntot + own] * (f[idx] - kv);
        }
      }
    }
    if constexpr (MODE == 1) {
      double v[1] = {dj};
      block_sum<1>(v, partial + (long long)blockIdx.x * MAPLE_MAX_CASES + j);
    }
  }
}

template <int ND, typename T>
void fine_apply(const Launch& L, int mode, int ncase, const T* Ke, const T* rho, const T* u,
                const T* f, const T* dinv, T omega, T* out, double* partial, hipStream_t s) {
  dim3 grid(L.nblocks), block(256);
  switch (mode) {
    case 0: k_fine_apply<ND, T, 0><<<grid, block, 0, s>>>(L.cells, L.blk_start, L.ncells, L.Ntot, ncase, Ke, rho, u, f, dinv, omega, out, partial); break;
    case 1: k_fine_apply<ND, T, 1><<<grid, block, 0, s>>>(L.cells, L.blk_start, L.ncells, L.Ntot, ncase, Ke, rho, u, f, dinv, omega, out, partial); break;
    case 2: k_fine_apply<ND, T, 2><<<grid, block, 0, s>>>(L.cells, L.blk_start, L.ncells, L.Ntot, ncase, Ke, rho, u, f, dinv, omega, out, partial); break;
    default: k_fine_apply<ND, T, 3><<<grid, block, 0, s>>>(L.cells, L.blk_start, L.ncells, L.Ntot, ncase, Ke, rho, u, f, dinv, omega, out, partial); break;
  }
}

template void fine_apply<3, double>(const Launch&, int, int, const double*, const double*, const double*, const double*, const double*, double, double*, double*, hipStream_t);
template void fine_apply<1, double>(const Launch&, int, int, const double*, const double*, const double*, const double*, const double*, double, double*, double*, hipStream_t);
template void fine_apply<3, float>(const Launch&, int, int, const float*, const float*, const float*, const float*, const float*, float, float*, double*, hipStream_t);
template void fine_apply<1, float>(const Launch&, int, int, const float*, const float*, const float*, const float*, const float*, float, float*, double*, hipStream_t);

// ---------------------------------------------------------------------------
// Node-centric helpers over the node grid (one thread per node, cell-aligned
// blocks of 256 nodes).

__device__ __forceinline__ bool node_of(const LevelCell* cells, const int* blk_start, int ncells,
                                        LevelCell& c, int& x, int& y, int& z, long long& gid) {
  const int ci = find_cell(blk_start, ncells, blockIdx.x);
  c = cells[ci];
  const long long n = (long long)(blockIdx.x - blk_start[ci]) * 256 + threadIdx.x;
  const long long nn = (long long)c.nx * c.ny * c.nz;
  if (n >= nn) return false;
  z = (int)(n % c.nz);
  const long long t = n / c.nz;
  y = (int)(t % c.ny);
  x = (int)(t / c.ny);
  gid = c.off + n;
  return true;
}

// F[j][d][n] = sum_{e owns n as corner k} rho_e Fe[ND*k+d][j]   (_utils.py:453-507)
template <int ND>
__global__ __launch_bounds__(256) void k_fine_rhs(const LevelCell* __restrict__ cells,
                                                  const int* __restrict__ blk_start, int ncells,
                                                  long long Ntot, int ncase, const double* __restrict__ Fe,
                                                  const double* __restrict__ rho, double* __restrict__ F) {
  LevelCell c; int x, y, z; long long gid;
  if (!node_of(cells, blk_start, ncells, c, x, y, z, gid)) return;
  Fe += (long long)find_cell(blk_start, ncells, blockIdx.x) * 8 * ND * ncase;
  double acc[MAPLE_MAX_CASES][ND];
#pragma unroll
  for (int j = 0; j < MAPLE_MAX_CASES; ++j)
#pragma unroll
    for (int d = 0; d < ND; ++d) acc[j][d] = 0.0;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const long long e = c.off + ((long long)wrap(x - hex_dx(k), c.nx) * c.ny + wrap(y - hex_dy(k), c.ny)) * c.nz +
                        wrap(z - hex_dz(k), c.nz);
    const double r = rho[e];
#pragma unroll
    for (int d = 0; d < ND; ++d)
#pragma unroll
      for (int j = 0; j < MAPLE_MAX_CASES; ++j)
        if (j < ncase) acc[j][d] = fma(r, Fe[(ND * k + d) * ncase + j], acc[j][d]);
  }
#pragma unroll
  for (int j = 0; j < MAPLE_MAX_CASES; ++j)
    if (j < ncase)
#pragma unroll
      for (int d = 0; d < ND; ++d) F[(long long)(j * ND + d) * Ntot + gid] = acc[j][d];
}

// Dinv[d][n] = 1 / sum_{e owns n as corner k} rho_e Ke[ND*k+d][ND*k+d]  (zero -> 1, EBE_GMG.py:388)
template <int ND, typename T>
__global__ __launch_bounds__(256) void k_fine_dinv(const LevelCell* __restrict__ cells,
                                                   const int* __restrict__ blk_start, int ncells,
                                                   long long Ntot, const double* __restrict__ Ke,
                                                   const double* __restrict__ rho, T* __restrict__ dinv) {
  LevelCell c; int x, y, z; long long gid;
  if (!node_of(cells, blk_start, ncells, c, x, y, z, gid)) return;
  Ke += (long long)find_cell(blk_start, ncells, blockIdx.x) * 64 * ND * ND;
  double acc[ND];
#pragma unroll
  for (int d = 0; d < ND; ++d) acc[d] = 0.0;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const long long e = c.off + ((long long)wrap(x - hex_dx(k), c.nx) * c.ny + wrap(y - hex_dy(k), c.ny)) * c.nz +
                        wrap(z - hex_dz(k), c.nz);
    const double r = rho[e];
#pragma unroll
    for (int d = 0; d < ND; ++d) acc[d] = fma(r, Ke[(ND * k + d) * (8 * ND) + ND * k + d], acc[d]);
  }
#pragma unroll
  for (int d = 0; d < ND; ++d) dinv[(long long)d * Ntot + gid] = (T)(acc[d] != 0.0 ? 1.0 / acc[d] : 1.0);
}

void fine_rhs(const Launch& L, int nd, int ncase, const double* Fe, const double* rho, double* F, hipStream_t s) {
  if (nd == 3) k_fine_rhs<3><<<L.nblocks, 256, 0, s>>>(L.cells, L.blk_start, L.ncells, L.Ntot, ncase, Fe, rho, F);
  else k_fine_rhs<1><<<L.nblocks, 256, 0, s>>>(L.cells, L.blk_start, L.ncells, L.Ntot, ncase, Fe, rho, F);
}

template <typename T>
void fine_dinv(const Launch& L, int nd, const double* Ke, const double* rho, T* dinv, hipStream_t s) {
  if (nd == 3) k_fine_dinv<3, T><<<L.nblocks, 256, 0, s>>>(L.cells, L.blk_start, L.ncells, L.Ntot, Ke, rho, dinv);
  else k_fine_dinv<1, T><<<L.nblocks, 256, 0, s>>>(L.cells, L.blk_start, L.ncells, L.Ntot, Ke, rho, dinv);
}
template void fine_dinv<double>(const Launch&, int, const double*, const double*, double*, hipStream_t);
template void fine_dinv<float>(const Launch&, int, const double*, const double*, float*, hipStream_t);
