// Level-0 thermal operator in node form for cells with nz = 64 (the thermal twin of
// EBEOperatorScalar.matvec, EBE_GMG_scalar.py:292-306).
//
// The conductivity element matrix of a cube is the Q1 Laplacian, whose entry for
// corners a, b depends only on their Hamming distance d(a, b) in {0, 1, 2, 3}
// (h/12 [4, 0, -1, -1] for unit conductivity).  Then
//
//   (K u)_n = sum_{m in the 27-neighbourhood of n} k_{|m - n|_1} S_nm u_m,
//   S_nm    = sum of rho_e over the elements holding both n and m,
//
// so a node needs its 8 element moduli, the 27 neighbour values and four
// coefficients per cell: no 64-entry Ke is held anywhere.  All three load cases
// run in one thread (the moduli and the S sums are shared by them), two
// x-adjacent nodes per thread share their neighbour columns, and a wavefront
// holds one periodic z-pencil, so the values at z-1 and z+1 come from the
// neighbouring lanes (DPP rotation) and each thread loads only its own z.  Per
// node and sweep the unique HBM bytes are u, f and the output per case plus rho
// and the smoother diagonal once: the kernel is bound by HBM.
//
// The cells' element matrices must have this structure (checked when they are
// uploaded); otherwise the element-row kernel runs.
#include "fine_util.h"

namespace {

// S of the offset (dx, dy, dz) from the moduli r[a][b][c] of the elements
// (x-1+a, y-1+b, z-1+c): element a holds x and x+dx iff dx = 0 or a = (dx + 1) / 2
template <typename T>
__device__ __forceinline__ T s_sum(const T (&r)[2][2][2], int dx, int dy, int dz) {
  T s = T(0);
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        if (dx != 0 && a != (dx + 1) / 2) continue;
        if (dy != 0 && b != (dy + 1) / 2) continue;
        if (dz != 0 && c != (dz + 1) / 2) continue;
        s += r[a][b][c];
      }
  return s;
}

}  // namespace

// MODE: 0 = K u, 1 = K u and partial (u, K u), 2 = f - K u, 3 = u + omega Dinv (f - K u).
// F0: the coefficient of Hamming distance 1 is zero (the Q1 Laplacian's: h/12 [4, 0, -1, -1]),
// so the six face neighbours carry no coupling and are skipped (21 of the 27 per node).
// Thread t of a cell owns the NX nodes (NX gx + k, y, z), t = (gx*ny + y)*64 + z: NX = 2 on
// the pair launch of the level-0 grid (neighbour columns shared by two nodes), NX = 1 on
// the node launch (half the live registers, twice the waves in flight).
template <typename T, int MODE, int NX, bool F0>
__global__ __launch_bounds__(256) void k_thermal64(
    const LevelCell* __restrict__ cells, const int* __restrict__ blk_start, int ncells, long long Ntot,
    int ncase, const double* __restrict__ coef, const T* __restrict__ rho, const T* __restrict__ u,
    const T* __restrict__ f, const T* __restrict__ dinv, T omega, T* __restrict__ out, double* __restrict__ partial,
    const int* __restrict__ done, int ndc, unsigned long long* __restrict__ work) {
  constexpr int NC = 3;
  count_work(cells, ncells, done, ndc, ncase, work);
  const int k = find_cell(blk_start, ncells, lbid());
  const LevelCell c = cells[k];
  bool live[NC];
  bool any = false;
#pragma unroll
  for (int j = 0; j < NC; ++j) {
    live[j] = j < ncase && !case_done(done, ndc, c.id, j);   // uniform over the block
    any = any || live[j];
  }
  if (!any) return;
  const T kc[4] = {(T)coef[4 * c.id], (T)coef[4 * c.id + 1], (T)coef[4 * c.id + 2], (T)coef[4 * c.id + 3]};
  // a wavefront is one whole pencil: all its lanes are valid or none is
  const long long t = (long long)(lbid() - blk_start[k]) * 256 + threadIdx.x;
  const bool valid = t < (long long)(c.nx / NX) * c.ny * 64;
  const int z = (int)(threadIdx.x & 63);
  int x0 = 0, y = 0;
  if (valid) {
    const long long tt = t >> 6;
    y = (int)(tt % c.ny);
    x0 = (int)(tt / c.ny) * NX;
  }
  int xs[NX + 2];
#pragma unroll
  for (int i = 0; i < NX + 2; ++i) xs[i] = wrap(x0 - 1 + i, c.nx);
  const int ys[3] = {wrap(y - 1, c.ny), y, wrap(y + 1, c.ny)};
  // moduli of the elements (x0-1+a, y-1+b, z-1+cc): cc = 1 loaded, cc = 0 from the lane of z-1
  T re[NX + 1][2][2];
#pragma unroll
  for (int a = 0; a < NX + 1; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      re[a][b][1] = rho[c.off + ((long long)xs[a] * c.ny + ys[b]) * 64 + z];
      re[a][b][0] = lane_from_zm1(re[a][b][1]);
    }
  // coupling weights w[kx][dx][dy][dz] = k_{|d|_1} S_nm of each node's 27 neighbours
  T w[NX][3][3][3];
#pragma unroll
  for (int kx = 0; kx < NX; ++kx) {
    T r[2][2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int cc = 0; cc < 2; ++cc) r[a][b][cc] = re[kx + a][b][cc];
#pragma unroll
    for (int dx = -1; dx <= 1; ++dx)
#pragma unroll
      for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
        for (int dz = -1; dz <= 1; ++dz) {
          const int d1 = (dx != 0) + (dy != 0) + (dz != 0);
          w[kx][dx + 1][dy + 1][dz + 1] = kc[d1] * s_sum(r, dx, dy, dz);
        }
  }
  T acc[NX][NC], uown[NX][NC];
#pragma unroll
  for (int kx = 0; kx < NX; ++kx)
#pragma unroll
    for (int j = 0; j < NC; ++j) acc[kx][j] = uown[kx][j] = T(0);
#pragma unroll
  for (int j = 0; j < NC; ++j) {
    if (!live[j]) continue;
    const T* __restrict__ uj = u + (long long)j * Ntot + c.off;
#pragma unroll
    for (int ix = 0; ix < NX + 2; ++ix)
#pragma unroll
      for (int dy = -1; dy <= 1; ++dy) {
        const T uc = uj[((long long)xs[ix] * c.ny + ys[dy + 1]) * 64 + z];
        const T um = lane_from_zm1(uc), up = lane_from_zp1(uc);
#pragma unroll
        for (int kx = 0; kx < NX; ++kx) {
          const int dx = ix - 1 - kx;
          if (dx < -1 || dx > 1) continue;
          if (dx == 0 && dy == 0) uown[kx][j] = uc;
          const bool zface = F0 && dx == 0 && dy == 0;               // (0, 0, -1), (0, 0, +1)
          const bool xyface = F0 && (dx != 0) + (dy != 0) == 1;      // (+-1, 0, 0), (0, +-1, 0)
          if (!zface) acc[kx][j] = fma(w[kx][dx + 1][dy + 1][0], um, acc[kx][j]);
          if (!xyface) acc[kx][j] = fma(w[kx][dx + 1][dy + 1][1], uc, acc[kx][j]);
          if (!zface) acc[kx][j] = fma(w[kx][dx + 1][dy + 1][2], up, acc[kx][j]);
        }
      }
  }
  double dj[NC];
#pragma unroll
  for (int j = 0; j < NC; ++j) dj[j] = 0.0;
  if (valid) {
#pragma unroll
    for (int kx = 0; kx < NX; ++kx) {
      const long long gid = c.off + ((long long)(x0 + kx) * c.ny + y) * 64 + z;
      T dv = T(0);
      if constexpr (MODE == 3) dv = dinv[gid];
#pragma unroll
      for (int j = 0; j < NC; ++j) {
        if (!live[j]) continue;
        const long long idx = (long long)j * Ntot + gid;
        const T kv = acc[kx][j];
        if constexpr (MODE == 0 || MODE == 1) {
          out[idx] = kv;
          if constexpr (MODE == 1) dj[j] += (double)uown[kx][j] * (double)kv;
        } else if constexpr (MODE == 2) {
          out[idx] = f[idx] - kv;
        } else {
          out[idx] = uown[kx][j] + omega * dv * (f[idx] - kv);
        }
      }
    }
  }
  if constexpr (MODE == 1) block_sum<NC>(dj, partial + (long long)lbid() * MAPLE_MAX_CASES);
}

template <typename T, int NX, bool F0>
void thermal_apply64_g(const Launch& L, int mode, int ncase, const double* coef, const T* rho, const T* u,
                       const T* f, const T* dinv, T omega, T* out, double* partial, hipStream_t s) {
  switch (mode) {
    case 0: k_thermal64<T, 0, NX, F0><<<L.nblocks, 256, 0, s>>>(L.cells, L.blk_start, L.ncells, L.Ntot, ncase, coef, rho, u, f, dinv, omega, out, partial, L.done, L.ndc, L.work); break;
    case 1: k_thermal64<T, 1, NX, F0><<<L.nblocks, 256, 0, s>>>(L.cells, L.blk_start, L.ncells, L.Ntot, ncase, coef, rho, u, f, dinv, omega, out, partial, L.done, L.ndc, L.work); break;
    case 2: k_thermal64<T, 2, NX, F0><<<L.nblocks, 256, 0, s>>>(L.cells, L.blk_start, L.ncells, L.Ntot, ncase, coef, rho, u, f, dinv, omega, out, partial, L.done, L.ndc, L.work); break;
    default: k_thermal64<T, 3, NX, F0><<<L.nblocks, 256, 0, s>>>(L.cells, L.blk_start, L.ncells, L.Ntot, ncase, coef, rho, u, f, dinv, omega, out, partial, L.done, L.ndc, L.work); break;
  }
}

// L.var 2: two nodes per thread (pair launch); 3: one node per thread (node launch)
template <typename T>
void thermal_apply64(const Launch& L, int mode, int ncase, const double* coef, const T* rho, const T* u, const T* f,
                     const T* dinv, T omega, T* out, double* partial, hipStream_t s, bool face0) {
  if (L.var == 3) {
    if (face0) thermal_apply64_g<T, 1, true>(L, mode, ncase, coef, rho, u, f, dinv, omega, out, partial, s);
    else thermal_apply64_g<T, 1, false>(L, mode, ncase, coef, rho, u, f, dinv, omega, out, partial, s);
  } else {
    if (face0) thermal_apply64_g<T, 2, true>(L, mode, ncase, coef, rho, u, f, dinv, omega, out, partial, s);
    else thermal_apply64_g<T, 2, false>(L, mode, ncase, coef, rho, u, f, dinv, omega, out, partial, s);
  }
}
template void thermal_apply64<double>(const Launch&, int, int, const double*, const double*, const double*, const double*, const double*, double, double*, double*, hipStream_t, bool);
template void thermal_apply64<float>(const Launch&, int, int, const double*, const float*, const float*, const float*, const float*, float, float*, double*, hipStream_t, bool);
