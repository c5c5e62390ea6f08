// Level-0 elastic operator, node-block form with one node per thread (variants 2, 3).
#include "fine_util.h"

// ---------------------------------------------------------------------------
// Node-block form of the level-0 elastic operator, every load case per lane.
//
//   (K u)_n = sum_{m in the 27-neighbourhood of n} K_nm u_m,
//   K_nm    = sum_{elements e holding n and m} rho_e Ke[corner_e(n), corner_e(m)]  (3x3)
//
// The same sum as the element-row kernel, regrouped: each lane forms its node's
// 27 blocks K_nm from the 8 element moduli around it (every entry of Ke is used
// once: 576 FMA), then applies them to the load cases (243 FMA per case).  The
// blocks are shared by the cases, so six cases cost 576 + 6*243 = 2034 FMA per
// node instead of 6*576 = 3456.  One thread per node over the node launch
// (cell-aligned blocks of 256 consecutive node ids: z-pencils, coalesced).
// NCG: load cases per thread; blockIdx.y selects the group of NCG cases.
template <typename T, int MODE, int NCG>
__global__ __launch_bounds__(256) void k_fine_node(
    const LevelCell* __restrict__ cells, const int* __restrict__ blk_start, int ncells,
    long long Ntot, int ncase, const T* __restrict__ Ke, const T* __restrict__ rho,
    const T* __restrict__ u, const T* __restrict__ f, const T* __restrict__ dinv, T omega,
    T* __restrict__ out, double* __restrict__ partial, const int* __restrict__ done, int ndc,
    unsigned long long* __restrict__ work) {
  count_work(cells, ncells, done, ndc, ncase, work);
  LevelCell c; int x, y, z; long long gid;
  const bool valid = node_of(cells, blk_start, ncells, c, x, y, z, gid);
  const int j0 = blockIdx.y * NCG;
  bool live[NCG];
  bool any = false;
#pragma unroll
  for (int q = 0; q < NCG; ++q) {
    live[q] = (j0 + q < ncase) && !case_done(done, ndc, c.id, j0 + q);   // uniform over the block
    any = any || live[q];
  }
  if (!any) return;   // every case of this group has converged: nothing is read again
  Ke += (long long)c.id * 576;
  if (!valid) { x = y = z = 0; gid = c.off; }   // in-range coordinates; the store is masked
  const int xs[3] = {wrap(x - 1, c.nx), x, wrap(x + 1, c.nx)};
  const int ys[3] = {wrap(y - 1, c.ny), y, wrap(y + 1, c.ny)};
  const int zs[3] = {wrap(z - 1, c.nz), z, wrap(z + 1, c.nz)};
  // moduli of the 8 elements (x-1+a, y-1+b, z-1+cc) around the node
  T r[2][2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int cc = 0; cc < 2; ++cc) r[a][b][cc] = rho[c.off + ((long long)xs[a] * c.ny + ys[b]) * c.nz + zs[cc]];

  T acc[NCG][3], uown[NCG][3];
#pragma unroll
  for (int q = 0; q < NCG; ++q)
#pragma unroll
    for (int i = 0; i < 3; ++i) acc[q][i] = uown[q][i] = T(0);

#pragma unroll
  for (int dx = -1; dx <= 1; ++dx)
#pragma unroll
    for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
      for (int dz = -1; dz <= 1; ++dz) {
        // K_nm: elements (a,b,cc) around n that also hold m = n + (dx,dy,dz)
        T Kb[3][3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int k = 0; k < 3; ++k) Kb[i][k] = T(0);
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int cc = 0; cc < 2; ++cc) {
              const int mx = 1 - a + dx, my = 1 - b + dy, mz = 1 - cc + dz;
              if (mx < 0 || mx > 1 || my < 0 || my > 1 || mz < 0 || mz > 1) continue;
              const int cn = hex_corner(1 - a, 1 - b, 1 - cc), cm = hex_corner(mx, my, mz);
              const T re = r[a][b][cc];
#pragma unroll
              for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int k = 0; k < 3; ++k) Kb[i][k] = fma(re, Ke[(3 * cn + i) * 24 + 3 * cm + k], Kb[i][k]);
            }
        const long long m = c.off + ((long long)xs[dx + 1] * c.ny + ys[dy + 1]) * c.nz + zs[dz + 1];
#pragma unroll
        for (int q = 0; q < NCG; ++q) {
          if (!live[q]) continue;
          const T* __restrict__ uq = u + (long long)(j0 + q) * 3 * Ntot + m;
          const T u0 = uq[0], u1 = uq[Ntot], u2 = uq[2 * Ntot];
          if (dx == 0 && dy == 0 && dz == 0) {
            uown[q][0] = u0; uown[q][1] = u1; uown[q][2] = u2;
          }
#pragma unroll
          for (int i = 0; i < 3; ++i) acc[q][i] = fma(Kb[i][0], u0, fma(Kb[i][1], u1, fma(Kb[i][2], u2, acc[q][i])));
        }
      }

  double dj[NCG];
#pragma unroll
  for (int q = 0; q < NCG; ++q) {
    dj[q] = 0.0;
    if (!live[q] || !valid) continue;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const long long idx = (long long)((j0 + q) * 3 + i) * Ntot + gid;
      const T kv = acc[q][i];
      if constexpr (MODE == 0 || MODE == 1) {
        out[idx] = kv;
        if constexpr (MODE == 1) dj[q] += (double)uown[q][i] * (double)kv;
      } else if constexpr (MODE == 2) {
        out[idx] = f[idx] - kv;
      } else {
        out[idx] = uown[q][i] + omega * dinv[(long long)i * Ntot + gid] * (f[idx] - kv);
      }
    }
  }
  if constexpr (MODE == 1) block_sum<NCG>(dj, partial + (long long)lbid() * MAPLE_MAX_CASES + j0);
}

template <typename T, int NCG>
void fine_apply_node_g(const Launch& L, int mode, int ncase, const T* Ke, const T* rho, const T* u, const T* f,
                       const T* dinv, T omega, T* out, double* partial, hipStream_t s) {
  dim3 grid(L.nblocks, (ncase + NCG - 1) / NCG), block(256);
  switch (mode) {
    case 0: k_fine_node<T, 0, NCG><<<grid, block, 0, s>>>(L.cells, L.blk_start, L.ncells, L.Ntot, ncase, Ke, rho, u, f, dinv, omega, out, partial, L.done, L.ndc, L.work); break;
    case 1: k_fine_node<T, 1, NCG><<<grid, block, 0, s>>>(L.cells, L.blk_start, L.ncells, L.Ntot, ncase, Ke, rho, u, f, dinv, omega, out, partial, L.done, L.ndc, L.work); break;
    case 2: k_fine_node<T, 2, NCG><<<grid, block, 0, s>>>(L.cells, L.blk_start, L.ncells, L.Ntot, ncase, Ke, rho, u, f, dinv, omega, out, partial, L.done, L.ndc, L.work); break;
    default: k_fine_node<T, 3, NCG><<<grid, block, 0, s>>>(L.cells, L.blk_start, L.ncells, L.Ntot, ncase, Ke, rho, u, f, dinv, omega, out, partial, L.done, L.ndc, L.work); break;
  }
}

// variant 2: six cases per thread; variant 3: three cases per thread (twice the
// blocks, the block assembly done twice)
template <typename T>
void fine_apply_node(const Launch& L, int mode, int ncase, const T* Ke, const T* rho, const T* u, const T* f,
                     const T* dinv, T omega, T* out, double* partial, hipStream_t s) {
  if (L.var == 3) fine_apply_node_g<T, 3>(L, mode, ncase, Ke, rho, u, f, dinv, omega, out, partial, s);
  else fine_apply_node_g<T, 6>(L, mode, ncase, Ke, rho, u, f, dinv, omega, out, partial, s);
}
template void fine_apply_node<double>(const Launch&, int, int, const double*, const double*, const double*, const double*, const double*, double, double*, double*, hipStream_t);
template void fine_apply_node<float>(const Launch&, int, int, const float*, const float*, const float*, const float*, const float*, float, float*, double*, hipStream_t);
