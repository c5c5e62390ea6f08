// Level-0 elastic operator for nz = 64 cells: node blocks with z neighbours by lane rotation (variant 8).
#include "fine_util.h"

// ---------------------------------------------------------------------------
// Variant 8: the 2x1 node-block kernel (three cases per thread) for cells with
// nz = 64, where a wavefront holds exactly one periodic z-pencil (thread t =
// (gx*ny + y)*64 + z).  Each lane gathers the neighbour columns at its own z only
// and takes the values at z-1 and z+1 from the neighbouring lanes by a rotation of
// the whole wavefront (DPP wave_ror:1 / wave_rol:1, periodic like the pencil):
// 4*3 gathers per case component instead of 4*3*3, and the element moduli at z-1
// likewise.  The sums are formed in the same order as variant 5.

template <typename T, int MODE, int NCG>
__global__ __launch_bounds__(256) void k_fine_nodeb64(
    const LevelCell* __restrict__ cells, const int* __restrict__ blk_start, int ncells,
    long long Ntot, int ncase, const T* __restrict__ Ke, const T* __restrict__ rho,
    const T* __restrict__ u, const T* __restrict__ f, const T* __restrict__ dinv, T omega,
    T* __restrict__ out, double* __restrict__ partial, const int* __restrict__ done, int ndc,
    unsigned long long* __restrict__ work) {
  count_work(cells, ncells, done, ndc, ncase, work);
  constexpr int NX = 2;
  const int k = find_cell(blk_start, ncells, lbid());
  const LevelCell c = cells[k];
  const int j0 = blockIdx.y * NCG;
  bool live[NCG];
  bool any = false;
#pragma unroll
  for (int q = 0; q < NCG; ++q) {
    live[q] = (j0 + q < ncase) && !case_done(done, ndc, c.id, j0 + q);   // uniform over the block
    any = any || live[q];
  }
  if (!any) return;
  Ke += (long long)c.id * 576;
  // nz = 64: a wavefront is one whole pencil, so a wave is all valid or all idle
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
  // moduli of the elements (x0-1+a, y-1+b, z-1+cc): cc = 1 gathered, cc = 0 from the lane of z-1
  T r[NX + 1][2][2];
#pragma unroll
  for (int a = 0; a < NX + 1; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      r[a][b][1] = rho[c.off + ((long long)xs[a] * c.ny + ys[b]) * 64 + z];
      r[a][b][0] = lane_from_zm1(r[a][b][1]);
    }

  T acc[NX][NCG][3];
#pragma unroll
  for (int kx = 0; kx < NX; ++kx)
#pragma unroll
    for (int q = 0; q < NCG; ++q)
#pragma unroll
      for (int i = 0; i < 3; ++i) acc[kx][q][i] = T(0);

#pragma unroll
  for (int ix = 0; ix < NX + 2; ++ix)
#pragma unroll
    for (int dy = -1; dy <= 1; ++dy) {
      const long long m = c.off + ((long long)xs[ix] * c.ny + ys[dy + 1]) * 64 + z;
      T uc[NCG][3], um[NCG][3], up[NCG][3];
#pragma unroll
      for (int q = 0; q < NCG; ++q) {
        if (!live[q]) continue;
        const T* __restrict__ uq = u + (long long)(j0 + q) * 3 * Ntot + m;
        uc[q][0] = uq[0]; uc[q][1] = uq[Ntot]; uc[q][2] = uq[2 * Ntot];
#pragma unroll
        for (int e = 0; e < 3; ++e) {
          um[q][e] = lane_from_zm1(uc[q][e]);
          up[q][e] = lane_from_zp1(uc[q][e]);
        }
      }
#pragma unroll
      for (int dz = -1; dz <= 1; ++dz)
#pragma unroll
        for (int kx = 0; kx < NX; ++kx) {
          const int dx = ix - 1 - kx;
          if (dx < -1 || dx > 1) continue;
          T Kb[3][3];
#pragma unroll
          for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int kk = 0; kk < 3; ++kk) Kb[i][kk] = T(0);
#pragma unroll
          for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
              for (int cc = 0; cc < 2; ++cc) {
                const int mx = 1 - a + dx, my = 1 - b + dy, mz = 1 - cc + dz;
                if (mx < 0 || mx > 1 || my < 0 || my > 1 || mz < 0 || mz > 1) continue;
                const int cn = hex_corner(1 - a, 1 - b, 1 - cc), cm = hex_corner(mx, my, mz);
                const T re = r[kx + a][b][cc];
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                  for (int kk = 0; kk < 3; ++kk) Kb[i][kk] = fma(re, Ke[(3 * cn + i) * 24 + 3 * cm + kk], Kb[i][kk]);
              }
#pragma unroll
          for (int q = 0; q < NCG; ++q) {
            if (!live[q]) continue;
            const T* uv = dz < 0 ? um[q] : (dz == 0 ? uc[q] : up[q]);
#pragma unroll
            for (int i = 0; i < 3; ++i)
              acc[kx][q][i] = fma(Kb[i][0], uv[0], fma(Kb[i][1], uv[1], fma(Kb[i][2], uv[2], acc[kx][q][i])));
          }
        }
    }

  double dj[NCG];
#pragma unroll
  for (int q = 0; q < NCG; ++q) dj[q] = 0.0;
  if (valid) {
#pragma unroll
    for (int kx = 0; kx < NX; ++kx) {
      const long long gid = c.off + ((long long)(x0 + kx) * c.ny + y) * 64 + z;
#pragma unroll
      for (int q = 0; q < NCG; ++q) {
        if (!live[q]) continue;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          const long long idx = (long long)((j0 + q) * 3 + i) * Ntot + gid;
          const T kv = acc[kx][q][i];
          if constexpr (MODE == 0 || MODE == 1) {
            out[idx] = kv;
            if constexpr (MODE == 1) dj[q] += (double)u[idx] * (double)kv;
          } else if constexpr (MODE == 2) {
            out[idx] = f[idx] - kv;
          } else {
            out[idx] = u[idx] + omega * dinv[(long long)i * Ntot + gid] * (f[idx] - kv);
          }
        }
      }
    }
  }
  if constexpr (MODE == 1) block_sum<NCG>(dj, partial + (long long)lbid() * MAPLE_MAX_CASES + j0);
}

template <typename T>
void fine_apply_nodeb64(const Launch& L, int mode, int ncase, const T* Ke, const T* rho, const T* u, const T* f,
                        const T* dinv, T omega, T* out, double* partial, hipStream_t s) {
  constexpr int NCG = 3;
  dim3 grid(L.nblocks, (ncase + NCG - 1) / NCG), block(256);
  switch (mode) {
    case 0: k_fine_nodeb64<T, 0, NCG><<<grid, block, 0, s>>>(L.cells, L.blk_start, L.ncells, L.Ntot, ncase, Ke, rho, u, f, dinv, omega, out, partial, L.done, L.ndc, L.work); break;
    case 1: k_fine_nodeb64<T, 1, NCG><<<grid, block, 0, s>>>(L.cells, L.blk_start, L.ncells, L.Ntot, ncase, Ke, rho, u, f, dinv, omega, out, partial, L.done, L.ndc, L.work); break;
    case 2: k_fine_nodeb64<T, 2, NCG><<<grid, block, 0, s>>>(L.cells, L.blk_start, L.ncells, L.Ntot, ncase, Ke, rho, u, f, dinv, omega, out, partial, L.done, L.ndc, L.work); break;
    default: k_fine_nodeb64<T, 3, NCG><<<grid, block, 0, s>>>(L.cells, L.blk_start, L.ncells, L.Ntot, ncase, Ke, rho, u, f, dinv, omega, out, partial, L.done, L.ndc, L.work); break;
  }
}
template void fine_apply_nodeb64<double>(const Launch&, int, int, const double*, const double*, const double*, const double*, const double*, double, double*, double*, hipStream_t);
template void fine_apply_nodeb64<float>(const Launch&, int, int, const float*, const float*, const float*, const float*, const float*, float, float*, double*, hipStream_t);

