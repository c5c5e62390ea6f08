// Node blocks with NX x NZ nodes per thread: the kernel template of the level-0
// elastic operator variants 4-7 and its launcher.  Each variant is instantiated
// in a translation unit of its own (fine_nodeb_v*.hip) so the build compiles
// them in parallel; fine_nodeb.hip dispatches on the variant.
#pragma once
#include "fine_util.h"


// Node blocks, NX x NZ nodes per thread (variants 4-7).  Neighbouring nodes share
// most of their 27-neighbourhoods, so each neighbour value is gathered once for
// all the thread's nodes: (NX+2)*3*(NZ+2) gathers of every case's components for
// NX*NZ nodes, against 27 per node one node at a time (2x1: 18 per node, 2x2: 12).
// The launch has blocks of 256 threads; thread t of a cell owns the nodes
// (NX*gx + kx, y, NZ*gz + kz), t = (gx*ny + y)*(nz/NZ) + gz (nx, nz even).  The
// block assembly stays per node: 576 FMA per node and case group.
template <typename T, int MODE, int NCG, int NX, int NZ>
__global__ __launch_bounds__(256) void k_fine_nodeb(
    const LevelCell* __restrict__ cells, const int* __restrict__ blk_start, int ncells,
    long long Ntot, int ncase, const T* __restrict__ Ke, const T* __restrict__ rho,
    const T* __restrict__ u, const T* __restrict__ f, const T* __restrict__ dinv, T omega,
    T* __restrict__ out, double* __restrict__ partial, const int* __restrict__ done, int ndc,
    unsigned long long* __restrict__ work) {
  count_work(cells, ncells, done, ndc, ncase, work);
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
  const int nzt = c.nz / NZ;
  const long long t = (long long)(lbid() - blk_start[k]) * 256 + threadIdx.x;
  const bool valid = t < (long long)(c.nx / NX) * c.ny * nzt;
  int x0 = 0, y = 0, z0 = 0;
  if (valid) {
    z0 = (int)(t % nzt) * NZ;
    const long long tt = t / nzt;
    y = (int)(tt % c.ny);
    x0 = (int)(tt / c.ny) * NX;
  }
  int xs[NX + 2], zs[NZ + 2];
#pragma unroll
  for (int i = 0; i < NX + 2; ++i) xs[i] = wrap(x0 - 1 + i, c.nx);
#pragma unroll
  for (int i = 0; i < NZ + 2; ++i) zs[i] = wrap(z0 - 1 + i, c.nz);
  const int ys[3] = {wrap(y - 1, c.ny), y, wrap(y + 1, c.ny)};
  // moduli of the elements (x0-1+a, y-1+b, z0-1+cc)
  T r[NX + 1][2][NZ + 1];
#pragma unroll
  for (int a = 0; a < NX + 1; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int cc = 0; cc < NZ + 1; ++cc) r[a][b][cc] = rho[c.off + ((long long)xs[a] * c.ny + ys[b]) * c.nz + zs[cc]];

  T acc[NX][NZ][NCG][3];
#pragma unroll
  for (int kx = 0; kx < NX; ++kx)
#pragma unroll
    for (int kz = 0; kz < NZ; ++kz)
#pragma unroll
      for (int q = 0; q < NCG; ++q)
#pragma unroll
        for (int i = 0; i < 3; ++i) acc[kx][kz][q][i] = T(0);

#pragma unroll
  for (int ix = 0; ix < NX + 2; ++ix)
#pragma unroll
    for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
      for (int iz = 0; iz < NZ + 2; ++iz) {
        const long long m = c.off + ((long long)xs[ix] * c.ny + ys[dy + 1]) * c.nz + zs[iz];
        T uv[NCG][3];
#pragma unroll
        for (int q = 0; q < NCG; ++q) {
          if (!live[q]) continue;
          const T* __restrict__ uq = u + (long long)(j0 + q) * 3 * Ntot + m;
          uv[q][0] = uq[0]; uv[q][1] = uq[Ntot]; uv[q][2] = uq[2 * Ntot];
        }
#pragma unroll
        for (int kx = 0; kx < NX; ++kx)
#pragma unroll
          for (int kz = 0; kz < NZ; ++kz) {
            const int dx = ix - 1 - kx, dz = iz - 1 - kz;   // neighbour offset from own node (kx, kz)
            if (dx < -1 || dx > 1 || dz < -1 || dz > 1) continue;
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
                  const T re = r[kx + a][b][kz + cc];
#pragma unroll
                  for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int kk = 0; kk < 3; ++kk) Kb[i][kk] = fma(re, Ke[(3 * cn + i) * 24 + 3 * cm + kk], Kb[i][kk]);
                }
#pragma unroll
            for (int q = 0; q < NCG; ++q) {
              if (!live[q]) continue;
#pragma unroll
              for (int i = 0; i < 3; ++i)
                acc[kx][kz][q][i] =
                    fma(Kb[i][0], uv[q][0], fma(Kb[i][1], uv[q][1], fma(Kb[i][2], uv[q][2], acc[kx][kz][q][i])));
            }
          }
      }

  double dj[NCG];
#pragma unroll
  for (int q = 0; q < NCG; ++q) dj[q] = 0.0;
  if (valid) {
#pragma unroll
    for (int kx = 0; kx < NX; ++kx)
#pragma unroll
      for (int kz = 0; kz < NZ; ++kz) {
        const long long gid = c.off + ((long long)(x0 + kx) * c.ny + y) * c.nz + z0 + kz;
#pragma unroll
        for (int q = 0; q < NCG; ++q) {
          if (!live[q]) continue;
#pragma unroll
          for (int i = 0; i < 3; ++i) {
            const long long idx = (long long)((j0 + q) * 3 + i) * Ntot + gid;
            const T kv = acc[kx][kz][q][i];
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

template <typename T, int NCG, int NX, int NZ>
void fine_apply_nodeb_g(const Launch& L, int mode, int ncase, const T* Ke, const T* rho, const T* u, const T* f,
                        const T* dinv, T omega, T* out, double* partial, hipStream_t s) {
  dim3 grid(L.nblocks, (ncase + NCG - 1) / NCG), block(256);
  switch (mode) {
    case 0: k_fine_nodeb<T, 0, NCG, NX, NZ><<<grid, block, 0, s>>>(L.cells, L.blk_start, L.ncells, L.Ntot, ncase, Ke, rho, u, f, dinv, omega, out, partial, L.done, L.ndc, L.work); break;
    case 1: k_fine_nodeb<T, 1, NCG, NX, NZ><<<grid, block, 0, s>>>(L.cells, L.blk_start, L.ncells, L.Ntot, ncase, Ke, rho, u, f, dinv, omega, out, partial, L.done, L.ndc, L.work); break;
    case 2: k_fine_nodeb<T, 2, NCG, NX, NZ><<<grid, block, 0, s>>>(L.cells, L.blk_start, L.ncells, L.Ntot, ncase, Ke, rho, u, f, dinv, omega, out, partial, L.done, L.ndc, L.work); break;
    default: k_fine_nodeb<T, 3, NCG, NX, NZ><<<grid, block, 0, s>>>(L.cells, L.blk_start, L.ncells, L.Ntot, ncase, Ke, rho, u, f, dinv, omega, out, partial, L.done, L.ndc, L.work); break;
  }
}

