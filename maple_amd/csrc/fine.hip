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
#include "fine_util.h"

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
// VAR 0: the compiler may keep all 18*ND neighbour values live at once;
// VAR 1: a compiler barrier after each element stages the loads element by element
//        (4 x 8*ND loads from L1 instead of 18*ND), trading re-loads for fewer live
//        registers and more waves per SIMD.
template <int ND, typename T, int MODE, int VAR>
__global__ __launch_bounds__(256) void k_fine_apply(
    const LevelCell* __restrict__ cells, const int* __restrict__ blk_start, int ncells,
    long long Ntot, int ncase, const T* __restrict__ Ke, const T* __restrict__ rho,
    const T* __restrict__ u, const T* __restrict__ f, const T* __restrict__ dinv, T omega,
    T* __restrict__ out, double* __restrict__ partial, const int* __restrict__ done, int ndc,
    unsigned long long* __restrict__ work) {
  count_work(cells, ncells, done, ndc, ncase, work);
  constexpr int NE = 8 * ND;
  const int k = find_cell(blk_start, ncells, lbid());
  const LevelCell c = cells[k];
  if (case_done(done, ndc, c.id, blockIdx.y)) return;   // converged: nothing of this case is read again
  const int wave = (lbid() - blk_start[k]) * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const LaneMap m = lane_map(c, wave, lane);
  Ke += (long long)c.id * NE * NE;   // per-cell element matrix (h = 1/r scaling)

  // Every lane (also an idle one) holds in-range coordinates, so loads need no
  // guard; only the store is masked.
  const int xs[3] = {wrap(m.x - 1, c.nx), m.x, wrap(m.x + 1, c.nx)};
  const int ys[3] = {wrap(m.y - 1, c.ny), m.y, wrap(m.y + 1, c.ny)};
  const int zs[2] = {m.ze, wrap(m.ze + 1, c.nz)};
  int lb[3][3];   // node offset of line (X, Y) inside the cell
#pragma unroll
  for (int X = 0; X < 3; ++X)
#pragma unroll
    for (int Y = 0; Y < 3; ++Y) lb[X][Y] = (xs[X] * c.ny + ys[Y]) * c.nz;
  const T* __restrict__ rc = rho + c.off;
  auto nid = [&](int X, int Y, int Z) -> long long { return c.off + lb[X][Y] + zs[Z]; };

  T r4[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) r4[a][b] = rc[lb[a][b] + zs[0]];

  const long long own = nid(1, 1, 0);
  // One load case per block row: a loop over cases would let the compiler hoist all
  // 8*ND*ND*8 Ke operands out of it and spill them from SGPRs into VGPR lanes.
  const int j = blockIdx.y;
  const T* __restrict__ uj = u + (long long)j * ND * Ntot + c.off;
  T lo[ND], up[ND];
#pragma unroll
  for (int d = 0; d < ND; ++d) lo[d] = up[d] = T(0);
  T uown[ND];
#pragma unroll
  for (int a = 0; a < 2; ++a) {
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      // the 8*ND displacements of element (x-1+a, y-1+b, ze): only this element is
      // held in registers (L1 serves the overlap with the neighbouring elements)
      T ue[8 * ND];
#pragma unroll
      for (int mm = 0; mm < 8; ++mm)
#pragma unroll
        for (int e = 0; e < ND; ++e)
          ue[ND * mm + e] = uj[(long long)e * Ntot + (lb[a + hex_dx(mm)][b + hex_dy(mm)] + zs[hex_dz(mm)])];
      if (a == 1 && b == 1) {
#pragma unroll
        for (int e = 0; e < ND; ++e) uown[e] = ue[e];   // corner 0 of element (x,y,ze) is node (x,y,ze)
      }
      const int kL = hex_corner(1 - a, 1 - b, 0), kU = hex_corner(1 - a, 1 - b, 1);
      // Runtime loop over the 2*ND rows of this element: the Ke row index changes
      // every iteration, so its 8*ND scalar operands are loaded right before use
      // instead of being hoisted (and spilled) together with every other row.
#pragma nounroll
      for (int rr = 0; rr < 2 * ND; ++rr) {
        const int row = rr < ND ? ND * kL + rr : ND * kU + (rr - ND);
        const T* __restrict__ Kr = Ke + row * NE;
        // three independent partial sums keep several FMAs in flight
        T t0 = T(0), t1 = T(0), t2 = T(0);
#pragma unroll
        for (int q = 0; q < NE; q += 3) {
          t0 = fma(Kr[q], ue[q], t0);
          if (q + 1 < NE) t1 = fma(Kr[q + 1], ue[q + 1], t1);
          if (q + 2 < NE) t2 = fma(Kr[q + 2], ue[q + 2], t2);
        }
        const T t = r4[a][b] * ((t0 + t1) + t2);
        // rr is uniform: a scalar branch picks the accumulator
        if (ND == 3) {
          switch (rr) {
            case 0: lo[0] += t; break;
            case 1: lo[ND > 1 ? 1 : 0] += t; break;
            case 2: lo[ND > 2 ? 2 : 0] += t; break;
            case 3: up[0] += t; break;
            case 4: up[ND > 1 ? 1 : 0] += t; break;
            default: up[ND > 2 ? 2 : 0] += t; break;
          }
        } else {
          if (rr == 0) lo[0] += t; else up[0] += t;
        }
      }
      if constexpr (VAR == 1) asm volatile("" ::: "memory");
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
        if constexpr (MODE == 1) dj += (double)uown[d] * (double)kv;
      } else if constexpr (MODE == 2) {
        out[idx] = f[idx] - kv;
      } else {
        out[idx] = uown[d] + omega * dinv[(long long)d * Ntot + own] * (f[idx] - kv);
      }
    }
  }
  if constexpr (MODE == 1) {
    double v[1] = {dj};
    block_sum<1>(v, partial + (long long)lbid() * MAPLE_MAX_CASES + j);
  }
}

template <int ND, typename T, int VAR>
void fine_apply_v(const Launch& L, int mode, int ncase, const T* Ke, const T* rho, const T* u,
                  const T* f, const T* dinv, T omega, T* out, double* partial, hipStream_t s) {
  dim3 grid(L.nblocks, ncase), block(256);
  switch (mode) {
    case 0: k_fine_apply<ND, T, 0, VAR><<<grid, block, 0, s>>>(L.cells, L.blk_start, L.ncells, L.Ntot, ncase, Ke, rho, u, f, dinv, omega, out, partial, L.done, L.ndc, L.work); break;
    case 1: k_fine_apply<ND, T, 1, VAR><<<grid, block, 0, s>>>(L.cells, L.blk_start, L.ncells, L.Ntot, ncase, Ke, rho, u, f, dinv, omega, out, partial, L.done, L.ndc, L.work); break;
    case 2: k_fine_apply<ND, T, 2, VAR><<<grid, block, 0, s>>>(L.cells, L.blk_start, L.ncells, L.Ntot, ncase, Ke, rho, u, f, dinv, omega, out, partial, L.done, L.ndc, L.work); break;
    default: k_fine_apply<ND, T, 3, VAR><<<grid, block, 0, s>>>(L.cells, L.blk_start, L.ncells, L.Ntot, ncase, Ke, rho, u, f, dinv, omega, out, partial, L.done, L.ndc, L.work); break;
  }
}

template <int ND, typename T>
void fine_apply(const Launch& L, int mode, int ncase, const T* Ke, const T* rho, const T* u,
                const T* f, const T* dinv, T omega, T* out, double* partial, hipStream_t s) {
  if (L.var == 1) fine_apply_v<ND, T, 1>(L, mode, ncase, Ke, rho, u, f, dinv, omega, out, partial, s);
  else fine_apply_v<ND, T, 0>(L, mode, ncase, Ke, rho, u, f, dinv, omega, out, partial, s);
}

template void fine_apply<3, double>(const Launch&, int, int, const double*, const double*, const double*, const double*, const double*, double, double*, double*, hipStream_t);
template void fine_apply<1, double>(const Launch&, int, int, const double*, const double*, const double*, const double*, const double*, double, double*, double*, hipStream_t);
template void fine_apply<3, float>(const Launch&, int, int, const float*, const float*, const float*, const float*, const float*, float, float*, double*, hipStream_t);
template void fine_apply<1, float>(const Launch&, int, int, const float*, const float*, const float*, const float*, const float*, float, float*, double*, hipStream_t);


// F[j][d][n] = sum_{e owns n as corner k} rho_e Fe[ND*k+d][j]   (_utils.py:453-507)
template <int ND>
__global__ __launch_bounds__(256) void k_fine_rhs(const LevelCell* __restrict__ cells,
                                                  const int* __restrict__ blk_start, int ncells,
                                                  long long Ntot, int ncase, const double* __restrict__ Fe,
                                                  const double* __restrict__ rho, double* __restrict__ F) {
  LevelCell c; int x, y, z; long long gid;
  if (!node_of(cells, blk_start, ncells, c, x, y, z, gid)) return;
  Fe += (long long)c.id * 8 * ND * ncase;
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
// L1: the diagonal is replaced by the absolute row sum sum_j |K[n,d][j]| ("l1-Jacobi"),
// which bounds the spectrum of Dinv K by 1 for any density field.
template <int ND, typename T, bool L1>
__global__ __launch_bounds__(256) void k_fine_dinv(const LevelCell* __restrict__ cells,
                                                   const int* __restrict__ blk_start, int ncells,
                                                   long long Ntot, const double* __restrict__ Ke,
                                                   const double* __restrict__ rho, T* __restrict__ dinv) {
  LevelCell c; int x, y, z; long long gid;
  if (!node_of(cells, blk_start, ncells, c, x, y, z, gid)) return;
  Ke += (long long)c.id * 64 * ND * ND;
  double acc[ND];
#pragma unroll
  for (int d = 0; d < ND; ++d) acc[d] = 0.0;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const long long e = c.off + ((long long)wrap(x - hex_dx(k), c.nx) * c.ny + wrap(y - hex_dy(k), c.ny)) * c.nz +
                        wrap(z - hex_dz(k), c.nz);
    const double r = rho[e];
#pragma unroll
    for (int d = 0; d < ND; ++d) {
      double kv;
      if constexpr (L1) {
        kv = 0.0;
        for (int j = 0; j < 8 * ND; ++j) kv += fabs(Ke[(ND * k + d) * (8 * ND) + j]);
      } else {
        kv = Ke[(ND * k + d) * (8 * ND) + ND * k + d];
      }
      acc[d] = fma(r, kv, acc[d]);
    }
  }
#pragma unroll
  for (int d = 0; d < ND; ++d) dinv[(long long)d * Ntot + gid] = (T)(acc[d] != 0.0 ? 1.0 / acc[d] : 1.0);
}

void fine_rhs(const Launch& L, int nd, int ncase, const double* Fe, const double* rho, double* F, hipStream_t s) {
  if (nd == 3) k_fine_rhs<3><<<L.nblocks, 256, 0, s>>>(L.cells, L.blk_start, L.ncells, L.Ntot, ncase, Fe, rho, F);
  else k_fine_rhs<1><<<L.nblocks, 256, 0, s>>>(L.cells, L.blk_start, L.ncells, L.Ntot, ncase, Fe, rho, F);
}

template <typename T>
void fine_dinv(const Launch& L, int nd, bool l1, const double* Ke, const double* rho, T* dinv, hipStream_t s) {
  if (nd == 3 && l1) k_fine_dinv<3, T, true><<<L.nblocks, 256, 0, s>>>(L.cells, L.blk_start, L.ncells, L.Ntot, Ke, rho, dinv);
  else if (nd == 3) k_fine_dinv<3, T, false><<<L.nblocks, 256, 0, s>>>(L.cells, L.blk_start, L.ncells, L.Ntot, Ke, rho, dinv);
  else if (l1) k_fine_dinv<1, T, true><<<L.nblocks, 256, 0, s>>>(L.cells, L.blk_start, L.ncells, L.Ntot, Ke, rho, dinv);
  else k_fine_dinv<1, T, false><<<L.nblocks, 256, 0, s>>>(L.cells, L.blk_start, L.ncells, L.Ntot, Ke, rho, dinv);
}
template void fine_dinv<double>(const Launch&, int, bool, const double*, const double*, double*, hipStream_t);
template void fine_dinv<float>(const Launch&, int, bool, const double*, const double*, float*, hipStream_t);

