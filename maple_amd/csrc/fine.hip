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

// Work of one operator launch, for kernel timing: the first wavefront of block
// (0, 0) adds the elements x load cases the launch processes (listed cells, cases
// not converged when it runs) to *work.  The convergence flags do not change while
// an operator launch runs, so every block of it sees the same ones.
__device__ __forceinline__ void count_work(const LevelCell* __restrict__ cells, int ncells,
                                           const int* __restrict__ done, int ndc, int ncase,
                                           unsigned long long* __restrict__ work) {
  if (work == nullptr || blockIdx.x != 0 || blockIdx.y != 0 || threadIdx.x >= 64) return;
  unsigned long long w = 0;
  for (int k = threadIdx.x; k < ncells; k += 64) {
    const LevelCell c = cells[k];
    const unsigned long long ne = (unsigned long long)c.nx * c.ny * c.nz;
    for (int j = 0; j < ncase; ++j)
      if (!(done != nullptr && done[c.id * ndc + j] != 0)) w += ne;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) w += __shfl_xor(w, o, 64);
  if (threadIdx.x == 0) atomicAdd(work, w);
}

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

// ---------------------------------------------------------------------------
// Node-centric helpers over the node grid (one thread per node, cell-aligned
// blocks of 256 nodes).

__device__ __forceinline__ bool node_of(const LevelCell* cells, const int* blk_start, int ncells,
                                        LevelCell& c, int& x, int& y, int& z, long long& gid) {
  const int ci = find_cell(blk_start, ncells, lbid());
  c = cells[ci];
  const long long n = (long long)(lbid() - blk_start[ci]) * 256 + threadIdx.x;
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

// variants 4, 5: 2x1 nodes, six / three cases per thread; 6, 7: 2x2 nodes, three / two
template <typename T>
void fine_apply_nodeb(const Launch& L, int mode, int ncase, const T* Ke, const T* rho, const T* u, const T* f,
                      const T* dinv, T omega, T* out, double* partial, hipStream_t s) {
  switch (L.var) {
    case 4: fine_apply_nodeb_g<T, 6, 2, 1>(L, mode, ncase, Ke, rho, u, f, dinv, omega, out, partial, s); break;
    case 5: fine_apply_nodeb_g<T, 3, 2, 1>(L, mode, ncase, Ke, rho, u, f, dinv, omega, out, partial, s); break;
    case 6: fine_apply_nodeb_g<T, 3, 2, 2>(L, mode, ncase, Ke, rho, u, f, dinv, omega, out, partial, s); break;
    default: fine_apply_nodeb_g<T, 2, 2, 2>(L, mode, ncase, Ke, rho, u, f, dinv, omega, out, partial, s); break;
  }
}
template void fine_apply_nodeb<double>(const Launch&, int, int, const double*, const double*, const double*, const double*, const double*, double, double*, double*, hipStream_t);
template void fine_apply_nodeb<float>(const Launch&, int, int, const float*, const float*, const float*, const float*, const float*, float, float*, double*, hipStream_t);

// ---------------------------------------------------------------------------
// Variant 8: the 2x1 node-block kernel (three cases per thread) for cells with
// nz = 64, where a wavefront holds exactly one periodic z-pencil (thread t =
// (gx*ny + y)*64 + z).  Each lane gathers the neighbour columns at its own z only
// and takes the values at z-1 and z+1 from the neighbouring lanes by a rotation of
// the whole wavefront (DPP wave_ror:1 / wave_rol:1, periodic like the pencil):
// 4*3 gathers per case component instead of 4*3*3, and the element moduli at z-1
// likewise.  The sums are formed in the same order as variant 5.
namespace {
__device__ __forceinline__ float lane_from_zm1(float v) {   // value of the lane holding z-1
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x13C, 0xF, 0xF, false));
}
__device__ __forceinline__ float lane_from_zp1(float v) {   // value of the lane holding z+1
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x134, 0xF, 0xF, false));
}
__device__ __forceinline__ double lane_from_zm1(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffLL), 0x13C, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x13C, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ double lane_from_zp1(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffLL), 0x134, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x134, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
}  // namespace

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
