// Level-0 elastic operator in FP32 on the matrix cores (variant 9): the element
// product of EBEOperator.matvec (EBE_GMG.py:412-430), v_e = Ke (rho_e u_e) with one
// shared 24 x 24 Ke per cell, as v_mfma_f32_16x16x4f32 tiles.
#include <type_traits>

#include "fine_util.h"

// ---------------------------------------------------------------------------
// Work split.  A workgroup of four wavefronts owns four consecutive node pencils
// (x, y0..y0+3, all z) of one cell over a run of MF_X consecutive x planes, for one
// group of three load cases, and marches along x.  At step s it takes the element
// plane ex = x0 - 1 + s.  Node (X, y, z) is corner (a, b, c) of element
// (X - a, y - b, z - c), so for the pencil y of a wavefront the element pencils
// (ex, y) and (ex, y - 1) contribute through the 12 rows of Ke whose corner has
// b = 0 and b = 1 respectively (a, c in {0, 1}, three components each).
//
// One MFMA tile is 16 rows (M) x 16 elements (N) x 4 Ke columns (K):
//   M: row 4 g + r holds component r (< 3; r = 3 is a zero row) of corner
//      (a, b, c) with g = 2 ((a + parity) mod 2) + c, b fixed per tile;
//   N: 16 consecutive z of the element pencil (nz / 16 tiles per pencil);
//   K: Ke column 4 ks + k (ks = 0..5), the node corner / component of u_e.
// A is the cell's Ke (rows permuted as above: two parities x two b x six ks, 24
// registers per lane); B is rho_e * u, read from LDS.  Both b tiles accumulate into
// the same 16 x 16 result, so a result row holds the sum over the two element pencils.
//
// Rows with a = 0 complete node plane ex (the rows with a = 1 of element plane
// ex - 1 are already in them); rows with a = 1 start node plane ex + 1.  The parity
// in the row map swaps the two row groups every step, so the partial of plane
// ex + 1 stays in its registers and the next step's a = 0 rows land on it.  After
// a step the completed plane is read out of its row groups (c = 0 rows from element
// z, c = 1 rows from element z - 1, moved across lanes with ds_bpermute), the mode
// is applied, the node pencil is written with one coalesced store per component,
// and the row groups are cleared.
//
// LDS: a ring of three node planes (the six pencils y0-1..y0+4 the workgroup reads,
// three load cases x three components) and of three element-modulus planes (pencils
// y0-1..y0+3).  While step s reads the planes of ex and ex + 1, the loads of plane
// ex + 2 are in flight; they are stored after the step's products and a barrier
// closes the step.
//
// Arithmetic per element and load case: 2 tiles of 16 rows for 24 Ke rows (75 % of
// the MFMA work is useful), plus one extra plane per MF_X planes.  Deterministic:
// every sum is formed in a fixed order.  FP32 only: the FP64 operator keeps the
// vector-FMA node-block kernels (FP64 MFMA has no rate over FP64 FMA on this chip).

#define MF_X 16
#define MF_NCG 3
#define MF_ROWS (3 * MF_NCG)   // staged rows per node pencil: case x component

typedef float mf_f4 __attribute__((ext_vector_type(4)));

// Row stride of the staged node planes: 16 mod 64 floats, so the four lane groups of a
// B read (four rows, sixteen consecutive z each) fall on different LDS banks.
template <int NT>
struct MfStride {
  static constexpr int S = 16 * NT;
};

template <int NT>
struct MfShared {
  float u[3][6][MF_ROWS][MfStride<NT>::S];   // node planes: ring slot, pencil y0-1+p, row, z
  float r[3][5][16 * NT];            // element planes: ring slot, pencil y0-1+p, z
  float f[2][4][MF_ROWS][16 * NT];   // right-hand side of the plane being written: slot, pencil y0+p, row, z
  float d[2][4][3][16 * NT];         // its smoother diagonal: slot, pencil y0+p, component, z
};

// Global loads into registers, one row of NZ consecutive z per wavefront and
// instruction (lane = z; the row index is uniform, so its address is scalar): node
// plane X (u: six pencils x nine rows, rows w, w + 4, ... for wavefront w), element
// plane XE (rho: five pencils, pencil w and, for wavefront 0, pencil 4) and, for the
// plane written next, node plane XF's f (pencil y0 + w, nine rows) and Dinv (three rows).
template <int NT>
struct MfPrefetch {
  static constexpr int NU = (6 * MF_ROWS + 3) / 4;
  float u[NU], r[2], f[MF_ROWS], d[3];
};

// Per-wavefront row offsets of the prefetch (uniform, 32-bit: the launch checks that
// 18 Ntot fits), plane offsets added per step.  Rows past the end and load cases past
// ncase read a valid row instead (their values are never stored); lanes past nz read z = 0.
template <int NT>
struct MfRows {
  int u[MfPrefetch<NT>::NU], r[2], f[MF_ROWS], d[3];
};

template <int NT>
__device__ __forceinline__ void mf_rows(MfRows<NT>& ro, const LevelCell& c, int y0, long long Ntot, int j0, int ncase) {
  constexpr int NZ = 16 * NT;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
#pragma unroll
  for (int i = 0; i < MfPrefetch<NT>::NU; ++i) {
    const int row = min(w + 4 * i, 6 * MF_ROWS - 1);
    const int p = row / MF_ROWS, qc = row - MF_ROWS * p;
    const int q = min(qc / 3, ncase - 1 - j0), comp = qc - 3 * (qc / 3);
    ro.u[i] = (int)(((j0 + q) * 3 + comp) * Ntot + c.off) + wrap(y0 - 1 + p, c.ny) * NZ;
  }
#pragma unroll
  for (int i = 0; i < 2; ++i) ro.r[i] = (int)c.off + wrap(y0 - 1 + min(w + 4 * i, 4), c.ny) * NZ;
#pragma unroll
  for (int qc = 0; qc < MF_ROWS; ++qc) {
    const int q = min(qc / 3, ncase - 1 - j0), comp = qc - 3 * (qc / 3);
    ro.f[qc] = (int)(((j0 + q) * 3 + comp) * Ntot + c.off) + (y0 + w) * NZ;
  }
#pragma unroll
  for (int comp = 0; comp < 3; ++comp) ro.d[comp] = (int)(comp * Ntot + c.off) + (y0 + w) * NZ;
}

// Global loads into registers, one row of nz consecutive z per wavefront and
// instruction (lane = z): node plane X (u: six pencils x nine rows, rows w, w + 4, ...
// for wavefront w), element plane XE (rho: five pencils, pencil w and, for wavefront 0,
// pencil 4) and, for the plane written next, node plane XF's f (pencil y0 + w, nine
// rows) and Dinv (three rows).
template <int MODE, int NT>
__device__ __forceinline__ void mf_load(MfPrefetch<NT>& pf, const MfRows<NT>& ro, bool do_u, bool do_r, bool do_f,
                                        int X, int XE, int XF, const LevelCell& c, const float* __restrict__ u,
                                        const float* __restrict__ rho, const float* __restrict__ f,
                                        const float* __restrict__ dinv) {
  constexpr int NZ = 16 * NT;
  const int z = NT == 4 ? (int)(threadIdx.x & 63) : (int)(threadIdx.x & 63) % NZ;
  if (do_u) {
    const int px = X * c.ny * NZ;
#pragma unroll
    for (int i = 0; i < MfPrefetch<NT>::NU; ++i) pf.u[i] = (u + (ro.u[i] + px))[z];
  }
  if (do_r) {
    const int px = XE * c.ny * NZ;
#pragma unroll
    for (int i = 0; i < 2; ++i) pf.r[i] = (rho + (ro.r[i] + px))[z];
  }
  if (MODE >= 2 && do_f) {
    const int px = XF * c.ny * NZ;
#pragma unroll
    for (int qc = 0; qc < MF_ROWS; ++qc) pf.f[qc] = (f + (ro.f[qc] + px))[z];
  }
  if (MODE == 3 && do_f) {
    const int px = XF * c.ny * NZ;
#pragma unroll
    for (int comp = 0; comp < 3; ++comp) pf.d[comp] = (dinv + (ro.d[comp] + px))[z];
  }
}

template <int MODE, int NT>
__device__ __forceinline__ void mf_store(const MfPrefetch<NT>& pf, bool do_u, bool do_r, bool do_f, int us, int rs,
                                         int fs, MfShared<NT>& sh) {
  constexpr int NZ = 16 * NT;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), z = threadIdx.x & 63;
  if (z >= NZ) return;
  if (do_u) {
#pragma unroll
    for (int i = 0; i < MfPrefetch<NT>::NU; ++i) {
      const int row = w + 4 * i;
      if (row < 6 * MF_ROWS) sh.u[us][row / MF_ROWS][row % MF_ROWS][z] = pf.u[i];
    }
  }
  if (do_r) {
#pragma unroll
    for (int i = 0; i < 2; ++i)
      if (w + 4 * i < 5) sh.r[rs][w + 4 * i][z] = pf.r[i];
  }
  if (MODE >= 2 && do_f) {
#pragma unroll
    for (int qc = 0; qc < MF_ROWS; ++qc) sh.f[fs][w][qc][z] = pf.f[qc];
  }
  if (MODE == 3 && do_f) {
#pragma unroll
    for (int comp = 0; comp < 3; ++comp) sh.d[fs][w][comp][z] = pf.d[comp];
  }
}

// B operands of one (b, ks) group: lane 16 k + n holds, per z tile and case, u at the
// node of Ke column 4 ks + k of element z = 16 t + n
template <int NT>
__device__ __forceinline__ void mf_read_b(float (&bu)[NT][MF_NCG], const float* __restrict__ src, int zsh, int j) {
  constexpr int NZ = 16 * NT;
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    int zz = 16 * t + j + zsh;
    zz = zz >= NZ ? zz - NZ : zz;
#pragma unroll
    for (int q = 0; q < MF_NCG; ++q) bu[t][q] = src[q * 3 * MfStride<NT>::S + zz];
  }
}

// One step: element plane ex = x0 - 1 + s, node planes in ring slots s % 3 (ex) and
// (s + 1) % 3 (ex + 1), moduli in slot s % 3, f / Dinv of node plane ex in slot s % 2.
template <int MODE, int NT, int PAR>
__device__ __forceinline__ void mf_step(
    int s, const LevelCell& c, int y, int x0, long long Ntot, int j0, const bool (&live)[MF_NCG],
    const float (&A)[2][2][6], const int (&boff)[2][6], int bsel, int zsh, const MfShared<NT>& sh, float omega,
    float* __restrict__ out, mf_f4 (&acc)[MF_NCG][NT]) {
  constexpr int NZ = 16 * NT;
  const int lane = threadIdx.x & 63;
  const int g = lane >> 4, j = lane & 15;
  const int wy = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int s0 = s % 3, s1 = (s + 1) % 3;
  const float* su0 = &sh.u[s0][0][0][0];
  const float* su1 = &sh.u[s1][0][0][0];
  // element moduli of the two element pencils (y, y - 1 -> window pencils wy + 1, wy), per z tile
  float re[2][NT];
#pragma unroll
  for (int b = 0; b < 2; ++b)
#pragma unroll
    for (int t = 0; t < NT; ++t) re[b][t] = sh.r[s0][wy + 1 - b][16 * t + j];
  // the twelve (b, ks) groups, B operands read one group ahead.  Per element pencil b
  // the six K steps accumulate Ke u_e from zero; the element moduli scale the result's
  // columns (lane n of a tile holds column n: element z = 16 t + n) into acc.
  float bu[2][NT][MF_NCG];
  mf_read_b<NT>(bu[0], ((bsel & 1) ? su1 : su0) + boff[0][0], zsh & 1, j);
#pragma unroll
  for (int b = 0; b < 2; ++b) {
    mf_f4 tb[MF_NCG][NT];
#pragma unroll
    for (int ks = 0; ks < 6; ++ks) {
      const int gi = 6 * b + ks;
      if (gi + 1 < 12) {
        const int b1 = (gi + 1) / 6, k1 = (gi + 1) % 6;
        mf_read_b<NT>(bu[(gi + 1) & 1], ((bsel >> (6 * b1 + k1)) & 1 ? su1 : su0) + boff[b1][k1], (zsh >> k1) & 1, j);
      }
      // keep the next group's LDS reads ahead of this group's products
      __builtin_amdgcn_sched_barrier(0);
      const float a = A[PAR][b][ks];
#pragma unroll
      for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int q = 0; q < MF_NCG; ++q)
          tb[q][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, bu[gi & 1][t][q], ks == 0 ? mf_f4{0.f, 0.f, 0.f, 0.f}
                                                                                     : tb[q][t], 0, 0, 0);
    }
#pragma unroll
    for (int q = 0; q < MF_NCG; ++q)
#pragma unroll
      for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int r = 0; r < 3; ++r) acc[q][t][r] = fmaf(re[b][t], tb[q][t][r], acc[q][t][r]);   // row 3 is zero
  }
  // Row groups 2 PAR (c = 0) and 2 PAR + 1 (c = 1) hold node plane ex (a = 0 rows).  A
  // transpose of the four tile registers over the four 16-lane rows (permlane32 and
  // permlane16 swaps) puts tile t of row group g into lane row t of register g: lane z
  // then holds element z's rows, the c = 0 part of node z and the c = 1 part of node
  // z + 1, which moves one lane up (periodic in z).
  const int g0 = 2 * PAR;
  if (s > 0) {
    const int xn = wrap(x0 - 1 + s, c.nx);   // owned for s = 1..MF_X
    const int z = lane;
    const bool zok = z < NZ;
    const float* sf = &sh.f[s & 1][wy][0][0];
    const float* sd = &sh.d[s & 1][wy][0][0];
    const float* su = &sh.u[s0][wy + 1][0][0];   // u of node pencil (xn, y)
    const long long pen = c.off + ((long long)xn * c.ny + y) * NZ;   // uniform
#pragma unroll
    for (int q = 0; q < MF_NCG; ++q) {
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        unsigned R[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) R[t] = t < NT ? __float_as_uint(acc[q][t < NT ? t : 0][r]) : 0u;
        auto x02 = __builtin_amdgcn_permlane32_swap(R[0], R[2], false, false);
        auto x13 = __builtin_amdgcn_permlane32_swap(R[1], R[3], false, false);
        auto y01 = __builtin_amdgcn_permlane16_swap(x02[0], x13[0], false, false);   // rows 0, 1
        auto y23 = __builtin_amdgcn_permlane16_swap(x02[1], x13[1], false, false);   // rows 2, 3
        const float c0 = __uint_as_float(PAR == 0 ? y01[0] : y23[0]);
        float c1 = __uint_as_float(PAR == 0 ? y01[1] : y23[1]);
        if constexpr (NT == 4) {
          c1 = lane_from_zm1(c1);
        } else {
          const float last = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(c1), NZ - 1));
          c1 = lane_from_zm1(c1);
          c1 = z == 0 ? last : c1;
        }
        const float kv = c0 + c1;
        if (live[q] && zok) {
          const int row = q * 3 + r;
          float* o = out + (long long)((j0 + q) * 3 + r) * Ntot + pen;
          if constexpr (MODE == 0) o[z] = kv;
          else if constexpr (MODE == 2) o[z] = sf[row * NZ + z] - kv;
          else o[z] = su[row * MfStride<NT>::S + z] + omega * sd[r * NZ + z] * (sf[row * NZ + z] - kv);
        }
      }
    }
  }
  // clear the row groups of plane ex (at s = 0: plane x0 - 1, not this workgroup's)
  const bool clr = (g >> 1) == PAR;
#pragma unroll
  for (int q = 0; q < MF_NCG; ++q)
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
      for (int r = 0; r < 3; ++r) acc[q][t][r] = clr ? 0.f : acc[q][t][r];   // row 3 stays zero
}

template <int MODE, int NT>
__global__ __launch_bounds__(256, 2) void k_fine_mfma(
    const LevelCell* __restrict__ cells, const int* __restrict__ blk_start, int ncells,
    long long Ntot, int ncase, const float* __restrict__ Ke, const float* __restrict__ rho,
    const float* __restrict__ u, const float* __restrict__ f, const float* __restrict__ dinv, float omega,
    float* __restrict__ out, const int* __restrict__ done, int ndc, unsigned long long* __restrict__ work) {
  constexpr int NZ = 16 * NT;
  __shared__ MfShared<NT> sh;
  count_work(cells, ncells, done, ndc, ncase, work);   // work: set on one launch of an application
  const int k = find_cell(blk_start, ncells, lbid());
  const LevelCell c = cells[k];
  if (c.nz != NZ) return;                 // this launch serves the cells of NT z tiles
  const int j0 = blockIdx.y * MF_NCG;
  bool live[MF_NCG];
  bool any = false;
#pragma unroll
  for (int q = 0; q < MF_NCG; ++q) {
    live[q] = (j0 + q < ncase) && !case_done(done, ndc, c.id, j0 + q);   // uniform over the block
    any = any || live[q];
  }
  if (!any) return;
  // workgroup -> (x run, four pencils); the whole workgroup stays to the end (barriers)
  const int wb = lbid() - blk_start[k];
  const int quads = c.ny / 4;
  const int y0 = (wb % quads) * 4;
  const int x0 = (wb / quads) * MF_X;
  const int wy = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int y = y0 + wy;
  const int lane = threadIdx.x & 63;
  const int g = lane >> 4, j = lane & 15;
  Ke += (long long)c.id * 576;

  // A operands (lane 16 k + m supplies A[m][k] = Ke[row(m)][4 ks + k]) and the LDS
  // offsets of the B operands (lane 16 k + n reads the node of column 4 ks + k)
  float A[2][2][6];
  int boff[2][6];
  int bsel = 0, zsh = 0;   // bit 6 b + ks: node plane ex + 1; bit ks: node z + 1
  {
    const int gm = j >> 2, rm = j & 3;
#pragma unroll
    for (int ks = 0; ks < 6; ++ks) {
      const int kk = 4 * ks + g;
      const int cm = kk / 3, comp = kk - 3 * (kk / 3);
      zsh |= hex_dz(cm) << ks;
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        bsel |= hex_dx(cm) << (6 * b + ks);
        const int pw = wy + 1 - b + hex_dy(cm);     // window pencil of node y - b + dy
        boff[b][ks] = (pw * MF_ROWS + comp) * MfStride<NT>::S;
#pragma unroll
        for (int par = 0; par < 2; ++par) {
          float a = 0.f;
          if (rm < 3) a = Ke[(3 * hex_corner((gm >> 1) ^ par, b, gm & 1) + rm) * 24 + kk];
          A[par][b][ks] = a;
        }
      }
    }
  }
  mf_f4 acc[MF_NCG][NT];
#pragma unroll
  for (int q = 0; q < MF_NCG; ++q)
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[q][t] = mf_f4{0.f, 0.f, 0.f, 0.f};

  // prologue: node planes x0 - 1, x0 (u slots 0, 1), element plane x0 - 1 (rho slot 0),
  // f / Dinv of node plane x0 (slot 1, written at step 1)
  MfRows<NT> ro;
  mf_rows<NT>(ro, c, y0, Ntot, j0, ncase);
  MfPrefetch<NT> pf;
  mf_load<MODE, NT>(pf, ro, true, true, true, wrap(x0 - 1, c.nx), wrap(x0 - 1, c.nx), x0, c, u, rho, f, dinv);
  mf_store<MODE, NT>(pf, true, true, true, 0, 0, 1, sh);
  mf_load<MODE, NT>(pf, ro, true, false, false, x0, 0, 0, c, u, rho, f, dinv);
  mf_store<MODE, NT>(pf, true, false, false, 1, 0, 0, sh);
  __syncthreads();
  // step s < MF_X prefetches node plane x0 + 1 + s, element plane x0 + s and f / Dinv of
  // node plane x0 + s (the last ones the run reads: node plane x0 + MF_X, element plane
  // x0 + MF_X - 1, f of node plane x0 + MF_X - 1)
  auto step = [&](int s, auto par) {
    constexpr int P = decltype(par)::value;
    const bool pre = s < MF_X;
    const bool pref = s >= 1 && s < MF_X;
    mf_load<MODE, NT>(pf, ro, pre, pre, pref, wrap(x0 + 1 + s, c.nx), wrap(x0 + s, c.nx), x0 + s, c, u, rho, f,
                      dinv);
    mf_step<MODE, NT, P>(s, c, y, x0, Ntot, j0, live, A, boff, bsel, zsh, sh, omega, out, acc);
    mf_store<MODE, NT>(pf, pre, pre, pref, (s + 2) % 3, (s + 1) % 3, (s + 1) & 1, sh);
    __syncthreads();
  };
  step(0, std::integral_constant<int, 0>{});
  for (int s = 1; s <= MF_X; s += 2) {
    step(s, std::integral_constant<int, 1>{});
    step(s + 1, std::integral_constant<int, 0>{});
  }
}

bool fine_mfma_applies(const LevelCell& c) {
  return c.nx == c.ny && c.ny == c.nz && c.nz % 16 == 0 && c.nz <= 64 && c.nx % MF_X == 0 && c.ny % 4 == 0;
}

int fine_mfma_blocks(const LevelCell& c) { return (c.nx / MF_X) * (c.ny / 4); }

template <int MODE, int NT>
static void mf_launch(const Launch& L, int ncase, const float* Ke, const float* rho, const float* u, const float* f,
                      const float* dinv, float omega, float* out, unsigned long long*& work, hipStream_t s) {
  dim3 grid(L.nblocks, (ncase + MF_NCG - 1) / MF_NCG), block(256);
  k_fine_mfma<MODE, NT><<<grid, block, 0, s>>>(L.cells, L.blk_start, L.ncells, L.Ntot, ncase, Ke, rho, u, f, dinv,
                                               omega, out, L.done, L.ndc, work);
  work = nullptr;   // the first launch counts the whole application's work
}

template <int MODE>
static void mf_launch_all(const Launch& L, int ncase, const float* Ke, const float* rho, const float* u,
                          const float* f, const float* dinv, float omega, float* out, hipStream_t s) {
  // one launch per pencil length present in the batch (L.var_mask bit t: cells with t + 1 tiles)
  unsigned long long* work = L.work;
  if (L.var_mask & 8) mf_launch<MODE, 4>(L, ncase, Ke, rho, u, f, dinv, omega, out, work, s);
  if (L.var_mask & 4) mf_launch<MODE, 3>(L, ncase, Ke, rho, u, f, dinv, omega, out, work, s);
  if (L.var_mask & 2) mf_launch<MODE, 2>(L, ncase, Ke, rho, u, f, dinv, omega, out, work, s);
  if (L.var_mask & 1) mf_launch<MODE, 1>(L, ncase, Ke, rho, u, f, dinv, omega, out, work, s);
}

void fine_apply_mfma(const Launch& L, int mode, int ncase, const float* Ke, const float* rho, const float* u,
                     const float* f, const float* dinv, float omega, float* out, hipStream_t s) {
  switch (mode) {
    case 0: mf_launch_all<0>(L, ncase, Ke, rho, u, f, dinv, omega, out, s); break;
    case 2: mf_launch_all<2>(L, ncase, Ke, rho, u, f, dinv, omega, out, s); break;
    default: mf_launch_all<3>(L, ncase, Ke, rho, u, f, dinv, omega, out, s); break;
  }
}
