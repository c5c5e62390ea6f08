// Kernel launchers of the cell-solver library (device code lives in *.hip).
#pragma once

#include <string>

#include "common.h"

// A launch over every cell of a batch: blk_start (device, ncells+1 entries)
// maps a block to its cell; the blocks of one cell are contiguous.
struct Launch {
  const LevelCell* cells = nullptr;
  const int* blk_start = nullptr;
  int ncells = 0;
  int nblocks = 0;
  long long Ntot = 0;
  // optional convergence flags done[cell * ndc + case]: blocks of converged
  // (cell, case) pairs return at once, so a batch pays only for the cases
  // still iterating
  const int* done = nullptr;
  int ndc = 0;
  // kernel variant of the level-0 elastic operator, picked by the tuner: element rows
  // per pencil lane (0: neighbour values in registers, 1: staged element by element)
  // or node blocks (2: six load cases per thread, 3: three; 4, 5: the same with two
  // x-adjacent nodes per thread; 6, 7: 2x2 nodes in x and z, three / two cases;
  // 8: as 5 with z neighbours taken from the neighbouring lanes, cells with nz = 64;
  // 9: FP32 element products on the matrix cores, FP64 as 8 or 5)
  int var = 0;
  // MFMA operator launch: bit t set when the listed cells include cells of nz = 16 (t + 1)
  int var_mask = 0;
  // level-0 operator timing: when set, the launch adds the elements x load cases it
  // processes (listed cells, cases not converged) to *work
  unsigned long long* work = nullptr;
};

__device__ __forceinline__ bool case_done(const int* done, int ndc, int ci, int j) {
  return done != nullptr && done[ci * ndc + j] != 0;
}
__device__ __forceinline__ bool cell_done(const int* done, int ndc, int ci) {
  if (done == nullptr) return false;
  for (int j = 0; j < ndc; ++j)
    if (!done[ci * ndc + j]) return false;
  return true;
}

// Records the calling thread's last error (maple_last_error) and returns code.
int maple_record_error(int code, const std::string& msg);

// ---- level 0 (fine.hip) ---------------------------------------------------
int fine_waves(const LevelCell& c);  // wavefronts per cell of k_fine_apply (4 per block)

// mode 0: out = K u; 1: out = K u, partial[blk*6+j] = (u_j, K u_j) over the block;
// 2: out = f - K u; 3: out = u + omega Dinv (f - K u)
template <int ND, typename T>
void fine_apply(const Launch& L, int mode, int ncase, const T* Ke, const T* rho, const T* u,
                const T* f, const T* dinv, T omega, T* out, double* partial, hipStream_t s);
// Node-block form of the elastic operator (variants 2 and 3), same modes; L is the
// node launch (256 consecutive nodes per block).
template <typename T>
void fine_apply_node(const Launch& L, int mode, int ncase, const T* Ke, const T* rho, const T* u,
                     const T* f, const T* dinv, T omega, T* out, double* partial, hipStream_t s);
// Node blocks, 2x1 (variants 4, 5) or 2x2 (6, 7) nodes per thread; L: blocks of 256
// threads over the cell's nodes / (nodes per thread) threads.
template <typename T>
void fine_apply_nodeb(const Launch& L, int mode, int ncase, const T* Ke, const T* rho, const T* u,
                      const T* f, const T* dinv, T omega, T* out, double* partial, hipStream_t s);
// Variant 8: as variant 5 for batches whose cells all have nz = 64 (a wavefront is
// one z-pencil; z neighbours come from the neighbouring lanes).  L: the pair launch.
template <typename T>
void fine_apply_nodeb64(const Launch& L, int mode, int ncase, const T* Ke, const T* rho, const T* u,
                        const T* f, const T* dinv, T omega, T* out, double* partial, hipStream_t s);
// Variant 9, FP32 only: the element product Ke (rho_e u_e) on v_mfma_f32_16x16x4f32, a
// wavefront per node pencil marching along x (modes 0, 2, 3).  Cubic cells with
// r % 16 == 0 and r <= 64 (fine_mfma_applies); L: the launch of fine_mfma_blocks
// blocks (4 wavefronts) per cell.
bool fine_mfma_applies(const LevelCell& c);
int fine_mfma_blocks(const LevelCell& c);
void fine_apply_mfma(const Launch& L, int mode, int ncase, const float* Ke, const float* rho, const float* u,
                     const float* f, const float* dinv, float omega, float* out, hipStream_t s);
// Thermal operator in node form for nz = 64 cells whose element matrix is the Q1
// Laplacian (entry k[d] for corners at Hamming distance d; coef: 4 per cell), same
// modes, all (<= 3) load cases per thread; L: the pair launch.  face0: k[1] = 0 in every
// cell (the face neighbours couple nothing and are skipped).
template <typename T>
void thermal_apply64(const Launch& L, int mode, int ncase, const double* coef, const T* rho, const T* u, const T* f,
                     const T* dinv, T omega, T* out, double* partial, hipStream_t s, bool face0);
void fine_rhs(const Launch& L, int nd, int ncase, const double* Fe, const double* rho, double* F, hipStream_t s);
template <typename T>
void fine_dinv(const Launch& L, int nd, bool l1, const double* Ke, const double* rho, T* dinv, hipStream_t s);

// ---- coarse levels (coarse.hip) --------------------------------------------
// Stencil S of a level, symmetric half: S[((d-13)*ND*ND + i*ND + j)*Ntot + n] is the (i,j)
// entry of the block coupling node n to node n+off(d), d = (dx+1)*9 + (dy+1)*3 + (dz+1),
// for d = 13..26; block d < 13 of node n is the transpose of block 26-d of n+off(d).
template <int ND, typename T>
void stencil_apply(const Launch& L, int mode, int ncase, const T* S, const T* u, const T* f,
                   const T* dinv, T omega, T* out, double* partial, hipStream_t s);
template <typename T>
void stencil_dinv(const Launch& L, int nd, bool l1, const T* S, T* dinv, hipStream_t s);
// All 27 blocks of a half-stored stencil, out[(d*nd*nd + i*nd + j)*Ntot + n] (FP64).
template <typename T>
void stencil_expand(const Launch& L, int nd, const T* S, double* out, hipStream_t s);
// level-1 stencil straight from the voxel densities: block(N, d) =
// sum over coarse elements E containing N and N+d of sum_c rho_child(E,c) G_c[..]
template <typename T>
void stencil_level1(const Launch& Lc, const Launch& Lf, int nd, const double* G, const double* rho, T* S, hipStream_t s);
// Galerkin P^T A P on stencils (fine stencil Sf -> coarse stencil Sc)
template <typename T>
void stencil_galerkin(const Launch& Lc, const Launch& Lf, int nd, const T* Sf, T* Sc, hipStream_t s);
// b_c = P^T r_f (periodic trilinear); out_f += P e_c
template <typename T>
void restrict_pt(const Launch& Lc, const Launch& Lf, int nd, int ncase, const T* rf, T* bc, hipStream_t s);
template <typename T>
void prolong_add(const Launch& Lf, const Launch& Lc, int nd, int ncase, const T* ec, T* uf, hipStream_t s);

// ---- dense coarsest level (dense.hip) --------------------------------------
struct DenseCell {
  int n;            // matrix order (ND * nodes)
  int nt;           // tiles per side (n / 64 rounded up)
  long long aoff;   // offset of the n x n row-major matrix in the dense buffer
  long long voff;   // first node of the cell in the coarsest-level arrays
};
// gamma: ncells doubles of scratch (the constraint weight of each cell)
template <typename T>
void dense_assemble_g(const Launch& Lc, const DenseCell* dcells, int maxn, int nd, const T* S, double* A,
                      double* gamma, hipStream_t s);
// In-place blocked Gauss-Jordan inverse of SPD matrices, one step per pivot tile.
void dense_inverse(const DenseCell* dcells, int ncells, int max_nt, double* A, double* work, hipStream_t s);
template <typename T>
void dense_apply(const Launch& Lc, const DenseCell* dcells, int max_n, int nd, int ncase, const double* Ainv,
                 const T* b, T* x, hipStream_t s);

// ---- vectors and PCG scalars (vec.hip) --------------------------------------
struct CaseScalars {   // one per (cell, case)
  double ff;           // ||f||^2
  double tgt2;         // squared residual target
  double rz, rz_new, pq, rr;
  double rzo;          // (r_new, z_old): the flexible (Polak-Ribiere) correction of beta
  double rrs;          // squared residual of a supplied start (learned warm start)
  double alpha, beta;
  int done, iters;
  int src;             // start of the last solve: 0 zero, 1 previous fields, 2 supplied start,
                       // 3 projected on earlier solutions
  int brk;             // stopped early: 1 breakdown (p^T K p <= 0 on MAPLE_CG_RESTARTS directions in
                       // a row), 2 stagnation (no new residual minimum in MAPLE_CG_STALL iterations,
                       // or the residual 1e6 times above its minimum)
  int nrst;            // consecutive restarts (p^T K p <= 0): the next direction is p = z
  int it_min;          // iteration of the smallest residual so far
  double rr_min;       // smallest squared residual so far
  double zmean[3];     // per component: mean of the preconditioned residual over the cell's nodes
};
#define MAPLE_CG_RESTARTS 3
#define MAPLE_CG_STALL 400
// Block CG state of one cell (bcg.hip): Gram matrices of the current step (two
// NC x NC, row-major, at 0 and 36), the step matrices, the Cholesky factor of
// P^T K P, the orthonormalising transform and the active directions.
struct BlockScalars {
  double gram[72];
  double alpha[36], beta[36], T[36], lg[36];
  int active[MAPLE_MAX_CASES];
  int done, pad_;
};
// ng = 2: A^T B0, A^T B1 into bs[].gram[0..36), [36..72); ng = 1: A^T B0 and the
// per-(case, component) sums of A (its translation content) at gram[36..)
void bcg_gram(const Launch& L, int nd, int nc, int ng, const double* A, const double* B0, const double* B1,
              double* partial, BlockScalars* bs, hipStream_t s);
// alpha = (P^T Q)^-1 P^T R from gram = (P^T Q, P^T R)
void bcg_alpha(int ncells, int nc, BlockScalars* bs, hipStream_t s);
// X += P alpha; R -= Q alpha; partial[blk*6 + j] = (r_j, r_j)
void bcg_update_xr(const Launch& L, int nd, int nc, const BlockScalars* bs, const double* P, const double* Q,
                   double* X, double* R, double* partial, hipStream_t s);
// per cell: done when every case meets its target; iteration counts
void bcg_check(int ncells, int nc, BlockScalars* bs, CaseScalars* sc, int* done_flags, hipStream_t s);
// beta = -(P^T Q)^-1 Q^T Z from gram = Z^T Q
void bcg_beta(int ncells, int nc, BlockScalars* bs, hipStream_t s);
// P = orth(Z - translations + P beta) (first: no P term); the translation sums of Z
// must be in bs[].gram[36..) (bcg_gram with ng = 1)
void bcg_new_direction(const Launch& L, int nd, int nc, bool first, BlockScalars* bs, const double* Z, double* P,
                       double* partial, hipStream_t s);
void bcg_init(int ncells, int nc, BlockScalars* bs, CaseScalars* sc, int* done_flags, hipStream_t s);

// sum partial[blk*6 + j] over each cell's blocks into dst (a member of CaseScalars)
// (over the cells the launch lists, with the launch that wrote the partials)
void reduce_partials(const Launch& L, int ncase, const double* partial, CaseScalars* sc, int which, hipStream_t s);
// SC_RR_LIVE: as SC_RR, but only for (cell, case) pairs not yet stopped (a stopped
// pair's last residual stays; the update passes skip it and sum zero for it)
enum { SC_FF = 0, SC_RZ_NEW, SC_PQ, SC_RR, SC_RZO, SC_RRS, SC_RR_LIVE };
// elementwise (cell-aligned 256-node blocks):
template <typename T>
void vec_scale_dinv(const Launch& L, int nd, int ncase, T omega, const T* dinv, const T* b, T* x, hipStream_t s);
// partial (a_j, b_j)
template <typename Ta, typename Tb>
void vec_dot(const Launch& L, int nd, int ncase, const Ta* a, const Tb* b, double* partial, hipStream_t s);
// u += alpha p; r -= alpha q; partial (r, r)
void pcg_update_ur(const Launch& L, int nd, int ncase, const CaseScalars* sc, const double* p, const double* q,
                   double* u, double* r, double* partial, hipStream_t s);
// Start projected on earlier solutions: basis vectors h_a and their products g_a = K h_a
// (device pointers, a < k <= MAPLE_PROJ_MAX).  Per (cell, case) the combination
// sum_a c_a h_a that minimises the energy norm of the error (G c = b with G_ab = (h_a, K h_b),
// b_a = (h_a, f); vectors whose Cholesky pivot falls below MAPLE_PROJ_TOL G_aa are left out)
// goes to us, its residual f - sum_a c_a g_a to rs, (rs, rs) per case to partial.
// gpart: L.nblocks x ncase x MAPLE_PROJ_NG doubles; coef: ncells x ncase x MAPLE_PROJ_MAX.
#define MAPLE_PROJ_MAX 6
#define MAPLE_PROJ_NG (MAPLE_PROJ_MAX * (MAPLE_PROJ_MAX + 1) / 2 + MAPLE_PROJ_MAX)
#define MAPLE_PROJ_TOL 1e-10
struct ProjBasis {
  const double* h[MAPLE_PROJ_MAX];
  const double* g[MAPLE_PROJ_MAX];
};
int proj_start(const Launch& L, int nd, int ncase, int k, const ProjBasis& B, const double* f, double* gpart,
               double* coef, double* us, double* rs, double* partial, hipStream_t s);
// Start of a solve, per (cell, case): the candidate with the smallest residual
// among zero (residual f), the previous fields u (residual r, if have_prev) and a
// supplied start us (residual rs, if have_start); u and r receive the choice.
// start_src: the source recorded for the supplied start (2 supplied, 3 projected).
void pcg_pick_start(const Launch& L, int ncells, int nd, int ncase, bool have_prev, bool have_start, CaseScalars* sc,
                    double* u, double* r, const double* us, const double* rs, const double* f, hipStream_t s,
                    int start_src = 2);
// (r_j, z_j) into partial[blk*6 + j], and the sums of z over the block's nodes per (case,
// component) into zpart[blk*18 + j*nd + i]
template <typename Tz>
void pcg_dot_rz(const Launch& L, int nd, int ncase, const double* r, const Tz* z, double* partial, double* zpart,
                hipStream_t s);
// sc.zmean of every (cell, case) the launch lists: its zpart sums over the cell's blocks / nodes
void pcg_reduce_zmean(const Launch& L, int nd, int ncase, const double* zpart, CaseScalars* sc, hipStream_t s);
// p = (z - zmean) + beta p (p not read where beta = 0): the search directions stay free of
// the rigid translations, the null space of the periodic operator, so p^T K p measures
// the direction's energy instead of the rounding of a translation
void pcg_update_p(const Launch& L, int nd, int ncase, const CaseScalars* sc, const double* z, double* p, hipStream_t s);
// mixed form, (nd, ncase) = (3, 6) or (1, 3) only (else -1, nothing launched): the
// update also writes r32 = (float) r and, if zold is not null, partial2 = (r, zold)
int pcg_update_ur_mixed(const Launch& L, int nd, int ncase, const CaseScalars* sc, const double* p, const double* q,
                        double* u, double* r, float* r32, const float* zold, double* partial, double* partial2,
                        hipStream_t s);
int pcg_update_p32(const Launch& L, int nd, int ncase, const CaseScalars* sc, const float* z, double* p,
                   hipStream_t s);
// per (cell, case) scalar steps
void pcg_init_scalars(int ncells, int ncase, double rtol, const double* atol2, CaseScalars* sc, int* done_flags,
                      hipStream_t s);
// alpha = rz / pq.  A live case with pq <= 0 (or NaN) takes no step and restarts
// from p = z (beta = 0) and raises *restart_flag; after MAPLE_CG_RESTARTS such
// directions in a row it is a breakdown: stopped (done, brk) with its fields and
// residual as they are
void pcg_alpha(int ncells, int ncase, CaseScalars* sc, int* done_flags, int* restart_flag, hipStream_t s);
void pcg_check(int ncells, int ncase, CaseScalars* sc, int* done_flags, hipStream_t s);
// flexible: beta = (rz_new - rzo) / rz (valid for a variable or slightly non-symmetric preconditioner)
void pcg_beta(int ncells, int ncase, bool flexible, CaseScalars* sc, hipStream_t s);
template <typename Td, typename Ts>
void vec_copy(long long n, const Ts* src, Td* dst, hipStream_t s);
template <typename T>
void vec_zero(long long n, T* dst, hipStream_t s);
// dst = src on the (cell, case) pairs a launch with convergence flags still iterates
template <typename Td, typename Ts>
void vec_copy_cases(const Launch& L, int nd, int ncase, const Ts* src, Td* dst, hipStream_t s);

// ---- homogenization, sensitivities, filter, OC (homog.hip) -----------------
// S_e packed upper triangle (nu = nc(nc+1)/2 values per element) and
// per-block partials of rho_e S_e
void element_energy(const Launch& L, int nd, int ncase, const double* Ke, const double* X0, const double* rho,
                    const double* u, double* Se, double* partial, hipStream_t s);
void reduce_energy(const int* blk_start, int ncells, int nu, const double* partial, int nc, double* CH, hipStream_t s);
void sensitivity(const Launch& L, int nu3, const double* Se3, const double* G6, int nu1, const double* Se1,
                 const double* G3, double* dJdrho, hipStream_t s);
void box_filter(const Launch& L, int k, const double* in, double* out, hipStream_t s);
// mean over reflections (kind 1, axes bit mask) or the 48 cube symmetries (kind 2)
void symmetrize(const Launch& L, int kind, int axes, const double* in, double* out, hipStream_t s);
struct OCState {
  double l1, l2, last_lam;
  double volfrac, move, xmin;
  double lam[64];      // the 63 midpoints of this pass (heap order, index 1..63)
  double c[64];        // 1/sqrt(lam)
};
// a = x sqrt(max(dJ, 1e-12)); 256-element blocks
void oc_prepare(const Launch& L, const double* x, const double* dJ, double* a, hipStream_t s);
void oc_start(int ncells, OCState* st, hipStream_t s);
// 1024-element blocks; partial[blk*64 + t]
void oc_pass(const Launch& L, const OCState* st, const double* x, const double* a, double* partial, hipStream_t s);
void oc_walk(const int* blk_start, const LevelCell* cells, int ncells, int depth, const double* partial,
             OCState* st, hipStream_t s);
void oc_final(const Launch& L, const OCState* st, const double* x, const double* a, double* xnew, hipStream_t s);
// the same passes forming a = x sqrt(max(dJ, 1e-12)) on the fly from dJ (no a array)
void oc_pass_dj(const Launch& L, const OCState* st, const double* x, const double* dJ, double* partial,
                hipStream_t s);
void oc_final_dj(const Launch& L, const OCState* st, const double* x, const double* dJ, double* xnew, hipStream_t s);

// ---- AM penalties (penalty.hip) ----------------------------------------------------
// Overhang penalty of every cell (cubic, r <= 64, rad <= 7): P[cell]; dP/drho into grad
// (Etot, may be null); tape: Etot doubles (the support layers, kept for the adjoint).
void overhang_penalty(const LevelCell* cells, int ncells, int rmax, const double* rho, int rad, int periodic,
                      double* tape, double* P, double* grad, hipStream_t s);
// Trapped-void penalty: buffers over the cells' supercells (unit cell tiled tx x ty x tz)
struct CavityBuffers {
  Launch L;                 // supercell voxels, 256 per block (L.cells: one supercell per cell)
  double beta = 20.0, thr = 0.5;
  int mcr = 0, rad = 1, nsteps = 1, periodic = 1, faces = 0, focus = 0;
  int* region = nullptr;    // 3 per cell: the unit cell's sides (the counted centre tile)
  double *voidv = nullptr, *vc = nullptr, *er = nullptr, *reach[2] = {nullptr, nullptr};
  double *g[2] = {nullptr, nullptr}, *gvc = nullptr, *gvoid = nullptr, *partial = nullptr;
  short *argE = nullptr, *argD = nullptr;
  unsigned char* tape = nullptr;   // nsteps x voxels
};
// P[cell] and (grad not null) dP/drho over the unit cells; Lu: the level-0 node launch
int cavity_penalty(const CavityBuffers& B, const LevelCell* uc, const Launch& Lu, const double* rho, double* P,
                   double* grad, hipStream_t s);

// ---- island correction of the preconditioner (islands.hip) -------------------
#define MAPLE_ISL_MAXE 16     // elements of an island
#define MAPLE_ISL_MAXN 32     // nodes of an island (its block is at most 96 x 96)
#define MAPLE_ISL_ROUNDS 8    // label propagation rounds (even)
struct Island {
  int cell;                      // cell id
  int ne, nn;                    // elements; nodes (0: not corrected)
  int pad_;
  int elems[MAPLE_ISL_MAXE];     // global element ids, sorted
  int nodes[MAPLE_ISL_MAXN];     // global node ids, sorted
};
// L0: the full level-0 node launch; lab, lab2, cnt, bad, slot: Etot ints each;
// nisl (device) counts the islands found (slots beyond maxi are dropped)
void island_detect(const Launch& L0, const double* rho, double thr, int* lab, int* lab2, int* cnt, int* bad,
                   int* slot, Island* isl, int* nisl, int maxi, hipStream_t s);
void island_nodes(const LevelCell* cells, Island* isl, int n, hipStream_t s);
// inv: n x (nd*MAPLE_ISL_MAXN)^2 doubles; valid: n ints
void island_setup(int nd, const LevelCell* cells, Island* isl, int n, const double* Ke, const double* rho,
                  double* inv, int* valid, hipStream_t s);
void island_apply(int nd, int ncase, const Island* isl, const int* valid, int n, const double* inv, long long Ntot,
                  const int* done, int ndc, const double* r, double* z, hipStream_t s);
void island_apply(int nd, int ncase, const Island* isl, const int* valid, int n, const double* inv, long long Ntot,
                  const int* done, int ndc, const double* r, float* z, hipStream_t s);
