/* C ABI of the MI355X (gfx950) periodic inverse-homogenization library.
 *
 * Drop-in boundary for the reference's cell solve and homogenization
 * (xing-yuu/MAPLE, src/maple/core).  All array arguments are plain pointers:
 * "dev" arrays live in device memory (e.g. torch.Tensor.data_ptr() of a CUDA
 * tensor), "host" arrays in host memory.  `stream` is a hipStream_t (NULL =
 * the default stream); every kernel of a call is enqueued on it.  Every
 * function returns 0 on success or a negative MAPLE_E* code; the message of
 * the last failure on the calling thread is maple_last_error().  No function
 * aborts or exits the process.
 *
 * Batches: a context holds `ncells` periodic cells, possibly of different
 * resolution.  Per-element and per-node device arrays are the cells'
 * arrays concatenated in cell order; inside a cell the element / node id is
 * (x*ny + y)*nz + z (z fastest), as in build_unitcell_batch
 * (main_IHD_v3.py:175-197).  Multi-case fields are stored plane by plane:
 * value (case j, component c, node n) at [(j*nd + c)*Ntot + n], nd = 3
 * (elastic, 6 cases) or 1 (thermal, 3 cases).
 */
#ifndef MAPLE_HIP_H
#define MAPLE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MAPLE_OK 0
#define MAPLE_EINVAL (-1)      /* bad argument */
#define MAPLE_EDEVICE (-2)     /* no usable GPU / HIP runtime error */
#define MAPLE_ENOMEM (-3)      /* device allocation failed */
#define MAPLE_ENOCONV (-4)     /* CG hit maxit before rtol (results still written) */
#define MAPLE_ENONFINITE (-5)  /* NaN / Inf in the input or the iteration */
#define MAPLE_EBREAKDOWN (-6)  /* CG met p^T K p <= 0 on a live load case (operator or preconditioner
                                  not SPD there): that case stopped unconverged; results still written */

#define MAPLE_ELASTIC 1
#define MAPLE_THERMAL 2

#define MAPLE_PREC_FP64 0      /* plain method: every kernel in FP64 */
#define MAPLE_PREC_MIXED 1     /* FP64 CG (residual, operator, dot products, C_H) refining with an
                                  FP32 multigrid preconditioner (level-0 element contraction,
                                  stencils, smoothing); the coarsest inverse stays FP64 */

typedef struct maple_ctx maple_ctx;

/* Message of the last failure on this thread ("" if none). */
const char* maple_last_error(void);

/* Library version string. */
const char* maple_version(void);

/* Number of HIP devices visible (0 without a GPU); never fails. */
int maple_device_count(void);

/* Element matrices of the reference (isotropic_elastic_tensor / hexahedron /
 * hexahedron_scalar, _utils.py:39-196) for a cube of side 1/r, row-major FP64
 * host arrays: Ke24 (24x24), Fe24 (24x6), X0_24 (24x6), Ke8 (8x8), Fe8 (8x3),
 * X0_8 (8x3).  Pure host code. */
int maple_hex8_matrices(int r, double E0, double nu, double kappa0, double* Ke24, double* Fe24, double* X0_24,
                        double* Ke8, double* Fe8, double* X0_8);

/* Create a context for `ncells` cells with node grids dims[3*i..3*i+2] on
 * `device`, `nlevels` multigrid levels (every dim must be divisible by
 * 2^(nlevels-1)), precision MAPLE_PREC_*.  The element matrices default to the
 * reference's material (E0=1, nu=0.3, kappa0=1, model.py:18, 25).
 * Replaces IHD.__init__ (model.py:13-32) and the per-call hierarchy build of
 * GMGSolver.solve (EBE_GMG.py:790-849). */
int maple_ctx_create(int device, int ncells, const int* dims_host, int nlevels, int precision, maple_ctx** out);
int maple_ctx_destroy(maple_ctx* ctx);

/* Override the element matrices of one cell (host arrays, layouts as in
 * maple_hex8_matrices; any pointer may be NULL to keep that physics). */
int maple_set_element_matrices(maple_ctx* ctx, int cell, const double* Ke24, const double* Fe24, const double* X0_24,
                               const double* Ke8, const double* Fe8, const double* X0_8);

/* Solver knobs: Jacobi weight, sweeps per V-cycle side, and whether each call
 * starts CG from the previous call's fields (1) or from zero (0, the reference,
 * EBE_GMG.py:858). */
int maple_set_solver(maple_ctx* ctx, double omega, int nsmooth, int warm_start);

/* Jacobi sweeps of the V-cycle before (pre >= 1, the first from a zero start)
 * and after (post >= 0) the coarse-grid correction; maple_set_solver sets both
 * to nsmooth.  Unequal counts make the V-cycle nonsymmetric: use flexible CG
 * (maple_set_cg(ctx, 1); the mixed form's default).  The reference's recursion
 * uses smooth_iter per level (EBE_GMG.py:886-984). */
int maple_set_sweeps(maple_ctx* ctx, int pre, int post);

/* Weights of the smoother sweeps: with lo = 0 (the default) every sweep takes the
 * solver's omega; with 0 < lo < 1 the k sweeps of each side take the reciprocals of the
 * k Chebyshev nodes of [lo, 1], which damp the error over that part of the l1-Jacobi
 * spectrum (contained in (0, 1]) by the degree-k Chebyshev polynomial instead of
 * (1 - omega lambda)^k, for the same operator applications.  Pre-sweeps use the weights
 * in rising, post-sweeps in falling order; with equal counts the V-cycle stays
 * symmetric.  Needs the l1-Jacobi smoother.  The reference smooths with 8-colour
 * Gauss-Seidel or weighted Jacobi (EBE_GMG.py:395-459); this has no counterpart there. */
int maple_set_chebyshev(maple_ctx* ctx, double lo);

/* Smoother of the V-cycle: MAPLE_SMOOTH_JACOBI (weighted point Jacobi; the
 * reference's smooth_jacobi, EBE_GMG.py:395-411) or MAPLE_SMOOTH_L1_JACOBI
 * (the diagonal replaced by the absolute row sum, convergent for every
 * density field; the default). */
#define MAPLE_SMOOTH_JACOBI 0
#define MAPLE_SMOOTH_L1_JACOBI 1
int maple_set_smoother(maple_ctx* ctx, int kind);

/* Island correction of the preconditioner (off by default, the reference's
 * V-cycle alone).  On: every homogenization finds the small solid islands of
 * the density field -- connected clusters (26-neighbour) of at most 16 elements
 * with rho > threshold and at most 32 nodes, touching the rest of the cell only
 * through elements at or below the threshold -- inverts each island's block of
 * the level-0 operator (its nodes, every element touching them) in FP64, and
 * adds z += A_i^{-1} r|_i to every preconditioner application.  The
 * preconditioner stays SPD, so CG converges to the same tolerance; islands held
 * by 1e-9 voids otherwise cost CG iterations for each of their soft modes.  Up
 * to 256 islands per cell are corrected; the first call with on = 1 allocates
 * the island buffers.  No counterpart in the reference, whose V-cycle
 * (EBE_GMG.py:886-984) has no such term. */
int maple_set_islands(maple_ctx* ctx, int on, double threshold);
/* Islands corrected in the last homogenization (all cells). */
int maple_num_islands(const maple_ctx* ctx);

/* CG variant: 0 = Fletcher-Reeves beta, 1 = flexible (Polak-Ribiere) beta, which
 * stays valid for a preconditioner that is not exactly symmetric (one extra dot
 * product per iteration), -1 = flexible in MAPLE_PREC_MIXED only (default),
 * 2 = block CG: the load cases of a cell share one Krylov space (search
 * directions K-conjugate block to block, orthonormalised with dependent
 * directions dropped), so a slow mode common to the cases is paid for once;
 * a cell stops when all its cases meet the tolerance.  Block steps solve with
 * P^T K P, whose conditioning follows the design's contrast, so after 200 block
 * iterations per-case CG takes over from the block's fields.  The reference
 * solves the cases one after another (model.py:111-124). */
int maple_set_cg(maple_ctx* ctx, int flexible);

/* When a homogenization asks for both physics, solve them at the same time:
 * the thermal solve runs on a second stream of the context, driven by a second
 * host thread, and the call returns when both are done (on = 1, the default), or
 * one after the other on the caller's stream (on = 0). */
int maple_set_concurrency(maple_ctx* ctx, int on);

/* Run the CG iterations of maple_homogenize as hipGraphs (on = 1) or as single
 * launches (on = 0, the default).  Each iteration is two captured segments (the
 * operator, step length, update and convergence check; the preconditioner, next
 * direction) replayed on the caller's stream; the convergence poll and the launch
 * tables of the live cells stay on the host between them, and a segment is captured
 * again whenever what its launches were captured with changes (the live cells, the
 * FP64 fallback, the islands, any setting).  Same arithmetic in the same order as the
 * single launches: results are bitwise equal.  With graphs on, homogenize copies rho
 * into the context first (one more device buffer of the level-0 size).  Kernel timing
 * (maple_timing_enable) runs the iterations as single launches.  The reference
 * launches every op of its V-cycle from Python (EBE_GMG.py:886-984); this has no
 * counterpart there. */
int maple_set_graphs(maple_ctx* ctx, int on);
/* Segments captured and graph launches made so far by the context (either may be NULL). */
int maple_graph_counts(const maple_ctx* ctx, long long* captures, long long* launches);

/* MAPLE_PREC_MIXED only: a solve that has not converged after `iters` CG
 * iterations builds the FP64 hierarchy of the same density and finishes on the
 * FP64 V-cycle (0 = never; default 300).  An FP32 V-cycle cannot resolve the
 * soft modes of solid islands held by 1e-9 voids, where it stagnates (no
 * counterpart in the reference, whose solver runs in FP64 only). */
int maple_set_mixed_fallback(maple_ctx* ctx, int iters);

/* Rebuild the coarsest-level dense inverse every `every` homogenizations (1 =
 * always, the default).  In between, the previous density's SPD inverse serves
 * as the coarsest solve of the V-cycle: the preconditioner stays SPD and CG
 * still stops at the requested tolerance (no counterpart in the reference,
 * which rebuilds its whole hierarchy on every solve, EBE_GMG.py:790-849). */
int maple_set_coarse_refresh(maple_ctx* ctx, int every);

/* Start each solve projected on the last m solutions (0 = off, the default; at most
 * 5).  Per (cell, case) the solve starts from the combination of those solutions
 * and a supplied start (maple_set_start) that minimises the energy norm of the
 * error, (u* - u0)^T K (u* - u0): c solves the m x m (or m+1) Galerkin system
 * (h_a^T K h_b) c = h_a^T f, leaving out vectors that are numerically combinations
 * of newer ones, so the previous fields (warm start) and the supplied start are
 * special cases (with a history the previous fields are always in the basis, whatever
 * maple_set_solver's warm_start says).  Each solve costs one operator application per
 * basis vector; the context holds 2m+1 copies of each physics' fields.  Solves by the per-case CG
 * only (not block CG).  Every call clears the history.  No counterpart in the
 * reference, which starts every solve from zero (EBE_GMG.py:852-858). */
int maple_set_start_history(maple_ctx* ctx, int m);

/* Kernel variants for the tuner (maple_amd/tune.py times them and keeps the
 * fastest): MAPLE_KERNEL_FINE_OPERATOR (the level-0 elastic operator of
 * EBEOperator.matvec, EBE_GMG.py:412-430), variant 0 (element rows per pencil
 * lane, all neighbour values held in registers), 1 (the same, loads staged
 * element by element, fewer live registers), 2 (node blocks: each thread forms
 * its node's 27 3x3 coupling blocks from the 8 element moduli around it and
 * applies them to all six load cases), 3 (node blocks, three load cases per
 * thread), 4 or 5 (node blocks as 2 or 3 for two x-adjacent nodes per thread,
 * which gather their shared neighbours once), 6 or 7 (2x2 nodes in x and z per
 * thread, three or two load cases) or 8 (as 5 with the z neighbours taken from
 * the neighbouring lanes of a wavefront, which holds one z-pencil; cells with
 * nz = 64, else 5 is used) or 9 (the FP32 operator of the mixed form's
 * preconditioner as element products Ke (rho_e u_e) on the matrix cores,
 * v_mfma_f32_16x16x4f32, one wavefront per node pencil marching along x; cubic
 * cells with r a multiple of 16 up to 64; the FP64 operator runs as 8, or 5).
 * Default: 8. */
/* MAPLE_KERNEL_COARSE_OPERATOR: the elastic Galerkin stencil product of the
 * coarse levels (the coarse matvec of the V-cycle, EBE_GMG.py:886-984), variant
 * 0 (one thread per node row, three times the waves), 1 (one thread per node,
 * each neighbour gathered once for the three rows) or 2 (1 for launches of 256
 * or more blocks, 0 below; the default). */
/* MAPLE_KERNEL_THERMAL_OPERATOR: the level-0 thermal operator
 * (EBEOperatorScalar.matvec, EBE_GMG_scalar.py:292-306), variant 0 or 1 (element
 * rows per pencil lane, as variants 0 and 1 above), 2 or 3 (node form: all load
 * cases per thread, z neighbours from the neighbouring lanes, the element matrix's
 * four Hamming-distance entries instead of 64; 2: two x-adjacent nodes per thread,
 * 3: one; cells with nz = 64 whose thermal element matrix is a Q1 Laplacian, else
 * variant 0 runs).  Default: 2. */
#define MAPLE_KERNEL_FINE_OPERATOR 0
#define MAPLE_KERNEL_COARSE_OPERATOR 1
#define MAPLE_KERNEL_THERMAL_OPERATOR 2
int maple_set_variant(maple_ctx* ctx, int kernel, int variant);

/* Homogenize: for every cell, solve the 6 elastic unit-strain (and/or 3 thermal
 * unit-gradient) periodic cell problems K(rho) u = F(rho) by multigrid-
 * preconditioned CG to ||r|| <= max(rtol ||f||, 1e-14 ||Fe||_F sqrt(E)),
 * then reduce C_H = sum_e rho_e (X0-u_e)^T Ke (X0-u_e) and kappa_H likewise.
 *   rho_dev   : Etot doubles (SIMP-scaled element moduli, > 0)
 *   physics   : MAPLE_ELASTIC | MAPLE_THERMAL
 *   CH_dev    : ncells*36 doubles (row-major 6x6 per cell) or NULL
 *   kappa_dev : ncells*9 doubles or NULL
 *   iters_host: ncells*9 ints (6 elastic then 3 thermal CG counts) or NULL
 * Replaces IHD.TO_homo_3d / TO_homo_1d (model.py:98-159) with
 * GMGSolver.solve / GMGSolverScalar.solve (EBE_GMG.py:758-885,
 * EBE_GMG_scalar.py:617-725) and homogenized_batch_3d / _thermal
 * (model.py:33-95).  Returns MAPLE_ENOCONV if any case hit maxit, MAPLE_EBREAKDOWN if
 * any case broke down (both still write C_H / kappa_H of the fields reached). */
int maple_homogenize(maple_ctx* ctx, const double* rho_dev, int physics, double rtol, int maxit, double* CH_dev,
                     double* kappa_dev, int* iters_host, void* stream);

/* Element sensitivities of J from the last maple_homogenize:
 *   dJdrho[e] = <G6_cell, S_e> + <G3_cell, T_e>,  S_e = (X0-u_e)^T Ke (X0-u_e)
 * G6_dev (ncells*36) = dJ/dC_H, G3_dev (ncells*9) = dJ/dkappa (either NULL if
 * that physics is unused).  The converged form of autograd.grad(J, rho)
 * through TO_homo_3d / TO_homo_1d (main_IHD_v3.py:688). */
int maple_sensitivity(maple_ctx* ctx, const double* G6_dev, const double* G3_dev, double* dJdrho_dev, void* stream);

/* Periodic box mean of odd width k over each cell (avg_filter3d_x_periodic,
 * main_IHD_v3.py:342-359).  in and out: Etot doubles, may not alias. */
int maple_filter_box(maple_ctx* ctx, const double* in_dev, double* out_dev, int k, void* stream);

/* Symmetry projection of an element field (symmetrize_vec, main_IHD_v3.py:221-337):
 * kind 0 none (copy), 1 mirror about the centre planes of the axes in the bit
 * mask `axes` (1 x, 2 y, 4 z), 2 the 48 symmetries of the cube (cubic cells).
 * in and out: Etot doubles, may not alias.  The projection is self-adjoint, so
 * it also maps dJ/dx_sym to dJ/dx. */
int maple_symmetrize(maple_ctx* ctx, int kind, int axes, const double* in_dev, double* out_dev, void* stream);

/* OC update with nbisect bisection steps on lambda (oc_update,
 * main_IHD_v3.py:362-380): volfrac/move/xmin are per-cell host arrays. */
int maple_oc_update(maple_ctx* ctx, const double* x_dev, const double* dJdx_dev, const double* volfrac_host,
                    const double* move_host, const double* xmin_host, int nbisect, double* xnew_dev, void* stream);

/* The OC update of maple_oc_update without a context and without allocating:
 * no counterpart in the reference; for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  `ncells` cells with node grids dims_host[3*i..3*i+2] on
 * `device`, arrays as in maple_oc_update (x, dJdx, xnew: device, all cells
 * concatenated; volfrac/move/xmin: per-cell host arrays).  The scratch (bisection
 * state, partial sums, launch tables, about 4.3 MB) is a static device array that
 * each device holds until the process ends; calls on one device share it, so calls
 * that overlap on one device (two streams or threads) race.  At most
 * MAPLE_NOALLOC_MAX_CELLS cells and MAPLE_NOALLOC_MAX_ELEMENTS elements in all
 * (else MAPLE_EINVAL, nothing launched).  Same results as maple_oc_update, bit
 * for bit. */
#define MAPLE_NOALLOC_MAX_CELLS 64
#define MAPLE_NOALLOC_MAX_ELEMENTS 8388608LL
int maple_oc_update_noalloc(int device, int ncells, const int* dims_host, const double* x_dev, const double* dJdx_dev,
                            const double* volfrac_host, const double* move_host, const double* xmin_host, int nbisect,
                            double* xnew_dev, void* stream);

/* ---- additive-manufacturing penalties ---------------------------------------
 * Differentiable penalties on the SIMP density of every cell (cubic cells, grid
 * (x, y, z) with the build direction +z), each with its gradient dP/drho, for
 * J_total = J - lambda_print P_print - lambda_cavity P_cavity (main_IHD_v3.py:686).
 * The gradients are the subgradients the reference's autograd takes through
 * max-pooling (first maximum of a window in raster order), minimum (ties halved)
 * and relu.  P_dev: ncells doubles; dP_dev: Etot doubles or NULL (value only). */

/* printability_penalty_overhang (printability.py:91-134): a layer is supported by
 * the dilated support of the layer below (lateral radius ceil(tan(90 - deg)),
 * at most 7; periodic_xy wraps the layer); cells up to 64^3. */
int maple_penalty_overhang(maple_ctx* ctx, const double* rho_dev, double overhang_deg, int periodic_xy, double* P_dev,
                           double* dP_dev, void* stream);

#define MAPLE_SEED_ALL 0
#define MAPLE_SEED_X 1
#define MAPLE_SEED_Y 2
#define MAPLE_SEED_Z 3
#define MAPLE_SEED_XY 4     /* the +/-z faces (drainage along z), as the reference's "xy" */
typedef struct {
  double void_thresh;       /* soft void indicator sigmoid(beta (void_thresh - rho)); 0.5 */
  double beta;              /* 20 */
  int min_channel_rad;      /* soft opening radius before the flood fill, 0..3 (0: none) */
  int rad;                  /* flood-fill neighbourhood radius, 0 or 1 */
  int iters;                /* flood-fill steps, 0: the supercell's largest side */
  int periodic;             /* one-cell form: periodic padding (1) or outside = 0 (0) */
  int seed_faces;           /* MAPLE_SEED_* */
  int tiles[3];             /* all 1: cavity_penalty_closed_void (printability.py:139-223);
                               else cavity_penalty_periodic_array (printability.py:229-323) on
                               the unit cell tiled tiles[0] x tiles[1] x tiles[2], outside = 0 */
  int focus_center;         /* periodic array with every tile count >= 3: count the centre cell only */
} maple_cavity_opts;

/* Configure the trapped-void penalty and allocate its buffers (per voxel of every
 * cell's supercell: 9 doubles, 2 shorts and one byte per flood-fill step). */
int maple_set_cavity(maple_ctx* ctx, const maple_cavity_opts* opts);
/* The trapped-void penalty as configured by maple_set_cavity. */
int maple_penalty_cavity(maple_ctx* ctx, const double* rho_dev, double* P_dev, double* dP_dev, void* stream);

/* RHS F(rho) of one physics (assemble_F_fast / assemble_F_fast_scalar,
 * _utils.py:453-507), in the plane-major layout. */
int maple_rhs(maple_ctx* ctx, int physics, const double* rho_dev, double* F_dev, void* stream);

/* Fine-level operator v = K(rho) u for `ncase` fields (EBEOperator.matvec,
 * EBE_GMG.py:412-430; EBEOperatorScalar.matvec, EBE_GMG_scalar.py:292-306). */
int maple_apply_operator(maple_ctx* ctx, int physics, const double* rho_dev, int ncase, const double* u_dev,
                         double* v_dev, void* stream);

/* Copy the fields of the last solve (plane-major, 6*3 or 3*1 planes of Ntot)
 * into u_dev, or seed the next solve's start from u_dev. */
int maple_get_fields(maple_ctx* ctx, int physics, double* u_dev, void* stream);
int maple_set_fields(maple_ctx* ctx, int physics, const double* u_dev, void* stream);

/* Supply a start for the next solve of one physics (plane-major fields at level
 * 0, as maple_get_fields).  Per (cell, case) that solve starts from whichever of
 * zero, the previous fields (when warm start is on) and this start has the
 * smallest residual ||f - K u||, so a poor start costs one operator application
 * and nothing else (no counterpart in the reference: its warm-start hook
 * init_list, EBE_GMG.py:648-757 and 852-858, is never fed, model.py:120). */
int maple_set_start(maple_ctx* ctx, int physics, const double* u_dev, void* stream);

/* Start chosen by the last solve of one physics, per (cell, case): 0 zero,
 * 1 previous fields, 2 supplied start, 3 projected start (maple_set_start_history)
 * (ncells*6 or ncells*3 ints). */
int maple_start_sources(maple_ctx* ctx, int physics, int* src_host);

/* fine_dev = P coarse_dev: periodic trilinear prolongation of plane-major
 * fields (all cases of one physics) from level `level` (>= 1) to level-1
 * (ProlongStencil.prolong, EBE_GMG.py:38-58). */
int maple_prolong(maple_ctx* ctx, int physics, int level, const double* coarse_dev, double* fine_dev, void* stream);

/* Sizes: total elements (= nodes) at level 0 over all cells; nodes at a level. */
long long maple_num_elements(const maple_ctx* ctx);
long long maple_num_nodes(const maple_ctx* ctx, int level);

/* Kernel timing with HIP events on the call's stream around every launch of
 * the level-0 operator kernels (enable, then read: count of launches, their
 * summed duration in ms and, if `work` is not NULL, the elements x load cases
 * those launches processed -- converged cells and cases are skipped, so a
 * launch's work varies; reading resets).  kind 0: FP64 elastic, 1: FP64
 * thermal, 2: FP32 elastic (mixed form's preconditioner).  No counterpart in
 * the reference. */
int maple_timing_enable(maple_ctx* ctx, int enable);
int maple_timing_read(maple_ctx* ctx, int kind, long long* count, double* total_ms, long long* work);

/* ---- multigrid internals ----------------------------------------------------
 * The hierarchy of one density field, piece by piece, for checks against a
 * dense Galerkin construction and for tools.  `prec` selects the FP64
 * hierarchy (MAPLE_PREC_FP64, every context) or the FP32 preconditioner of a
 * MAPLE_PREC_MIXED context; inputs and outputs are FP64 device arrays in the
 * plane-major layout of the level, every load case of the physics at once.
 * These calls clear the convergence flags of the last solve of that physics. */

/* Build the hierarchy of rho (Jacobi diagonals, Galerkin stencils, coarsest
 * inverse) without solving; the FP64 level 0 keeps reading rho_dev, which must
 * stay valid while the level-0 operator is used.  Replaces the per-solve build
 * of GMGSolver.solve (EBE_GMG.py:790-849). */
int maple_mg_setup(maple_ctx* ctx, int physics, const double* rho_dev, int prec, void* stream);

/* Assembled Galerkin operator P^T A P of level `level` >= 1 as the library's
 * 27-point block stencil: S[(d*nd*nd + i*nd + j)*N + n] is entry (i, j) of the
 * block coupling node n to node n + off(d), d = (dx+1)*9 + (dy+1)*3 + (dz+1)
 * (27*nd*nd*N_level values).  The reference's sparse_coarsen_ke
 * (EBE_GMG.py:226-337) builds the same operator as per-element matrices. */
int maple_mg_stencil(maple_ctx* ctx, int physics, int prec, int level, double* S_dev, void* stream);

/* Inverse smoother diagonal of level `level` < nlevels-1 (nd*N_level values):
 * 1/diag(A) for MAPLE_SMOOTH_JACOBI (_compute_diag_inverse, EBE_GMG.py:375-393,
 * zero -> 1), 1/(absolute row sum) for MAPLE_SMOOTH_L1_JACOBI. */
int maple_mg_dinv(maple_ctx* ctx, int physics, int prec, int level, double* dinv_dev, void* stream);

/* Operator of one level on every load case: mode 0 out = A u, 2 out = f - A u,
 * 3 out = u + omega Dinv (f - A u) (one smoother sweep; smooth_jacobi,
 * EBE_GMG.py:395-411).  Level 0 is the matrix-free voxel operator
 * (EBEOperator.matvec, EBE_GMG.py:412-430), the others their stencils. */
int maple_mg_apply(maple_ctx* ctx, int physics, int prec, int level, int mode, const double* u_dev,
                   const double* f_dev, double* out_dev, void* stream);

/* coarse = P^T fine: periodic full-weighting restriction from level-1 to
 * `level` (ProlongStencil.restrict_PT, EBE_GMG.py:59-79). */
int maple_mg_restrict(maple_ctx* ctx, int physics, int level, const double* fine_dev, double* coarse_dev,
                      void* stream);

/* z = M r: one application of the V-cycle preconditioner at level 0 (no island
 * correction), as inside the CG (GMGSolver._v_cycle, EBE_GMG.py:886-984). */
int maple_mg_vcycle(maple_ctx* ctx, int physics, int prec, const double* r_dev, double* z_dev, void* stream);

/* The coarsest level's dense inverse (A + gamma Z Z^T)^-1 of one cell, Z the
 * normalised rigid translations and gamma the mean diagonal of A, row-major
 * (order x order doubles, order = maple_mg_coarsest_order).  The reference's
 * coarsest level is smoothing only (EBE_GMG.py:918-929). */
int maple_mg_coarsest_inverse(maple_ctx* ctx, int physics, int cell, double* Ainv_dev, void* stream);
int maple_mg_coarsest_order(const maple_ctx* ctx, int physics, int cell);

/* ---- learned warm start ---------------------------------------------------
 * A 3-level periodic 3-D U-Net (channels 16-32-64, 3x3x3 convolutions with
 * circular padding, 2x2x2 mean pooling, nearest x2 upsampling, skip
 * concatenation, 1x1x1 head) that maps a density feature on an n^3 grid to the
 * 21 characteristic-field channels (elastic case j component c at 3*j + c,
 * thermal case j at 18 + j), trained and evaluated by hand-written BF16 MFMA
 * kernels with FP32 accumulation.  No counterpart in the reference: its
 * warm-start hook `init_list` (EBE_GMG.py:648-757) is never fed (model.py:120,
 * 149).  Tensors are channels-last: value (sample b, voxel (x,y,z), channel c)
 * at ((b*n + x)*n + y)*n + z)*C + c.  Parameters are one FP32 vector: per layer
 * in the order e1a e1b e2a e2b b1 b2 d2a d2b d1a d1b head, W[co][o][ci]
 * (o = (dx+1)*9 + (dy+1)*3 + (dz+1)) then b[co]. */
typedef struct maple_unet maple_unet;

/* Parameter count of the network. */
long long maple_unet_num_params(void);

/* Network on `device` for grids of n^3 voxels (n a multiple of 4 in [16, 128])
 * and batches of up to max_batch samples; every buffer is allocated here. */
int maple_unet_create(int device, int n, int max_batch, maple_unet** out);
int maple_unet_destroy(maple_unet* net);

/* Copy the FP32 parameter vector in (resets the Adam state) or out. */
int maple_unet_set_params(maple_unet* net, const float* theta_host);
int maple_unet_get_params(maple_unet* net, float* theta_host);

/* feat_dev: batch*n^3 FP32 -> out_dev: batch*n^3*21 FP32 (raw network output). */
int maple_unet_forward(maple_unet* net, const float* feat_dev, int batch, float* out_dev, void* stream);

/* Loss weighting: 0 = plain mean((y - t)^2); 1 = mean(w_v (y - t)^2) with
 * w_v = f_v^3 + 1e-3 per voxel, f the (BF16-rounded) input feature, which is
 * about the density under SIMP p = 3: the solver's residual sees a field error
 * in proportion to the local stiffness. */
int maple_unet_set_loss(maple_unet* net, int weighted);

/* Loss (per maple_unet_set_loss) over batch*n^3*21 values against BF16 targets
 * (raw bits, same layout as the output) and its gradient (FP32, parameter order)
 * into grad_dev (may be NULL).  Synchronizes the stream to read the loss. */
int maple_unet_loss_grad(maple_unet* net, const float* feat_dev, const uint16_t* target_dev, int batch,
                         double* loss_host, float* grad_dev, void* stream);

/* One Adam step (beta 0.9/0.999, eps 1e-8, learning rate lr) on the loss above;
 * loss_host receives the loss before the step. */
int maple_unet_train_step(maple_unet* net, const float* feat_dev, const uint16_t* target_dev, int batch, float lr,
                          double* loss_host, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* MAPLE_HIP_H */
