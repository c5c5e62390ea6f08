// C ABI of the library: contexts, the multigrid hierarchy, the multi-RHS
// preconditioned CG driver, homogenization, sensitivities, filter and OC.
//
// One context = one batch of periodic cells on one device.  Everything a
// density field needs (stencils of every level, the dense coarsest inverse,
// the CG vectors, the element energies) is allocated once at creation and
// reused by every call; calls allocate nothing.  Each call runs on the
// caller's stream; the only host synchronisations are the CG convergence
// polls (one-iteration lag, so the device never waits for the host) and the
// final read of the iteration counts.
#include "../../include/maple_hip.h"

#include <algorithm>
#include <initializer_list>
#include <type_traits>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "element.h"
#include "kernels.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(x)                                                                                  \
  do {                                                                                             \
    hipError_t e_ = (x);                                                                           \
    if (e_ != hipSuccess) return fail(MAPLE_EDEVICE, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define TRY(x)             \
  do {                     \
    int rc_ = (x);         \
    if (rc_ != 0) return rc_; \
  } while (0)

struct Level {
  std::vector<LevelCell> h;
  LevelCell* d = nullptr;
  long long Ntot = 0;
  int* dblk = nullptr;    // 256-node blocks
  int* dfblk = nullptr;   // level 0: operator blocks (4 waves)
  int* doblk = nullptr;   // level 0: 1024-element blocks (OC)
  int* dxblk = nullptr;   // level 0: 256 node pairs per block (2x1 node-block operator)
  int* dqblk = nullptr;   // level 0: 256 node quads per block (2x2 node-block operator)
  int* dmblk = nullptr;   // level 0: 4 node pencils x 16 planes per block (MFMA operator)
  Launch node, fine, oc, pair, quad, mfma;
  std::vector<int> cnt_node, cnt_fine, cnt_pair, cnt_quad, cnt_mfma;   // blocks per cell of the launches
};

// Multigrid hierarchy and V-cycle work vectors in precision T (FP64 for the
// plain method, FP32 for the mixed form's preconditioner).
template <typename T>
struct VC {
  std::vector<T*> S, dinv, X, B, R, W;
  T* Ke = nullptr;    // level-0 element matrices, per cell
  T* rho = nullptr;   // level-0 densities (the FP64 form uses the caller's array)
};

// CG's preconditioned residual: FP64 (FP64 V-cycle) or, in the mixed form, the FP32
// V-cycle's own output, never widened (the CG reads it as FP32).
struct ZVec {
  double* d = nullptr;
  float* f = nullptr;
};

// One captured segment of a CG iteration (hipGraph) and what the capture decided on the
// host: the key of everything baked into its kernel arguments, the preconditioned
// residual's buffer and the form of the fused update.
struct SegGraph {
  hipGraphExec_t exec = nullptr;
  unsigned long long key = ~0ULL;
  ZVec z;
  bool r32_ready = false, rzo_done = false;
};

struct Phys {
  int nd = 0, nc = 0;
  double *Ke = nullptr, *Fe = nullptr, *X0 = nullptr, *G = nullptr;   // per cell on device
  double* lap = nullptr;   // thermal: Q1-Laplacian coefficients k[0..3] per cell (node-form operator)
  bool laplacian = false;  // thermal: every cell's Ke is a Q1 Laplacian (k[Hamming distance])
  bool face0 = false;      // and its distance-1 entry is zero (to 1e-14 of the largest entry)
  std::vector<double> hKe, hFe, hX0;                                   // per cell on host
  VC<double> v64;
  VC<float> v32;
  double* z64 = nullptr;   // mixed form: the FP32 preconditioner's output, widened
  double *u = nullptr, *r = nullptr, *p = nullptr, *q = nullptr, *f = nullptr;
  double *A = nullptr, *work = nullptr, *gamma = nullptr;
  std::vector<DenseCell> hdc;
  DenseCell* ddc = nullptr;
  int maxn = 0, max_nt = 0;
  double* Se = nullptr;
  CaseScalars* sc = nullptr;
  std::vector<CaseScalars> hsc;
  int* done_d = nullptr;   // ncells*nc convergence flags, then the solve's CG restart flag
  int* done_h = nullptr;   // pinned, 2 slots of ncells*nc + 1
  std::vector<Launch> ln;  // per-level node launches that skip converged (cell, case) pairs
  Launch lf;               // level-0 operator launch, same
  Launch lx, lq, lm;       // level-0 node-pair, node-quad and MFMA operator launches, same
  double* atol2 = nullptr;
  bool have_fields = false, have_energy = false;
  double* ustart = nullptr;   // supplied start of the next solve
  double* partial = nullptr;  // reduction scratch of this physics' solve (the two solves may run concurrently)
  hipEvent_t poll[2] = {nullptr, nullptr};
  BlockScalars* bs = nullptr; // block CG state, one per cell
  // launch tables listing only the cells still iterating (per level, and the
  // level-0 operator's), their pinned host staging, and how many cells they list
  std::vector<LevelCell*> cc;
  std::vector<int*> cb;
  LevelCell* ccf = nullptr;
  int* cbf = nullptr;
  LevelCell* ccx = nullptr;
  int* cbx = nullptr;
  LevelCell* ccq = nullptr;
  int* cbq = nullptr;
  LevelCell* ccm = nullptr;
  int* cbm = nullptr;
  char* stage = nullptr;
  size_t stage_bytes = 0;
  int stage_slots = 0, stage_next = 0, nlive = 0;
  bool have_start = false;
  // start projected on earlier solutions (maple_set_start_history): a ring of the last
  // hist_m solutions, the products K h of a solve's basis, the Gram partials and the
  // coefficients
  double* hist = nullptr;
  double* khist = nullptr;
  double* gpart = nullptr;
  double* pcoef = nullptr;
  int nhist = 0, hpos = 0;
  double* isl_inv = nullptr;  // island correction: inverse blocks of this physics, per island
  int* isl_valid = nullptr;
  long long setups = 0;    // hierarchy builds so far (drives the coarsest-inverse refresh)
  bool built[2] = {false, false};   // a hierarchy exists: [0] FP64, [1] FP32
  // CG iterations as hipGraphs (maple_set_graphs): the two segments of an iteration for
  // the full launch tables ([0]) and for the current compacted ones ([1]), captured on
  // this physics' own stream (a capture cannot run on a null stream)
  SegGraph ga[2], gb[2];
  hipStream_t cap = nullptr;
  unsigned long long epoch = 0;   // compactions so far in this solve (0: full tables)
};

struct Timer {
  std::vector<hipEvent_t> ev;
  size_t used = 0;
};

}  // namespace

struct maple_ctx {
  int device = 0, ncells = 0, nlevels = 0, prec = 0;
  bool all_nz64 = false;   // every cell has nz = 64 (level-0 operator variant 8 applies)
  bool mfma_ok = false;    // every cell fits the MFMA operator (variant 9)
  int thermal_var = 2;     // thermal level-0 operator: 0/1 element rows, 2 node form where it applies
  std::vector<int> dims;
  std::vector<Level> lv;
  Phys ph[2];
  double* partial = nullptr;
  double omega = 1.6;   // l1-Jacobi: the spectrum of Dinv K lies in (0, 1], any omega < 2 converges
  int nsmooth = 2;          // Jacobi sweeps before the coarse correction (the first from zero)
  int npost = 2;            // and after it
  double cheb_lo = 0.0;     // > 0: Chebyshev weights over [cheb_lo, 1] of the l1-Jacobi spectrum
  std::vector<double> wpre, wpost;   // weight of each pre / post sweep
  int warm = 0;
  int smoother = MAPLE_SMOOTH_L1_JACOBI;
  int flexible = -1;   // -1: flexible CG in the mixed form only; 2: block CG
  int block_max = 200; // block CG iterations before per-case CG takes over
  int coarse_refresh = 1;   // rebuild the coarsest inverse every this many homogenizations
  int fallback = 300;       // mixed form: FP64 V-cycle after this many CG iterations (0: never)
  int hist_m = 0;           // > 0: solves start projected on the last hist_m solutions
  std::vector<void*> hist_allocs;
  OCState* ocst = nullptr;
  OCState* ocst_h = nullptr;
  double* oc_a = nullptr;
  hipStream_t side = nullptr;                  // the thermal solve's stream when both physics run
  hipEvent_t fork = nullptr, join = nullptr;
  bool concurrent = true;                      // solve elastic and thermal at the same time
  // island correction of the preconditioner (islands.hip)
  bool islands = false;
  double isl_thr = 1e-4;
  int isl_max = 0;                              // island slots (0: not allocated)
  int *isl_lab = nullptr, *isl_lab2 = nullptr, *isl_cnt = nullptr, *isl_bad = nullptr, *isl_slot = nullptr;
  int* isl_count_d = nullptr;
  int* isl_count_h = nullptr;                   // pinned
  Island* isl = nullptr;
  int nisl = 0;                                 // islands of the current density
  // AM penalties (penalty.hip)
  double* oh_tape = nullptr;                    // overhang: the support layers (Etot)
  CavityBuffers cav;
  bool cav_ready = false;
  std::vector<void*> cav_allocs;                // freed when the configuration changes
  bool timing = false;
  int graphs = 0;            // capture the CG iterations as hipGraphs
  double* rho_own = nullptr;   // graphs on: the density of the current homogenization (a fixed address)
  unsigned long long gen = 0;   // bumped by every setting that changes an iteration's kernels
  long long graph_captures = 0, graph_launches = 0;
  Timer timers[3];   // 0: FP64 elastic level-0 operator, 1: FP64 thermal, 2: FP32 elastic
  unsigned long long* work_d = nullptr;   // per timer: elements x load cases of the timed launches
  std::vector<void*> allocs;
  std::vector<void*> host_allocs;
};

namespace {

template <typename T>
int dalloc(maple_ctx* c, T** p, size_t n) {
  *p = nullptr;
  if (n == 0) n = 1;
  if (hipMalloc((void**)p, n * sizeof(T)) != hipSuccess) {
    (void)hipGetLastError();
    return fail(MAPLE_ENOMEM, "hipMalloc of " + std::to_string(n * sizeof(T)) + " bytes failed");
  }
  c->allocs.push_back((void*)*p);
  return 0;
}

int set_device(const maple_ctx* c) {
  HIPCHK(hipSetDevice(c->device));
  return 0;
}

// per_block == 0: blocks per cell from the MFMA operator's wavefront split
int make_launch(maple_ctx* c, Level& L, int per_block, bool fine_waves_mode, int** dptr, Launch* out,
                std::vector<int>* counts = nullptr) {
  std::vector<int> blk(c->ncells + 1, 0);
  if (counts) counts->assign(c->ncells, 0);
  for (int i = 0; i < c->ncells; ++i) {
    const LevelCell& lc = L.h[i];
    long long n = fine_waves_mode ? (long long)fine_waves(lc) : (long long)lc.nx * lc.ny * lc.nz;
    if (per_block == 0) blk[i + 1] = blk[i] + (fine_mfma_applies(lc) ? fine_mfma_blocks(lc) : 0);
    else blk[i + 1] = blk[i] + (int)((n + per_block - 1) / per_block);
    if (counts) (*counts)[i] = blk[i + 1] - blk[i];
  }
  TRY(dalloc(c, dptr, blk.size()));
  HIPCHK(hipMemcpy(*dptr, blk.data(), blk.size() * sizeof(int), hipMemcpyHostToDevice));
  out->cells = L.d;
  out->blk_start = *dptr;
  out->ncells = c->ncells;
  out->nblocks = blk.back();
  out->Ntot = L.Ntot;
  return 0;
}

int upload_phys_constants(maple_ctx* c, Phys& P) {
  const int ne = 8 * P.nd;
  std::vector<double> G((size_t)c->ncells * 8 * ne * ne);
  for (int i = 0; i < c->ncells; ++i) child_galerkin(P.nd, &P.hKe[(size_t)i * ne * ne], &G[(size_t)i * 8 * ne * ne]);
  HIPCHK(hipMemcpy(P.Ke, P.hKe.data(), P.hKe.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(P.Fe, P.hFe.data(), P.hFe.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(P.X0, P.hX0.data(), P.hX0.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(P.G, G.data(), G.size() * sizeof(double), hipMemcpyHostToDevice));
  if (P.v32.Ke) {
    std::vector<float> k32(P.hKe.begin(), P.hKe.end());
    HIPCHK(hipMemcpy(P.v32.Ke, k32.data(), k32.size() * sizeof(float), hipMemcpyHostToDevice));
  }
  // absolute residual floor per cell: (1e-14 ||Fe||_F sqrt(E))^2
  std::vector<double> at(c->ncells);
  for (int i = 0; i < c->ncells; ++i) {
    double fn = 0.0;
    for (int k = 0; k < ne * P.nc; ++k) fn += P.hFe[(size_t)i * ne * P.nc + k] * P.hFe[(size_t)i * ne * P.nc + k];
    const double E = (double)c->lv[0].h[i].nx * c->lv[0].h[i].ny * c->lv[0].h[i].nz;
    at[i] = 1e-28 * fn * E;
  }
  HIPCHK(hipMemcpy(P.atol2, at.data(), at.size() * sizeof(double), hipMemcpyHostToDevice));
  if (P.nd == 1) {
    // node-form thermal operator: Ke[a][b] = k[d(a, b)], d the Hamming distance of corners a, b
    std::vector<double> lap((size_t)c->ncells * 4, 0.0);
    bool ok = true, f0 = true;
    for (int i = 0; i < c->ncells && ok; ++i) {
      const double* K = &P.hKe[(size_t)i * 64];
      double* k = &lap[(size_t)i * 4];
      bool seen[4] = {false, false, false, false};
      double scale = 0.0;
      for (int e = 0; e < 64; ++e) scale = std::max(scale, std::fabs(K[e]));
      for (int a = 0; a < 8 && ok; ++a)
        for (int b = 0; b < 8 && ok; ++b) {
          const int d = (hex_dx(a) != hex_dx(b)) + (hex_dy(a) != hex_dy(b)) + (hex_dz(a) != hex_dz(b));
          if (!seen[d]) {
            k[d] = K[a * 8 + b];
            seen[d] = true;
          } else if (std::fabs(K[a * 8 + b] - k[d]) > 1e-14 * scale) {
            ok = false;
          }
        }
      if (std::fabs(k[1]) > 1e-14 * scale) f0 = false;
    }
    P.laplacian = ok;
    P.face0 = ok && f0;
    HIPCHK(hipMemcpy(P.lap, lap.data(), lap.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  return 0;
}

int alloc_phys(maple_ctx* c, Phys& P, int nd, int nc) {
  P.nd = nd;
  P.nc = nc;
  const int ne = 8 * nd;
  const int L = c->nlevels;
  TRY(dalloc(c, &P.Ke, (size_t)c->ncells * ne * ne));
  TRY(dalloc(c, &P.Fe, (size_t)c->ncells * ne * nc));
  TRY(dalloc(c, &P.X0, (size_t)c->ncells * ne * nc));
  TRY(dalloc(c, &P.G, (size_t)c->ncells * 8 * ne * ne));
  TRY(dalloc(c, &P.atol2, (size_t)c->ncells));
  TRY(dalloc(c, &P.lap, (size_t)c->ncells * 4));
  auto alloc_vc = [&](auto& V) -> int {
    using TT = typename std::remove_pointer<typename std::decay<decltype(V.S[0])>::type>::type;
    V.S.assign(L, nullptr);
    V.dinv.assign(L, nullptr);
    V.X.assign(L, nullptr);
    V.B.assign(L, nullptr);
    V.R.assign(L, nullptr);
    V.W.assign(L, nullptr);
    for (int l = 0; l < L; ++l) {
      const long long N = c->lv[l].Ntot;
      const size_t nv = (size_t)nc * nd * N;
      if (l >= 1) TRY(dalloc(c, (TT**)&V.S[l], (size_t)14 * nd * nd * N));   // symmetric half
      if (l < L - 1) {
        TRY(dalloc(c, (TT**)&V.dinv[l], (size_t)nd * N));
        TRY(dalloc(c, (TT**)&V.R[l], nv));
        TRY(dalloc(c, (TT**)&V.W[l], nv));
      }
      TRY(dalloc(c, (TT**)&V.X[l], nv));
      TRY(dalloc(c, (TT**)&V.B[l], nv));
    }
    return 0;
  };
  // the mixed form keeps an FP64 hierarchy too: it takes over a solve whose FP32
  // preconditioner stagnates
  TRY(alloc_vc(P.v64));
  if (c->prec != MAPLE_PREC_FP64) {
    TRY(alloc_vc(P.v32));
    TRY(dalloc(c, &P.v32.Ke, (size_t)c->ncells * ne * ne));
    TRY(dalloc(c, &P.v32.rho, (size_t)c->lv[0].Ntot));
    TRY(dalloc(c, &P.z64, (size_t)nc * nd * c->lv[0].Ntot));
  }
  const size_t n0 = (size_t)nc * nd * c->lv[0].Ntot;
  TRY(dalloc(c, &P.u, n0));
  TRY(dalloc(c, &P.r, n0));
  TRY(dalloc(c, &P.p, n0));
  TRY(dalloc(c, &P.q, n0));
  TRY(dalloc(c, &P.f, n0));
  TRY(dalloc(c, &P.ustart, n0));
  TRY(dalloc(c, &P.bs, (size_t)c->ncells));
  HIPCHK(hipMemset(P.u, 0, n0 * sizeof(double)));
  // dense coarsest level
  const Level& LC = c->lv[L - 1];
  P.hdc.resize(c->ncells);
  long long aoff = 0;
  for (int i = 0; i < c->ncells; ++i) {
    const LevelCell& lc = LC.h[i];
    DenseCell dc;
    dc.n = nd * lc.nx * lc.ny * lc.nz;
    dc.nt = (dc.n + 63) / 64;
    dc.aoff = aoff;
    dc.voff = lc.off;
    aoff += (long long)dc.n * dc.n;
    P.hdc[i] = dc;
    P.maxn = std::max(P.maxn, dc.n);
    P.max_nt = std::max(P.max_nt, dc.nt);
  }
  if (P.maxn > 12288) return fail(MAPLE_EINVAL, "coarsest level too large for the dense solve (" + std::to_string(P.maxn) + " unknowns); use more levels");
  TRY(dalloc(c, &P.A, (size_t)aoff));
  TRY(dalloc(c, &P.work, (size_t)c->ncells * 64 * 64));
  TRY(dalloc(c, &P.gamma, (size_t)c->ncells));
  TRY(dalloc(c, &P.ddc, (size_t)c->ncells));
  HIPCHK(hipMemcpy(P.ddc, P.hdc.data(), P.hdc.size() * sizeof(DenseCell), hipMemcpyHostToDevice));
  TRY(dalloc(c, &P.Se, (size_t)nc * (nc + 1) / 2 * c->lv[0].Ntot));
  TRY(dalloc(c, &P.sc, (size_t)c->ncells * nc));
  TRY(dalloc(c, &P.done_d, (size_t)c->ncells * nc + 1));
  P.ln.resize(L);
  for (int l = 0; l < L; ++l) {
    P.ln[l] = c->lv[l].node;
    P.ln[l].done = P.done_d;
    P.ln[l].ndc = nc;
  }
  P.lf = c->lv[0].fine;
  P.lf.done = P.done_d;
  P.lf.ndc = nc;
  P.lx = c->lv[0].pair;
  P.lx.done = P.done_d;
  P.lx.ndc = nc;
  P.lq = c->lv[0].quad;
  P.lq.done = P.done_d;
  P.lq.ndc = nc;
  P.lm = c->lv[0].mfma;
  P.lm.done = P.done_d;
  P.lm.ndc = nc;
  P.cc.assign(L, nullptr);
  P.cb.assign(L, nullptr);
  for (int l = 0; l < L; ++l) {
    TRY(dalloc(c, &P.cc[l], (size_t)c->ncells));
    TRY(dalloc(c, &P.cb[l], (size_t)c->ncells + 1));
  }
  TRY(dalloc(c, &P.ccf, (size_t)c->ncells));
  TRY(dalloc(c, &P.cbf, (size_t)c->ncells + 1));
  TRY(dalloc(c, &P.ccx, (size_t)c->ncells));
  TRY(dalloc(c, &P.cbx, (size_t)c->ncells + 1));
  TRY(dalloc(c, &P.ccq, (size_t)c->ncells));
  TRY(dalloc(c, &P.cbq, (size_t)c->ncells + 1));
  TRY(dalloc(c, &P.ccm, (size_t)c->ncells));
  TRY(dalloc(c, &P.cbm, (size_t)c->ncells + 1));
  auto r8 = [](size_t b) { return (b + 7) / 8 * 8; };
  P.stage_bytes = (size_t)(L + 4) * (r8(sizeof(LevelCell) * c->ncells) + r8(sizeof(int) * (c->ncells + 1)));
  P.stage_slots = c->ncells + 2;   // a solve compacts at most ncells - 1 times: no slot is reused in flight
  void* sp = nullptr;
  HIPCHK(hipHostMalloc(&sp, P.stage_bytes * P.stage_slots, hipHostMallocDefault));
  c->host_allocs.push_back(sp);
  P.stage = (char*)sp;
  P.nlive = c->ncells;
  P.hsc.resize((size_t)c->ncells * nc);
  void* hp = nullptr;
  HIPCHK(hipHostMalloc(&hp, sizeof(int) * 2 * (c->ncells * nc + 1), hipHostMallocDefault));
  c->host_allocs.push_back(hp);
  P.done_h = (int*)hp;
  return 0;
}

// ---- solver pieces ---------------------------------------------------------

// Weights of the smoother sweeps: omega for every sweep, or, with the Chebyshev schedule
// on, the reciprocals of the Chebyshev nodes of [lo, 1] (the l1-Jacobi spectrum of D^-1 K
// lies in (0, 1]): k sweeps x <- x + w_i D^-1 (f - K x) then damp the error by the degree-k
// Chebyshev polynomial of that interval, max |p| = 1 / T_k((1 + lo) / (1 - lo)), for the
// same operator applications as k Jacobi sweeps.  Pre-sweeps take the weights in rising
// order, post-sweeps in falling order; equal counts keep the V-cycle symmetric (the error
// propagators are polynomials in D^-1 K).
std::vector<double> sweep_weights(const maple_ctx* c, int k, bool rising) {
  std::vector<double> w(std::max(k, 1), c->omega);
  if (c->cheb_lo <= 0.0 || k < 1) return w;
  const double a = c->cheb_lo, b = 1.0, pi = 3.14159265358979323846;
  for (int i = 0; i < k; ++i) w[i] = 1.0 / (0.5 * (b + a) + 0.5 * (b - a) * std::cos((2.0 * i + 1.0) * pi / (2.0 * k)));
  std::sort(w.begin(), w.end());
  if (!rising) std::reverse(w.begin(), w.end());
  return w;
}

void set_sweep_weights(maple_ctx* c) {
  c->wpre = sweep_weights(c, c->nsmooth, true);
  c->wpost = sweep_weights(c, c->npost, false);
}

void timed_begin(maple_ctx* c, int kind, hipStream_t s) {
  if (!c->timing) return;
  Timer& t = c->timers[kind];
  if (t.used + 2 > t.ev.size()) {
    for (int k = 0; k < 256; ++k) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return;
      t.ev.push_back(e);
    }
  }
  (void)hipEventRecord(t.ev[t.used], s);
}

void timed_end(maple_ctx* c, int kind, hipStream_t s) {
  if (!c->timing) return;
  Timer& t = c->timers[kind];
  if (t.used + 2 > t.ev.size()) return;
  (void)hipEventRecord(t.ev[t.used + 1], s);
  t.used += 2;
}

// Launch of the level-0 operator of physics P: the pencil launch of the element-row
// kernels, or the node launch of the node-block kernels (elastic, variants 2-3).
// full: every cell and case (no convergence flags).  fp32: the FP32 operator (variant 9
// applies); the FP64 operator runs variant 9's choice as variant 8 (or 5).
Launch op_launch(maple_ctx* c, const Phys& P, bool full, bool fp32 = false) {
  if (P.nd == 1) {
    // thermal: node form (nz = 64 cells, Q1-Laplacian element matrices) or element rows
    int tv = c->thermal_var;
    if (tv >= 2 && !(c->all_nz64 && P.laplacian)) tv = 0;
    Launch L = tv == 2 ? (full ? c->lv[0].pair : P.lx)
             : tv == 3 ? (full ? c->lv[0].node : P.ln[0])
                       : (full ? c->lv[0].fine : P.lf);
    L.var = tv;
    return L;
  }
  int v = c->lv[0].fine.var;
  if (v == 9 && !(fp32 && c->mfma_ok)) v = 8;            // FP32 on the matrix cores only
  if (v == 8 && !c->all_nz64) v = 5;                     // the lane-rotation form needs nz = 64
  const int var = v;
  Launch L = var == 9 ? (full ? c->lv[0].mfma : P.lm)
           : var == 8 ? (full ? c->lv[0].pair : P.lx)
           : var >= 6 ? (full ? c->lv[0].quad : P.lq)
           : var >= 4 ? (full ? c->lv[0].pair : P.lx)
           : var >= 2 ? (full ? c->lv[0].node : P.ln[0])
                      : (full ? c->lv[0].fine : P.lf);
  L.var = var;
  return L;
}

// FP64 level-0 operator for the CG matvec and residuals (any precision mode).
void apply_fine64(maple_ctx* c, Phys& P, int mode, const double* rho, const double* u, const double* f, double* out,
                  double* partial, hipStream_t s) {
  const int kind = P.nd == 3 ? 0 : 1;
  Launch L = op_launch(c, P, mode == 0);
  L.work = c->timing ? c->work_d + kind : nullptr;   // the launch counts its own work
  const double* dinv = P.v64.dinv.empty() ? nullptr : P.v64.dinv[0];
  timed_begin(c, kind, s);
  if (P.nd == 3 && L.var == 8)
    fine_apply_nodeb64<double>(L, mode, P.nc, P.Ke, rho, u, f, dinv, c->omega, out, partial, s);
  else if (P.nd == 3 && L.var >= 4)
    fine_apply_nodeb<double>(L, mode, P.nc, P.Ke, rho, u, f, dinv, c->omega, out, partial, s);
  else if (P.nd == 3 && L.var >= 2)
    fine_apply_node<double>(L, mode, P.nc, P.Ke, rho, u, f, dinv, c->omega, out, partial, s);
  else if (P.nd == 3)
    fine_apply<3, double>(L, mode, P.nc, P.Ke, rho, u, f, dinv, c->omega, out, partial, s);
  else if (L.var >= 2)
    thermal_apply64<double>(L, mode, P.nc, P.lap, rho, u, f, dinv, c->omega, out, partial, s, P.face0);
  else
    fine_apply<1, double>(L, mode, P.nc, P.Ke, rho, u, f, dinv, c->omega, out, partial, s);
  timed_end(c, kind, s);
}

// Operator of level l of hierarchy V (mode 2: residual, 3: Jacobi sweep).
template <typename T>
void apply_t(maple_ctx* c, Phys& P, VC<T>& V, int l, int mode, const T* u, const T* f, T* out, hipStream_t s,
             double w = -1.0) {
  const T om = (T)(w > 0.0 ? w : c->omega);
  if (l == 0) {
    const int kind = P.nd == 3 ? (sizeof(T) == 8 ? 0 : 2) : (sizeof(T) == 8 ? 1 : -1);
    Launch L = op_launch(c, P, false, sizeof(T) == 4);
    L.work = c->timing && kind >= 0 ? c->work_d + kind : nullptr;
    if (kind >= 0) timed_begin(c, kind, s);
    if constexpr (sizeof(T) == 4) {
      if (P.nd == 3 && L.var == 9) {
        fine_apply_mfma(L, mode, P.nc, V.Ke, V.rho, u, f, V.dinv[0], om, out, s);
        if (kind >= 0) timed_end(c, kind, s);
        return;
      }
    }
    if (P.nd == 3 && L.var == 8)
      fine_apply_nodeb64<T>(L, mode, P.nc, V.Ke, V.rho, u, f, V.dinv[0], om, out, nullptr, s);
    else if (P.nd == 3 && L.var >= 4)
      fine_apply_nodeb<T>(L, mode, P.nc, V.Ke, V.rho, u, f, V.dinv[0], om, out, nullptr, s);
    else if (P.nd == 3 && L.var >= 2)
      fine_apply_node<T>(L, mode, P.nc, V.Ke, V.rho, u, f, V.dinv[0], om, out, nullptr, s);
    else if (P.nd == 3)
      fine_apply<3, T>(L, mode, P.nc, V.Ke, V.rho, u, f, V.dinv[0], om, out, nullptr, s);
    else if (L.var >= 2)
      thermal_apply64<T>(L, mode, P.nc, P.lap, V.rho, u, f, V.dinv[0], om, out, nullptr, s, P.face0);
    else
      fine_apply<1, T>(L, mode, P.nc, V.Ke, V.rho, u, f, V.dinv[0], om, out, nullptr, s);
    if (kind >= 0) timed_end(c, kind, s);
  } else {
    if (P.nd == 3)
      stencil_apply<3, T>(P.ln[l], mode, P.nc, V.S[l], u, f, V.dinv[l], om, out, nullptr, s);
    else
      stencil_apply<1, T>(P.ln[l], mode, P.nc, V.S[l], u, f, V.dinv[l], om, out, nullptr, s);
  }
}

// Hierarchy of one density field: Jacobi diagonals, Galerkin stencils, dense coarsest inverse.
// force_dense: rebuild the coarsest inverse now, outside the refresh cadence.
template <typename T>
void setup_t(maple_ctx* c, Phys& P, VC<T>& V, const double* rho, hipStream_t s, bool force_dense = false) {
  const int L = c->nlevels;
  const bool l1 = c->smoother == MAPLE_SMOOTH_L1_JACOBI;
  fine_dinv<T>(c->lv[0].node, P.nd, l1, P.Ke, rho, V.dinv[0], s);
  stencil_level1<T>(c->lv[1].node, c->lv[0].node, P.nd, P.G, rho, V.S[1], s);
  for (int l = 2; l < L; ++l) stencil_galerkin<T>(c->lv[l].node, c->lv[l - 1].node, P.nd, V.S[l - 1], V.S[l], s);
  for (int l = 1; l < L - 1; ++l) stencil_dinv<T>(c->lv[l].node, P.nd, l1, V.S[l], V.dinv[l], s);
  // A lagged coarsest inverse (an earlier density's SPD inverse) keeps the V-cycle SPD;
  // CG still stops at the same residual tolerance.
  if (force_dense || P.setups % c->coarse_refresh == 0) {
    dense_assemble_g<T>(c->lv[L - 1].node, P.ddc, P.maxn, P.nd, V.S[L - 1], P.A, P.gamma, s);
    dense_inverse(P.ddc, c->ncells, P.max_nt, P.A, P.work, s);
  }
  if (!force_dense) P.setups += 1;
}

void setup_fp64(maple_ctx* c, Phys& P, const double* rho, hipStream_t s, bool force_dense) {
  P.v64.Ke = P.Ke;
  P.v64.rho = const_cast<double*>(rho);
  setup_t<double>(c, P, P.v64, rho, s, force_dense);
  P.built[0] = true;
}

void setup_fp32(maple_ctx* c, Phys& P, const double* rho, hipStream_t s, bool force_dense) {
  vec_copy<float, double>(c->lv[0].Ntot, rho, P.v32.rho, s);
  setup_t<float>(c, P, P.v32, rho, s, force_dense);
  P.built[1] = true;
}

void setup_hierarchy(maple_ctx* c, Phys& P, const double* rho, hipStream_t s) {
  if (c->prec == MAPLE_PREC_FP64) setup_fp64(c, P, rho, s, false);
  else setup_fp32(c, P, rho, s, false);
}

// z = M b at level l (symmetric V-cycle, zero start); returns the buffer holding the result.
template <typename T>
T* vcycle_t(maple_ctx* c, Phys& P, VC<T>& V, int l, const T* b, hipStream_t s) {
  const int L = c->nlevels;
  if (l == L - 1) {
    dense_apply<T>(P.ln[l], P.ddc, P.maxn, P.nd, P.nc, P.A, b, V.X[l], s);
    return V.X[l];
  }
  T* x = V.X[l];
  T* t = V.W[l];
  vec_scale_dinv<T>(P.ln[l], P.nd, P.nc, (T)c->wpre[0], V.dinv[l], b, x, s);
  for (int k = 1; k < c->nsmooth; ++k) {
    apply_t<T>(c, P, V, l, 3, x, b, t, s, c->wpre[k]);
    std::swap(x, t);
  }
  apply_t<T>(c, P, V, l, 2, x, b, V.R[l], s);
  restrict_pt<T>(P.ln[l + 1], P.ln[l], P.nd, P.nc, V.R[l], V.B[l + 1], s);
  T* e = vcycle_t<T>(c, P, V, l + 1, V.B[l + 1], s);
  prolong_add<T>(P.ln[l], P.ln[l + 1], P.nd, P.nc, e, x, s);
  for (int k = 0; k < c->npost; ++k) {
    apply_t<T>(c, P, V, l, 3, x, b, t, s, c->wpost[k]);
    std::swap(x, t);
  }
  return x;
}

// Preconditioner on the CG residual r (FP64 in, FP64 out): the FP64 V-cycle, or the
// FP32 one between conversions.
// With the island correction on, z += A_i^{-1} r on every small island (FP64).
double* precondition(maple_ctx* c, Phys& P, const double* r, bool use64, hipStream_t s) {
  double* z;
  if (use64) {
    z = vcycle_t<double>(c, P, P.v64, 0, r, s);
  } else {
    vec_copy_cases<float, double>(P.ln[0], P.nd, P.nc, r, P.v32.B[0], s);
    float* z32 = vcycle_t<float>(c, P, P.v32, 0, P.v32.B[0], s);
    vec_copy_cases<double, float>(P.ln[0], P.nd, P.nc, z32, P.z64, s);
    z = P.z64;
  }
  if (c->islands && c->nisl > 0)
    island_apply(P.nd, P.nc, c->isl, P.isl_valid, c->nisl, P.isl_inv, c->lv[0].Ntot, P.done_d, P.nc, r, z, s);
  return z;
}

// r32_ready: the FP32 copy of r is already in the V-cycle's input (the mixed update wrote it)
ZVec precondition_cg(maple_ctx* c, Phys& P, const double* r, bool use64, bool r32_ready, hipStream_t s) {
  ZVec z;
  if (use64) {
    z.d = vcycle_t<double>(c, P, P.v64, 0, r, s);
    if (c->islands && c->nisl > 0)
      island_apply(P.nd, P.nc, c->isl, P.isl_valid, c->nisl, P.isl_inv, c->lv[0].Ntot, P.done_d, P.nc, r, z.d, s);
    return z;
  }
  if (!r32_ready) vec_copy_cases<float, double>(P.ln[0], P.nd, P.nc, r, P.v32.B[0], s);
  z.f = vcycle_t<float>(c, P, P.v32, 0, P.v32.B[0], s);
  if (c->islands && c->nisl > 0)
    island_apply(P.nd, P.nc, c->isl, P.isl_valid, c->nisl, P.isl_inv, c->lv[0].Ntot, P.done_d, P.nc, r, z.f, s);
  return z;
}

void dot_rz(Phys& P, const double* r, const ZVec& z, double* partial, hipStream_t s) {
  if (z.f) vec_dot<double, float>(P.ln[0], P.nd, P.nc, r, z.f, partial, s);
  else vec_dot<double, double>(P.ln[0], P.nd, P.nc, r, z.d, partial, s);
}

// (r, z) of every live case into SC_RZ_NEW and the per-component means of z into zmean
// (the next direction is projected off the rigid translations)
void dot_rz_mean(maple_ctx* c, Phys& P, const ZVec& z, hipStream_t s) {
  double* zpart = P.partial + (size_t)c->lv[0].node.nblocks * 2 * MAPLE_MAX_CASES;
  if (z.f) pcg_dot_rz<float>(P.ln[0], P.nd, P.nc, P.r, z.f, P.partial, zpart, s);
  else pcg_dot_rz<double>(P.ln[0], P.nd, P.nc, P.r, z.d, P.partial, zpart, s);
  reduce_partials(P.ln[0], P.nc, P.partial, P.sc, SC_RZ_NEW, s);
  pcg_reduce_zmean(P.ln[0], P.nd, P.nc, zpart, P.sc, s);
}

// p = (z - zmean) + beta p
void update_p(maple_ctx* c, Phys& P, const ZVec& z, hipStream_t s) {
  (void)c;
  if (z.f && pcg_update_p32(P.ln[0], P.nd, P.nc, P.sc, z.f, P.p, s) == 0) return;
  if (z.f) {   // shapes without a fused kernel: widen first
    vec_copy_cases<double, float>(P.ln[0], P.nd, P.nc, z.f, P.z64, s);
    pcg_update_p(P.ln[0], P.nd, P.nc, P.sc, P.z64, P.p, s);
  } else {
    pcg_update_p(P.ln[0], P.nd, P.nc, P.sc, z.d, P.p, s);
  }
}

// Launch tables over every cell (start of a solve).
void full_launches(maple_ctx* c, Phys& P) {
  for (size_t l = 0; l < P.ln.size(); ++l) {
    P.ln[l] = c->lv[l].node;
    P.ln[l].done = P.done_d;
    P.ln[l].ndc = P.nc;
  }
  P.lf = c->lv[0].fine;
  P.lf.done = P.done_d;
  P.lf.ndc = P.nc;
  P.lx = c->lv[0].pair;
  P.lx.done = P.done_d;
  P.lx.ndc = P.nc;
  P.lq = c->lv[0].quad;
  P.lq.done = P.done_d;
  P.lq.ndc = P.nc;
  P.lm = c->lv[0].mfma;
  P.lm.done = P.done_d;
  P.lm.ndc = P.nc;
  P.nlive = c->ncells;
  P.epoch = 0;
}

// Once cells have converged (host flags of the last poll), the iteration's launches
// list only the cells still iterating: a converged cell's blocks would only exit at
// once, but a 16-cell batch with one live cell would still dispatch sixteen cells'
// worth of them on every launch.  The tables are uploaded on the solve's stream, so
// the kernels already queued keep reading the previous ones.
int compact_launches(maple_ctx* c, Phys& P, const int* done_h, hipStream_t s) {
  std::vector<int> live;
  for (int i = 0; i < c->ncells; ++i) {
    bool d = true;
    for (int j = 0; j < P.nc; ++j) d = d && done_h[i * P.nc + j];
    if (!d) live.push_back(i);
  }
  const int n = (int)live.size();
  if (n == 0 || n >= P.nlive) return 0;
  char* st = P.stage + (size_t)(P.stage_next++ % P.stage_slots) * P.stage_bytes;
  auto r8 = [](size_t b) { return (b + 7) / 8 * 8; };
  auto build = [&](const Level& Lv, const std::vector<int>& cnt, LevelCell* dcells, int* dblk, Launch& out) -> int {
    LevelCell* hc = (LevelCell*)st;
    st += r8(sizeof(LevelCell) * c->ncells);
    int* hb = (int*)st;
    st += r8(sizeof(int) * (c->ncells + 1));
    hb[0] = 0;
    for (int k = 0; k < n; ++k) {
      hc[k] = Lv.h[live[k]];
      hb[k + 1] = hb[k] + cnt[live[k]];
    }
    HIPCHK(hipMemcpyAsync(dcells, hc, sizeof(LevelCell) * n, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(dblk, hb, sizeof(int) * (n + 1), hipMemcpyHostToDevice, s));
    out.cells = dcells;
    out.blk_start = dblk;
    out.ncells = n;
    out.nblocks = hb[n];
    return 0;
  };
  for (size_t l = 0; l < P.ln.size(); ++l) TRY(build(c->lv[l], c->lv[l].cnt_node, P.cc[l], P.cb[l], P.ln[l]));
  TRY(build(c->lv[0], c->lv[0].cnt_fine, P.ccf, P.cbf, P.lf));
  TRY(build(c->lv[0], c->lv[0].cnt_pair, P.ccx, P.cbx, P.lx));
  TRY(build(c->lv[0], c->lv[0].cnt_quad, P.ccq, P.cbq, P.lq));
  if (c->mfma_ok) TRY(build(c->lv[0], c->lv[0].cnt_mfma, P.ccm, P.cbm, P.lm));
  P.nlive = n;
  P.epoch += 1;   // the iteration's captured launches no longer apply
  return 0;
}

int poll_done(maple_ctx* c, Phys& P, int slot, bool* all) {
  HIPCHK(hipEventSynchronize(P.poll[slot]));
  const int n = c->ncells * P.nc;
  bool a = true;
  for (int i = 0; i < n; ++i) a = a && P.done_h[slot * (n + 1) + i];
  *all = a;
  return 0;
}

int enqueue_poll(maple_ctx* c, Phys& P, int slot, hipStream_t s) {
  const int n = c->ncells * P.nc;
  HIPCHK(hipMemcpyAsync(P.done_h + slot * (n + 1), P.done_d, sizeof(int) * (n + 1), hipMemcpyDeviceToHost, s));
  HIPCHK(hipEventRecord(P.poll[slot], s));
  return 0;
}

unsigned long long graph_key(std::initializer_list<unsigned long long> parts) {
  unsigned long long h = 1469598103934665603ULL;
  for (unsigned long long v : parts) h = (h ^ v) * 1099511628211ULL + (h >> 29);
  return h;
}

// Runs one CG segment: eagerly on s, or (graphs on, timing off) as the hipGraph captured
// for this key, capturing it first on the physics' own stream.  The body records its host
// decisions in G, which a replay reuses.
template <typename Body>
int run_segment(maple_ctx* c, Phys& P, SegGraph& G, unsigned long long key, hipStream_t s, Body&& body) {
  if (!c->graphs || c->timing) {
    body(s, G);
    return 0;
  }
  if (G.key != key || !G.exec) {
    if (G.exec) (void)hipGraphExecDestroy(G.exec);
    G.exec = nullptr;
    G.key = ~0ULL;
    if (!P.cap) HIPCHK(hipStreamCreateWithFlags(&P.cap, hipStreamNonBlocking));
    HIPCHK(hipStreamBeginCapture(P.cap, hipStreamCaptureModeThreadLocal));
    body(P.cap, G);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(P.cap, &g);
    if (e != hipSuccess || !g) return fail(MAPLE_EDEVICE, std::string("CG iteration capture failed: ") + hipGetErrorString(e));
    const hipError_t e2 = hipGraphInstantiate(&G.exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e2 != hipSuccess) {
      G.exec = nullptr;
      return fail(MAPLE_EDEVICE, std::string("CG iteration graph instantiation failed: ") + hipGetErrorString(e2));
    }
    G.key = key;
    __atomic_add_fetch(&c->graph_captures, 1, __ATOMIC_RELAXED);   // the two physics' threads
  }
  HIPCHK(hipGraphLaunch(G.exec, s));
  __atomic_add_fetch(&c->graph_launches, 1, __ATOMIC_RELAXED);
  return 0;
}

// Start projected on earlier solutions: the fields of the previous solve join the ring of
// the last hist_m solutions; the basis is the ring, newest first, and a supplied start.
// Per (cell, case) the start is the combination of the basis that minimises the energy
// norm of the error (proj_start); it and its residual go to P.ustart and P.q with
// (r, r) in CaseScalars::rrs, ready for pcg_pick_start.  P.f must hold the right-hand side.
int projected_start(maple_ctx* c, Phys& P, const double* rho, hipStream_t s, bool* done) {
  *done = false;
  const Level& L0 = c->lv[0];
  const size_t n0 = (size_t)P.nc * P.nd * L0.Ntot;
  const int m = c->hist_m;
  if (P.have_fields) {
    HIPCHK(hipMemcpyAsync(P.hist + (size_t)P.hpos * n0, P.u, n0 * sizeof(double), hipMemcpyDeviceToDevice, s));
    P.hpos = (P.hpos + 1) % m;
    P.nhist = std::min(P.nhist + 1, m);
  }
  ProjBasis B{};
  int k = 0;
  for (int t = 0; t < P.nhist; ++t) B.h[k++] = P.hist + (size_t)((P.hpos - 1 - t + 2 * m) % m) * n0;
  if (P.have_start) B.h[k++] = P.ustart;
  if (k == 0) return 0;
  for (int a = 0; a < k; ++a) {
    double* g = P.khist + (size_t)a * n0;
    apply_fine64(c, P, 0, rho, B.h[a], nullptr, g, nullptr, s);
    B.g[a] = g;
  }
  if (proj_start(L0.node, P.nd, P.nc, k, B, P.f, P.gpart, P.pcoef, P.ustart, P.q, P.partial, s))
    return fail(MAPLE_EINVAL, "projected start: unsupported basis or field shape");
  reduce_partials(L0.node, P.nc, P.partial, P.sc, SC_RRS, s);
  *done = true;
  return 0;
}

int pcg(maple_ctx* c, Phys& P, const double* rho, double rtol, int maxit, hipStream_t s, bool force_prev = false) {
  Level& L0 = c->lv[0];
  const bool flex = c->flexible < 0 ? c->prec != MAPLE_PREC_FP64 : c->flexible != 0;
  bool use64 = c->prec == MAPLE_PREC_FP64;
  bool want64 = false;   // a CG restart was seen: finish on the FP64 V-cycle
  g_err.clear();
  // second partial-sum area of the solve (the mixed update's fused flexible dot); the
  // third (dot_rz_mean's sums of z) follows it
  double* partial2 = P.partial + (size_t)c->lv[0].node.nblocks * MAPLE_MAX_CASES;
  vec_zero<int>((long long)c->ncells * P.nc + 1, P.done_d, s);   // no case is converged before the start
  full_launches(c, P);
  fine_rhs(L0.node, P.nd, P.nc, P.Fe, rho, P.f, s);
  vec_dot<double, double>(L0.node, P.nd, P.nc, P.f, P.f, P.partial, s);
  reduce_partials(L0.node, P.nc, P.partial, P.sc, SC_FF, s);
  // start: per (cell, case) the smallest residual among zero, the previous fields
  // (warm start) and a supplied start (learned warm start); with a solution history, the
  // previous fields and the supplied start are basis vectors of the projected start
  // instead.  A restart from the fields of an unfinished solve (force_prev) is no new
  // solution and keeps them.
  bool projected = false;
  if (c->hist_m > 0 && !force_prev) TRY(projected_start(c, P, rho, s, &projected));
  const bool have_prev = !projected && (c->warm || force_prev) && P.have_fields;
  if (have_prev) {
    apply_fine64(c, P, 2, rho, P.u, P.f, P.r, nullptr, s);
    vec_dot<double, double>(L0.node, P.nd, P.nc, P.r, P.r, P.partial, s);
    reduce_partials(L0.node, P.nc, P.partial, P.sc, SC_RR, s);
  }
  if (P.have_start && !projected) {
    apply_fine64(c, P, 2, rho, P.ustart, P.f, P.q, nullptr, s);
    vec_dot<double, double>(L0.node, P.nd, P.nc, P.q, P.q, P.partial, s);
    reduce_partials(L0.node, P.nc, P.partial, P.sc, SC_RRS, s);
  }
  pcg_pick_start(L0.node, c->ncells, P.nd, P.nc, have_prev, P.have_start || projected, P.sc, P.u, P.r, P.ustart,
                 P.q, P.f, s, projected ? 3 : 2);
  P.have_start = false;
  vec_dot<double, double>(L0.node, P.nd, P.nc, P.r, P.r, P.partial, s);
  reduce_partials(L0.node, P.nc, P.partial, P.sc, SC_RR, s);
  ZVec z = precondition_cg(c, P, P.r, use64, false, s);
  dot_rz_mean(c, P, z, s);
  pcg_init_scalars(c->ncells, P.nc, rtol, P.atol2, P.sc, P.done_d, s);   // beta = 0: p = z - zmean
  update_p(c, P, z, s);
  TRY(enqueue_poll(c, P, 0, s));
  bool all = false;
  TRY(poll_done(c, P, 0, &all));
  // An iteration is two segments: (A) q = K p, alpha, the update of u and r, the
  // convergence check; (B) the preconditioner, beta and the next direction.  With
  // graphs on, each runs as a captured hipGraph whose key covers everything baked into
  // its launches; the convergence poll and the launch compaction stay on the host between.
  auto segA = [&](hipStream_t st, SegGraph& G) {
    apply_fine64(c, P, 1, rho, P.p, nullptr, P.q, P.partial, st);
    reduce_partials(op_launch(c, P, false), P.nc, P.partial, P.sc, SC_PQ, st);
    pcg_alpha(c->ncells, P.nc, P.sc, P.done_d, P.done_d + c->ncells * P.nc, st);
    // mixed form: the update also writes the FP32 preconditioner's input and, from
    // the previous FP32 z, the flexible dot (r_new, z_old)
    G.r32_ready = G.rzo_done = false;
    if (!use64 && pcg_update_ur_mixed(P.ln[0], P.nd, P.nc, P.sc, P.p, P.q, P.u, P.r, P.v32.B[0],
                                      flex ? z.f : nullptr, P.partial, partial2, st) == 0) {
      G.r32_ready = true;
      G.rzo_done = flex && z.f != nullptr;
    } else {
      pcg_update_ur(P.ln[0], P.nd, P.nc, P.sc, P.p, P.q, P.u, P.r, P.partial, st);
    }
    reduce_partials(P.ln[0], P.nc, P.partial, P.sc, SC_RR_LIVE, st);
    if (G.rzo_done) reduce_partials(P.ln[0], P.nc, partial2, P.sc, SC_RZO, st);
    pcg_check(c->ncells, P.nc, P.sc, P.done_d, st);
  };
  bool r32_ready = false, rzo_done = false;
  auto segB = [&](hipStream_t st, SegGraph& G) {
    if (flex && !rzo_done) {
      dot_rz(P, P.r, z, P.partial, st);
      reduce_partials(P.ln[0], P.nc, P.partial, P.sc, SC_RZO, st);
    }
    G.z = precondition_cg(c, P, P.r, use64, r32_ready && !use64, st);
    dot_rz_mean(c, P, G.z, st);
    pcg_beta(c->ncells, P.nc, flex, P.sc, st);
    update_p(c, P, G.z, st);
  };
  for (int it = 0; it < maxit && !all; ++it) {
    const int gs = P.epoch == 0 ? 0 : 1;
    const unsigned long long kbase = graph_key({P.epoch, (unsigned long long)P.nlive, (unsigned long long)use64,
                                                (unsigned long long)flex, (unsigned long long)c->nisl,
                                                (unsigned long long)(uintptr_t)rho, c->gen});
    SegGraph& GA = P.ga[gs];
    TRY(run_segment(c, P, GA, graph_key({kbase, (unsigned long long)(uintptr_t)z.f,
                                         (unsigned long long)(uintptr_t)z.d}), s, segA));
    r32_ready = GA.r32_ready;
    rzo_done = GA.rzo_done;
    const int slot = (it + 1) & 1;
    TRY(enqueue_poll(c, P, slot, s));
    if (it + 1 < maxit) {
      // An FP32 V-cycle cannot resolve the soft modes of solid islands held by 1e-9
      // voids: a mixed solve still running after `fallback` iterations finishes on
      // the FP64 hierarchy of the same density (flexible CG absorbs the switch).
      // A restart (p^T K p <= 0 on a live case) under the FP32 preconditioner is the same
      // failure seen earlier: switch at once.
      if (!use64 && c->fallback > 0 && (it + 1 == c->fallback || want64)) {
        setup_fp64(c, P, rho, s, true);
        use64 = true;
      }
      SegGraph& GB = P.gb[gs];
      TRY(run_segment(c, P, GB, graph_key({kbase, (unsigned long long)use64, (unsigned long long)r32_ready,
                                           (unsigned long long)rzo_done, (unsigned long long)(uintptr_t)z.f,
                                           (unsigned long long)(uintptr_t)z.d}), s, segB));
      z = GB.z;
    }
    // the device is already working on the next iteration while we look at this one
    TRY(poll_done(c, P, slot, &all));
    want64 = want64 || P.done_h[slot * (c->ncells * P.nc + 1) + c->ncells * P.nc] != 0;
    if (!all) TRY(compact_launches(c, P, P.done_h + slot * (c->ncells * P.nc + 1), s));
  }
  P.have_fields = true;
  HIPCHK(hipMemcpyAsync(P.hsc.data(), P.sc, sizeof(CaseScalars) * P.hsc.size(), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  for (size_t i = 0; i < P.hsc.size(); ++i) {
    const CaseScalars& cs = P.hsc[i];
    if (!(cs.rr == cs.rr) || !std::isfinite(cs.ff))
      return fail(MAPLE_ENONFINITE, "non-finite residual in the cell solve (cell " + std::to_string(i / P.nc) +
                                        ", load case " + std::to_string(i % P.nc) + ", after " +
                                        std::to_string(cs.iters) + " iterations)");
  }
  for (size_t i = 0; i < P.hsc.size(); ++i)
    if (P.hsc[i].brk == 1)
      return fail(MAPLE_EBREAKDOWN, "CG breakdown (p^T K p <= 0) in cell " + std::to_string(i / P.nc) + ", load case " +
                                        std::to_string(i % P.nc) + " after " + std::to_string(P.hsc[i].iters) +
                                        " iterations; its fields are not converged");
  for (size_t i = 0; i < P.hsc.size(); ++i)
    if (P.hsc[i].brk == 2)
      return fail(MAPLE_ENOCONV, "CG stagnated before rtol in cell " + std::to_string(i / P.nc) + ", load case " +
                                     std::to_string(i % P.nc) + " (no new residual minimum after iteration " +
                                     std::to_string(P.hsc[i].it_min) + " of " + std::to_string(P.hsc[i].iters) +
                                     "): rtol is below the accuracy this operator allows");
  return all ? 0 : MAPLE_ENOCONV;
}

// Block CG: the cases of a cell share one Krylov space (bcg.hip).
int bpcg(maple_ctx* c, Phys& P, const double* rho, double rtol, int maxit, hipStream_t s) {
  Level& L0 = c->lv[0];
  const Launch& LN = P.ln[0];
  bool use64 = c->prec == MAPLE_PREC_FP64;
  vec_zero<int>((long long)c->ncells * P.nc + 1, P.done_d, s);
  full_launches(c, P);
  fine_rhs(L0.node, P.nd, P.nc, P.Fe, rho, P.f, s);
  vec_dot<double, double>(L0.node, P.nd, P.nc, P.f, P.f, P.partial, s);
  reduce_partials(L0.node, P.nc, P.partial, P.sc, SC_FF, s);
  const bool have_prev = c->warm && P.have_fields;
  if (have_prev) {
    apply_fine64(c, P, 2, rho, P.u, P.f, P.r, nullptr, s);
    vec_dot<double, double>(L0.node, P.nd, P.nc, P.r, P.r, P.partial, s);
    reduce_partials(L0.node, P.nc, P.partial, P.sc, SC_RR, s);
  }
  if (P.have_start) {
    apply_fine64(c, P, 2, rho, P.ustart, P.f, P.q, nullptr, s);
    vec_dot<double, double>(L0.node, P.nd, P.nc, P.q, P.q, P.partial, s);
    reduce_partials(L0.node, P.nc, P.partial, P.sc, SC_RRS, s);
  }
  pcg_pick_start(L0.node, c->ncells, P.nd, P.nc, have_prev, P.have_start, P.sc, P.u, P.r, P.ustart, P.q, P.f, s);
  P.have_start = false;
  vec_dot<double, double>(L0.node, P.nd, P.nc, P.r, P.r, P.partial, s);
  reduce_partials(L0.node, P.nc, P.partial, P.sc, SC_RR, s);
  pcg_init_scalars(c->ncells, P.nc, rtol, P.atol2, P.sc, P.done_d, s);
  bcg_init(c->ncells, P.nc, P.bs, P.sc, P.done_d, s);
  double* z = precondition(c, P, P.r, use64, s);
  bcg_gram(LN, P.nd, P.nc, 1, z, z, nullptr, P.partial, P.bs, s);   // translation content of Z
  bcg_new_direction(LN, P.nd, P.nc, true, P.bs, z, P.p, P.partial, s);
  TRY(enqueue_poll(c, P, 0, s));
  bool all = false;
  TRY(poll_done(c, P, 0, &all));
  for (int it = 0; it < maxit && !all; ++it) {
    apply_fine64(c, P, 1, rho, P.p, nullptr, P.q, P.partial, s);   // Q = K P (live cells only)
    bcg_gram(LN, P.nd, P.nc, 2, P.p, P.q, P.r, P.partial, P.bs, s);
    bcg_alpha(c->ncells, P.nc, P.bs, s);
    bcg_update_xr(LN, P.nd, P.nc, P.bs, P.p, P.q, P.u, P.r, P.partial, s);
    reduce_partials(LN, P.nc, P.partial, P.sc, SC_RR, s);
    bcg_check(c->ncells, P.nc, P.bs, P.sc, P.done_d, s);
    const int slot = (it + 1) & 1;
    TRY(enqueue_poll(c, P, slot, s));
    if (it + 1 < maxit) {
      if (!use64 && c->fallback > 0 && it + 1 == c->fallback) {
        setup_fp64(c, P, rho, s, true);
        use64 = true;
      }
      z = precondition(c, P, P.r, use64, s);
      bcg_gram(LN, P.nd, P.nc, 1, z, P.q, nullptr, P.partial, P.bs, s);
      bcg_beta(c->ncells, P.nc, P.bs, s);
      bcg_new_direction(LN, P.nd, P.nc, false, P.bs, z, P.p, P.partial, s);
    }
    TRY(poll_done(c, P, slot, &all));
    if (!all) TRY(compact_launches(c, P, P.done_h + slot * (c->ncells * P.nc + 1), s));
  }
  P.have_fields = true;
  HIPCHK(hipMemcpyAsync(P.hsc.data(), P.sc, sizeof(CaseScalars) * P.hsc.size(), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  for (const CaseScalars& cs : P.hsc)
    if (!(cs.rr == cs.rr) || !std::isfinite(cs.ff)) return fail(MAPLE_ENONFINITE, "non-finite residual in the cell solve");
  return all ? 0 : MAPLE_ENOCONV;
}

// The cell solve of one physics: per-case CG, or block CG that hands over to
// per-case CG (started from the block's fields) after block_max iterations.
// Block steps solve with the Gram matrix P^T K P, whose condition grows with the
// design's contrast; far below rtol 1e-7 a block can stall where per-case CG
// still converges.
int solve_cases(maple_ctx* c, Phys& P, const double* rho, double rtol, int maxit, hipStream_t s) {
  if (c->flexible != 2) return pcg(c, P, rho, rtol, maxit, s);
  const int nb = std::min(maxit, c->block_max);
  int rc = bpcg(c, P, rho, rtol, nb, s);
  if (rc != MAPLE_ENOCONV || nb >= maxit) return rc;
  std::vector<int> it0(P.hsc.size());
  for (size_t i = 0; i < P.hsc.size(); ++i) it0[i] = P.hsc[i].iters;
  rc = pcg(c, P, rho, rtol, maxit - nb, s, true);
  for (size_t i = 0; i < P.hsc.size(); ++i) P.hsc[i].iters += it0[i];
  return rc;
}

// One physics of a homogenization on stream s: hierarchy, cell solve, C_H (or
// kappa_H) and the element energies behind the sensitivities.
int solve_physics(maple_ctx* c, int p, const double* rho, double rtol, int maxit, double* out, int* iters_host,
                  hipStream_t s) {
  Phys& P = c->ph[p];
  setup_hierarchy(c, P, rho, s);
  if (c->islands && c->nisl > 0)
    island_setup(P.nd, c->lv[0].d, c->isl, c->nisl, P.Ke, rho, P.isl_inv, P.isl_valid, s);
  HIPCHK(hipGetLastError());
  int status = solve_cases(c, P, rho, rtol, maxit, s);
  if (status == MAPLE_ENOCONV) {
    if (g_err.rfind("CG stagnated", 0) == 0) g_err = std::string(p == 0 ? "elastic: " : "thermal: ") + g_err;
    else g_err = std::string(p == 0 ? "elastic" : "thermal") + " CG reached maxit before rtol";
  } else if (status == MAPLE_EBREAKDOWN) {
    g_err = std::string(p == 0 ? "elastic: " : "thermal: ") + g_err;
  } else if (status) {
    if (status == MAPLE_ENONFINITE) g_err = std::string(p == 0 ? "elastic: " : "thermal: ") + g_err;
    return status;
  }
  element_energy(c->lv[0].node, P.nd, P.nc, P.Ke, P.X0, rho, P.u, P.Se, P.partial, s);
  if (out) reduce_energy(c->lv[0].node.blk_start, c->ncells, P.nc * (P.nc + 1) / 2, P.partial, P.nc, out, s);
  HIPCHK(hipGetLastError());
  P.have_energy = true;
  if (iters_host)
    for (int i = 0; i < c->ncells; ++i)
      for (int j = 0; j < P.nc; ++j) iters_host[i * 9 + (p == 0 ? 0 : 6) + j] = P.hsc[i * P.nc + j].iters;
  return status;
}

// Islands of this density field (both physics use the same ones).  One host
// synchronisation: the island count sizes the setup and correction launches.
int find_islands(maple_ctx* c, const double* rho, hipStream_t s) {
  c->nisl = 0;
  if (!c->islands) return 0;
  vec_zero<int>(1, c->isl_count_d, s);
  island_detect(c->lv[0].node, rho, c->isl_thr, c->isl_lab, c->isl_lab2, c->isl_cnt, c->isl_bad, c->isl_slot,
                c->isl, c->isl_count_d, c->isl_max, s);
  HIPCHK(hipMemcpyAsync(c->isl_count_h, c->isl_count_d, sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  c->nisl = std::min(*c->isl_count_h, c->isl_max);
  island_nodes(c->lv[0].d, c->isl, c->nisl, s);
  HIPCHK(hipGetLastError());
  return 0;
}

// Outcomes of a solve whose results are still written: maxit reached, CG breakdown.
bool soft_status(int rc) { return rc == MAPLE_ENOCONV || rc == MAPLE_EBREAKDOWN; }

int check_ctx(const maple_ctx* c) {
  if (!c) return fail(MAPLE_EINVAL, "null context");
  return set_device(c);
}

}  // namespace

int maple_record_error(int code, const std::string& msg) { return fail(code, msg); }

// ---------------------------------------------------------------------------

extern "C" {

const char* maple_last_error(void) { return g_err.c_str(); }

const char* maple_version(void) { return "maple_hip 0.1 gfx950"; }

int maple_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return n;
}

int maple_hex8_matrices(int r, double E0, double nu, double kappa0, double* Ke24, double* Fe24, double* X0_24,
                        double* Ke8, double* Fe8, double* X0_8) {
  if (r <= 0) return fail(MAPLE_EINVAL, "r must be positive");
  if (!(nu > -1.0 && nu < 0.5)) return fail(MAPLE_EINVAL, "nu must lie in (-1, 0.5)");
  if (Ke24 && Fe24 && X0_24) hex8_elastic(r, E0, nu, Ke24, Fe24, X0_24);
  if (Ke8 && Fe8 && X0_8) hex8_thermal(r, kappa0, Ke8, Fe8, X0_8);
  return 0;
}

int maple_ctx_create(int device, int ncells, const int* dims_host, int nlevels, int precision, maple_ctx** out) {
  if (!out) return fail(MAPLE_EINVAL, "out is NULL");
  *out = nullptr;
  if (ncells <= 0 || !dims_host) return fail(MAPLE_EINVAL, "need at least one cell");
  if (nlevels < 2 || nlevels > 12) return fail(MAPLE_EINVAL, "nlevels must be in [2, 12]");
  if (precision != MAPLE_PREC_FP64 && precision != MAPLE_PREC_MIXED) return fail(MAPLE_EINVAL, "unknown precision");
  for (int i = 0; i < 3 * ncells; ++i) {
    const int d = dims_host[i];
    if (d <= 0 || d % (1 << (nlevels - 1)) != 0)
      return fail(MAPLE_EINVAL, "every cell dimension must be a positive multiple of 2^(nlevels-1)");
  }
  for (int i = 0; i < ncells; ++i)
    if (dims_host[3 * i] != dims_host[3 * i + 1] || dims_host[3 * i] != dims_host[3 * i + 2])
      return fail(MAPLE_EINVAL, "cells must be cubic (the element side is 1/r)");
  int ndev = maple_device_count();
  if (ndev <= 0) return fail(MAPLE_EDEVICE, "no HIP device available");
  if (device < 0 || device >= ndev) return fail(MAPLE_EINVAL, "device index out of range");
  maple_ctx* c = new maple_ctx();
  c->device = device;
  c->ncells = ncells;
  c->nlevels = nlevels;
  c->prec = precision;
  c->dims.assign(dims_host, dims_host + 3 * ncells);
  c->all_nz64 = true;
  for (int i = 0; i < ncells; ++i) c->all_nz64 = c->all_nz64 && dims_host[3 * i + 2] == 64;
  set_sweep_weights(c);
  int rc = set_device(c);
  auto bail = [&](int code) {
    std::string keep = g_err;
    maple_ctx_destroy(c);
    g_err = keep;
    return code;
  };
  if (rc) return bail(rc);
  c->lv.resize(nlevels);
  for (int l = 0; l < nlevels; ++l) {
    Level& L = c->lv[l];
    L.h.resize(ncells);
    long long off = 0;
    for (int i = 0; i < ncells; ++i) {
      LevelCell lc;
      lc.nx = c->dims[3 * i] >> l;
      lc.ny = c->dims[3 * i + 1] >> l;
      lc.nz = c->dims[3 * i + 2] >> l;
      lc.id = i;
      lc.off = off;
      off += (long long)lc.nx * lc.ny * lc.nz;
      L.h[i] = lc;
    }
    L.Ntot = off;
    if ((rc = dalloc(c, &L.d, ncells))) return bail(rc);
    if (hipMemcpy(L.d, L.h.data(), sizeof(LevelCell) * ncells, hipMemcpyHostToDevice) != hipSuccess)
      return bail(fail(MAPLE_EDEVICE, "upload of the cell table failed"));
    if ((rc = make_launch(c, L, 256, false, &L.dblk, &L.node, &L.cnt_node))) return bail(rc);
    if (l == 0) {
      if ((rc = make_launch(c, L, 4, true, &L.dfblk, &L.fine, &L.cnt_fine))) return bail(rc);
      if ((rc = make_launch(c, L, 1024, false, &L.doblk, &L.oc))) return bail(rc);
      if ((rc = make_launch(c, L, 512, false, &L.dxblk, &L.pair, &L.cnt_pair))) return bail(rc);
      if ((rc = make_launch(c, L, 1024, false, &L.dqblk, &L.quad, &L.cnt_quad))) return bail(rc);
      // MFMA operator: every cell must fit it, and the kernel's 32-bit offsets must cover
      // the 18 FP32 planes of the level (6 cases x 3 components)
      c->mfma_ok = 18LL * L.Ntot < (1LL << 31);
      for (int i = 0; i < ncells; ++i) c->mfma_ok = c->mfma_ok && fine_mfma_applies(L.h[i]);
      if ((rc = make_launch(c, L, 0, false, &L.dmblk, &L.mfma, &L.cnt_mfma))) return bail(rc);
      for (int i = 0; i < ncells; ++i)
        if (fine_mfma_applies(L.h[i])) L.mfma.var_mask |= 1 << (L.h[i].nz / 16 - 1);
      // default level-0 elastic operator: 2x1 node blocks, three load cases per thread,
      // z neighbours from the neighbouring lanes when every cell has nz = 64
      L.fine.var = 8;
    } else {
      L.node.var = 2;   // default coarse elastic operator: node or row threads by launch size
    }
  }
  // partial-sum scratch: 80 doubles per block of the largest launch (block CG: two 6x6 Grams)
  int maxb = std::max(c->lv[0].node.nblocks, c->lv[0].fine.nblocks);
  if ((rc = dalloc(c, &c->partial, (size_t)maxb * 80))) return bail(rc);
  if ((rc = alloc_phys(c, c->ph[0], 3, 6))) return bail(rc);
  if ((rc = alloc_phys(c, c->ph[1], 1, 3))) return bail(rc);
  for (int p = 0; p < 2; ++p) {
    Phys& P = c->ph[p];
    const int ne = 8 * P.nd;
    P.hKe.resize((size_t)ncells * ne * ne);
    P.hFe.resize((size_t)ncells * ne * P.nc);
    P.hX0.resize((size_t)ncells * ne * P.nc);
  }
  for (int i = 0; i < ncells; ++i) {
    const int r = c->dims[3 * i];
    hex8_elastic(r, 1.0, 0.3, &c->ph[0].hKe[(size_t)i * 576], &c->ph[0].hFe[(size_t)i * 144],
                 &c->ph[0].hX0[(size_t)i * 144]);
    hex8_thermal(r, 1.0, &c->ph[1].hKe[(size_t)i * 64], &c->ph[1].hFe[(size_t)i * 24], &c->ph[1].hX0[(size_t)i * 24]);
  }
  for (int p = 0; p < 2; ++p)
    if ((rc = upload_phys_constants(c, c->ph[p]))) return bail(rc);
  if ((rc = dalloc(c, &c->ocst, ncells))) return bail(rc);
  if ((rc = dalloc(c, &c->oc_a, (size_t)c->lv[0].Ntot))) return bail(rc);
  if ((rc = dalloc(c, &c->oh_tape, (size_t)c->lv[0].Ntot))) return bail(rc);
  void* hp = nullptr;
  if (hipHostMalloc(&hp, sizeof(OCState) * ncells, hipHostMallocDefault) != hipSuccess)
    return bail(fail(MAPLE_ENOMEM, "pinned host allocation failed"));
  c->host_allocs.push_back(hp);
  c->ocst_h = (OCState*)hp;
  for (Phys& P : c->ph) {
    if ((rc = dalloc(c, &P.partial, (size_t)maxb * 80))) return bail(rc);
    for (int k = 0; k < 2; ++k)
      if (hipEventCreateWithFlags(&P.poll[k], hipEventDisableTiming) != hipSuccess)
        return bail(fail(MAPLE_EDEVICE, "event creation failed"));
  }
  if ((rc = dalloc(c, &c->work_d, 3))) return bail(rc);
  if (hipMemset(c->work_d, 0, 3 * sizeof(unsigned long long)) != hipSuccess)
    return bail(fail(MAPLE_EDEVICE, "clearing the timing counters failed"));
  if (hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&c->fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->join, hipEventDisableTiming) != hipSuccess)
    return bail(fail(MAPLE_EDEVICE, "stream or event creation failed"));
  *out = c;
  return 0;
}

int maple_ctx_destroy(maple_ctx* c) {
  if (!c) return 0;
  (void)hipSetDevice(c->device);
  (void)hipDeviceSynchronize();
  for (void* p : c->allocs) (void)hipFree(p);
  for (void* p : c->cav_allocs) (void)hipFree(p);
  for (void* p : c->hist_allocs) (void)hipFree(p);
  for (void* p : c->host_allocs) (void)hipHostFree(p);
  for (Phys& P : c->ph) {
    for (int k = 0; k < 2; ++k)
      if (P.poll[k]) (void)hipEventDestroy(P.poll[k]);
    for (SegGraph* G : {&P.ga[0], &P.ga[1], &P.gb[0], &P.gb[1]})
      if (G->exec) (void)hipGraphExecDestroy(G->exec);
    if (P.cap) (void)hipStreamDestroy(P.cap);
  }
  if (c->fork) (void)hipEventDestroy(c->fork);
  if (c->join) (void)hipEventDestroy(c->join);
  if (c->side) (void)hipStreamDestroy(c->side);
  for (Timer& t : c->timers)
    for (hipEvent_t e : t.ev) (void)hipEventDestroy(e);
  delete c;
  return 0;
}

int maple_set_element_matrices(maple_ctx* c, int cell, const double* Ke24, const double* Fe24, const double* X0_24,
                               const double* Ke8, const double* Fe8, const double* X0_8) {
  TRY(check_ctx(c));
  if (cell < 0 || cell >= c->ncells) return fail(MAPLE_EINVAL, "cell index out of range");
  const double* src[2][3] = {{Ke24, Fe24, X0_24}, {Ke8, Fe8, X0_8}};
  for (int p = 0; p < 2; ++p) {
    Phys& P = c->ph[p];
    const int ne = 8 * P.nd;
    if (!src[p][0] && !src[p][1] && !src[p][2]) continue;
    if (!src[p][0] || !src[p][1] || !src[p][2]) return fail(MAPLE_EINVAL, "give Ke, Fe and X0 together");
    std::memcpy(&P.hKe[(size_t)cell * ne * ne], src[p][0], sizeof(double) * ne * ne);
    std::memcpy(&P.hFe[(size_t)cell * ne * P.nc], src[p][1], sizeof(double) * ne * P.nc);
    std::memcpy(&P.hX0[(size_t)cell * ne * P.nc], src[p][2], sizeof(double) * ne * P.nc);
    TRY(upload_phys_constants(c, P));
  }
  return 0;
}

int maple_set_solver(maple_ctx* c, double omega, int nsmooth, int warm_start) {
  if (!c) return fail(MAPLE_EINVAL, "null context");
  c->gen += 1;   // captured CG iterations of the old setting no longer apply
  if (!(omega > 0.0 && omega < 2.0)) return fail(MAPLE_EINVAL, "omega must lie in (0, 2)");
  if (nsmooth < 1 || nsmooth > 16) return fail(MAPLE_EINVAL, "nsmooth must lie in [1, 16]");
  c->omega = omega;
  c->nsmooth = nsmooth;
  c->npost = nsmooth;
  c->warm = warm_start ? 1 : 0;
  set_sweep_weights(c);
  return 0;
}

int maple_set_variant(maple_ctx* c, int kernel, int variant) {
  if (!c) return fail(MAPLE_EINVAL, "null context");
  c->gen += 1;   // captured CG iterations of the old setting no longer apply
  if (kernel == MAPLE_KERNEL_FINE_OPERATOR) {
    if (variant < 0 || variant > 9) return fail(MAPLE_EINVAL, "unknown variant");
    c->lv[0].fine.var = variant;
    for (Phys& P : c->ph) P.lf.var = variant;
    return 0;
  }
  if (kernel == MAPLE_KERNEL_THERMAL_OPERATOR) {
    if (variant < 0 || variant > 3) return fail(MAPLE_EINVAL, "unknown variant");
    c->thermal_var = variant;
    return 0;
  }
  if (kernel == MAPLE_KERNEL_COARSE_OPERATOR) {
    if (variant < 0 || variant > 2) return fail(MAPLE_EINVAL, "unknown variant");
    for (int l = 1; l < c->nlevels; ++l) {
      c->lv[l].node.var = variant;
      for (Phys& P : c->ph) P.ln[l].var = variant;
    }
    return 0;
  }
  return fail(MAPLE_EINVAL, "unknown kernel id");
}

int maple_set_coarse_refresh(maple_ctx* c, int every) {
  if (!c) return fail(MAPLE_EINVAL, "null context");
  if (every < 1) return fail(MAPLE_EINVAL, "refresh interval must be >= 1");
  c->coarse_refresh = every;
  for (Phys& P : c->ph) P.setups = 0;
  return 0;
}

int maple_set_start_history(maple_ctx* c, int m) {
  TRY(check_ctx(c));
  if (m < 0 || m > MAPLE_PROJ_MAX - 1) return fail(MAPLE_EINVAL, "history length must lie in [0, 5]");
  TRY(set_device(c));
  HIPCHK(hipDeviceSynchronize());
  for (void* p : c->hist_allocs) (void)hipFree(p);
  c->hist_allocs.clear();
  c->hist_m = 0;
  for (Phys& P : c->ph) {
    P.hist = P.khist = P.gpart = P.pcoef = nullptr;
    P.nhist = P.hpos = 0;
  }
  if (m == 0) return 0;
  auto grab = [&](double** p, size_t n) -> int {
    TRY(dalloc(c, p, n));
    c->allocs.pop_back();          // owned by the history list, freed when it is resized
    c->hist_allocs.push_back((void*)*p);
    return 0;
  };
  const int maxb = c->lv[0].node.nblocks;
  for (Phys& P : c->ph) {
    const size_t n0 = (size_t)P.nc * P.nd * c->lv[0].Ntot;
    TRY(grab(&P.hist, (size_t)m * n0));
    TRY(grab(&P.khist, (size_t)(m + 1) * n0));
    TRY(grab(&P.gpart, (size_t)maxb * P.nc * MAPLE_PROJ_NG));
    TRY(grab(&P.pcoef, (size_t)c->ncells * P.nc * MAPLE_PROJ_MAX));
  }
  c->hist_m = m;
  return 0;
}

int maple_set_concurrency(maple_ctx* c, int on) {
  if (!c) return fail(MAPLE_EINVAL, "null context");
  c->concurrent = on != 0;
  return 0;
}

int maple_set_mixed_fallback(maple_ctx* c, int iters) {
  if (!c) return fail(MAPLE_EINVAL, "null context");
  if (iters < 0) return fail(MAPLE_EINVAL, "fallback iteration count must be >= 0");
  c->fallback = iters;
  return 0;
}

int maple_set_cg(maple_ctx* c, int flexible) {
  if (!c) return fail(MAPLE_EINVAL, "null context");
  c->gen += 1;   // captured CG iterations of the old setting no longer apply
  if (flexible < -1 || flexible > 2) return fail(MAPLE_EINVAL, "CG variant must be -1, 0, 1 or 2");
  c->flexible = flexible;
  return 0;
}

int maple_set_islands(maple_ctx* c, int on, double threshold) {
  TRY(check_ctx(c));
  c->gen += 1;
  if (!(threshold > 0.0)) return fail(MAPLE_EINVAL, "island threshold must be positive");
  c->isl_thr = threshold;
  c->islands = false;
  if (!on) return 0;
  if (c->isl_max > 0) {
    c->islands = true;
    return 0;
  }
  TRY(set_device(c));
  const long long E = c->lv[0].Ntot;
  const int maxi = 256 * c->ncells;
  TRY(dalloc(c, &c->isl_lab, (size_t)E));
  TRY(dalloc(c, &c->isl_lab2, (size_t)E));
  TRY(dalloc(c, &c->isl_cnt, (size_t)E));
  TRY(dalloc(c, &c->isl_bad, (size_t)E));
  TRY(dalloc(c, &c->isl_slot, (size_t)E));
  TRY(dalloc(c, &c->isl_count_d, 1));
  TRY(dalloc(c, &c->isl, (size_t)maxi));
  for (Phys& P : c->ph) {
    const size_t nm = (size_t)P.nd * MAPLE_ISL_MAXN;
    TRY(dalloc(c, &P.isl_inv, (size_t)maxi * nm * nm));
    TRY(dalloc(c, &P.isl_valid, (size_t)maxi));
  }
  void* hp = nullptr;
  if (hipHostMalloc(&hp, sizeof(int), hipHostMallocDefault) != hipSuccess)
    return fail(MAPLE_ENOMEM, "pinned host allocation failed");
  c->host_allocs.push_back(hp);
  c->isl_count_h = (int*)hp;
  c->isl_max = maxi;
  c->islands = true;
  return 0;
}

int maple_num_islands(const maple_ctx* c) { return c ? c->nisl : 0; }

int maple_set_graphs(maple_ctx* c, int on) {
  TRY(check_ctx(c));
  if (on && !c->rho_own) {
    TRY(set_device(c));
    TRY(dalloc(c, &c->rho_own, (size_t)c->lv[0].Ntot));
  }
  c->graphs = on ? 1 : 0;
  return 0;
}

int maple_graph_counts(const maple_ctx* c, long long* captures, long long* launches) {
  TRY(check_ctx(c));
  if (captures) *captures = __atomic_load_n(&c->graph_captures, __ATOMIC_RELAXED);
  if (launches) *launches = __atomic_load_n(&c->graph_launches, __ATOMIC_RELAXED);
  return 0;
}

int maple_set_sweeps(maple_ctx* c, int pre, int post) {
  if (!c) return fail(MAPLE_EINVAL, "null context");
  c->gen += 1;   // captured CG iterations of the old setting no longer apply
  if (pre < 1 || pre > 16 || post < 0 || post > 16) return fail(MAPLE_EINVAL, "sweeps: pre in [1, 16], post in [0, 16]");
  c->nsmooth = pre;
  c->npost = post;
  set_sweep_weights(c);
  return 0;
}

int maple_set_chebyshev(maple_ctx* c, double lo) {
  if (!c) return fail(MAPLE_EINVAL, "null context");
  if (!(lo >= 0.0 && lo < 1.0)) return fail(MAPLE_EINVAL, "Chebyshev interval: lo must lie in [0, 1)");
  if (lo > 0.0 && c->smoother != MAPLE_SMOOTH_L1_JACOBI)
    return fail(MAPLE_EINVAL, "the Chebyshev schedule needs the l1-Jacobi smoother (spectrum in (0, 1])");
  c->gen += 1;
  c->cheb_lo = lo;
  set_sweep_weights(c);
  return 0;
}

int maple_set_smoother(maple_ctx* c, int kind) {
  if (!c) return fail(MAPLE_EINVAL, "null context");
  c->gen += 1;   // captured CG iterations of the old setting no longer apply
  if (kind != MAPLE_SMOOTH_JACOBI && kind != MAPLE_SMOOTH_L1_JACOBI) return fail(MAPLE_EINVAL, "unknown smoother");
  c->smoother = kind;
  if (kind != MAPLE_SMOOTH_L1_JACOBI) c->cheb_lo = 0.0;   // its spectrum bound is the l1 one
  set_sweep_weights(c);
  return 0;
}

int maple_homogenize(maple_ctx* c, const double* rho, int physics, double rtol, int maxit, double* CH,
                     double* kappa, int* iters_host, void* stream) {
  TRY(check_ctx(c));
  if (!rho) return fail(MAPLE_EINVAL, "rho is NULL");
  if (!(physics & (MAPLE_ELASTIC | MAPLE_THERMAL))) return fail(MAPLE_EINVAL, "no physics selected");
  if (!(rtol > 0.0) || maxit < 1) return fail(MAPLE_EINVAL, "need rtol > 0 and maxit >= 1");
  hipStream_t s = (hipStream_t)stream;
  if (c->graphs) {
    // the captured CG iterations read the density at a fixed address
    HIPCHK(hipMemcpyAsync(c->rho_own, rho, sizeof(double) * c->lv[0].Ntot, hipMemcpyDeviceToDevice, s));
    rho = c->rho_own;
  }
  TRY(find_islands(c, rho, s));
  const bool both = (physics & MAPLE_ELASTIC) && (physics & MAPLE_THERMAL);
  if (!both || !c->concurrent) {
    int status = 0;
    std::string msg;
    for (int p = 0; p < 2; ++p) {
      if (!(physics & (p == 0 ? MAPLE_ELASTIC : MAPLE_THERMAL))) continue;
      const int rc = solve_physics(c, p, rho, rtol, maxit, p == 0 ? CH : kappa, iters_host, s);
      if (soft_status(rc)) {
        if (status != MAPLE_EBREAKDOWN) {
          status = rc;
          msg = g_err;
        }
      } else if (rc) {
        return rc;
      }
    }
    if (status) g_err = msg;
    return status;
  }
  // The two physics share nothing but rho: the thermal solve runs on a second
  // stream, driven by a second host thread (each CG loop polls its own flags),
  // so its small kernels fill the CUs the elastic kernels leave idle.
  HIPCHK(hipEventRecord(c->fork, s));
  HIPCHK(hipStreamWaitEvent(c->side, c->fork, 0));
  int rc_th = 0;
  std::string err_th;
  std::thread th([&] {
    if (hipSetDevice(c->device) != hipSuccess) {
      rc_th = MAPLE_EDEVICE;
      err_th = "hipSetDevice failed";
      return;
    }
    rc_th = solve_physics(c, 1, rho, rtol, maxit, kappa, iters_host, c->side);
    if (rc_th) err_th = g_err;
  });
  const int rc_el = solve_physics(c, 0, rho, rtol, maxit, CH, iters_host, s);
  const std::string err_el = g_err;
  th.join();
  HIPCHK(hipEventRecord(c->join, c->side));
  HIPCHK(hipStreamWaitEvent(s, c->join, 0));
  if (rc_el && !soft_status(rc_el)) return fail(rc_el, err_el);
  if (rc_th && !soft_status(rc_th)) return fail(rc_th, err_th);
  // results are written for both soft outcomes; a breakdown outranks maxit
  if (rc_el == MAPLE_EBREAKDOWN || (rc_el && rc_th != MAPLE_EBREAKDOWN)) return fail(rc_el, err_el);
  if (rc_th) return fail(rc_th, err_th);
  return 0;
}

int maple_sensitivity(maple_ctx* c, const double* G6, const double* G3, double* out, void* stream) {
  TRY(check_ctx(c));
  if (!out) return fail(MAPLE_EINVAL, "output is NULL");
  if (G6 && !c->ph[0].have_energy) return fail(MAPLE_EINVAL, "no elastic solve to differentiate");
  if (G3 && !c->ph[1].have_energy) return fail(MAPLE_EINVAL, "no thermal solve to differentiate");
  if (!G6 && !G3) return fail(MAPLE_EINVAL, "give G6 and/or G3");
  sensitivity(c->lv[0].node, G6 ? 6 : 0, c->ph[0].Se, G6, G3 ? 3 : 0, c->ph[1].Se, G3, out, (hipStream_t)stream);
  HIPCHK(hipGetLastError());
  return 0;
}

int maple_filter_box(maple_ctx* c, const double* in, double* out, int k, void* stream) {
  TRY(check_ctx(c));
  if (k <= 0 || k % 2 == 0) return fail(MAPLE_EINVAL, "filter size must be positive and odd");
  if (!in || !out || in == out) return fail(MAPLE_EINVAL, "need distinct input and output");
  box_filter(c->lv[0].node, k, in, out, (hipStream_t)stream);
  HIPCHK(hipGetLastError());
  return 0;
}

int maple_symmetrize(maple_ctx* c, int kind, int axes, const double* in, double* out, void* stream) {
  TRY(check_ctx(c));
  if (!in || !out || in == out) return fail(MAPLE_EINVAL, "need distinct input and output");
  if (kind < 0 || kind > 2) return fail(MAPLE_EINVAL, "symmetry kind must be 0 (none), 1 (mirror) or 2 (cubic)");
  if (kind == 1 && (axes < 0 || axes > 7)) return fail(MAPLE_EINVAL, "mirror axes must be a bit mask in [0, 7]");
  if (kind == 2)
    for (int i = 0; i < c->ncells; ++i)
      if (c->dims[3 * i] != c->dims[3 * i + 1] || c->dims[3 * i] != c->dims[3 * i + 2])
        return fail(MAPLE_EINVAL, "cubic symmetry needs cubic cells");
  hipStream_t s = (hipStream_t)stream;
  if (kind == 0 || (kind == 1 && axes == 0)) {
    HIPCHK(hipMemcpyAsync(out, in, sizeof(double) * c->lv[0].Ntot, hipMemcpyDeviceToDevice, s));
    return 0;
  }
  symmetrize(c->lv[0].node, kind, axes, in, out, s);
  HIPCHK(hipGetLastError());
  return 0;
}

int maple_oc_update(maple_ctx* c, const double* x, const double* dJ, const double* volfrac, const double* move,
                    const double* xmin, int nbisect, double* xnew, void* stream) {
  TRY(check_ctx(c));
  if (!x || !dJ || !xnew || !volfrac || !move || !xmin) return fail(MAPLE_EINVAL, "null argument");
  if (nbisect < 1 || nbisect > 200) return fail(MAPLE_EINVAL, "nbisect must lie in [1, 200]");
  hipStream_t s = (hipStream_t)stream;
  HIPCHK(hipStreamSynchronize(s));   // the pinned staging buffer may still feed a previous call
  for (int i = 0; i < c->ncells; ++i) {
    c->ocst_h[i].volfrac = volfrac[i];
    c->ocst_h[i].move = move[i];
    c->ocst_h[i].xmin = xmin[i];
  }
  HIPCHK(hipMemcpyAsync(c->ocst, c->ocst_h, sizeof(OCState) * c->ncells, hipMemcpyHostToDevice, s));
  Level& L0 = c->lv[0];
  oc_prepare(L0.node, x, dJ, c->oc_a, s);
  oc_start(c->ncells, c->ocst, s);
  for (int done = 0; done < nbisect; done += 6) {
    const int depth = std::min(6, nbisect - done);
    oc_pass(L0.oc, c->ocst, x, c->oc_a, c->partial, s);
    oc_walk(L0.oc.blk_start, L0.d, c->ncells, depth, c->partial, c->ocst, s);
  }
  oc_final(L0.node, c->ocst, x, c->oc_a, xnew, s);
  HIPCHK(hipGetLastError());
  return 0;
}

int maple_rhs(maple_ctx* c, int physics, const double* rho, double* F, void* stream) {
  TRY(check_ctx(c));
  if (physics != MAPLE_ELASTIC && physics != MAPLE_THERMAL) return fail(MAPLE_EINVAL, "physics must be one of ELASTIC, THERMAL");
  if (!rho || !F) return fail(MAPLE_EINVAL, "null argument");
  Phys& P = c->ph[physics == MAPLE_ELASTIC ? 0 : 1];
  fine_rhs(c->lv[0].node, P.nd, P.nc, P.Fe, rho, F, (hipStream_t)stream);
  HIPCHK(hipGetLastError());
  return 0;
}

int maple_apply_operator(maple_ctx* c, int physics, const double* rho, int ncase, const double* u, double* v,
                         void* stream) {
  TRY(check_ctx(c));
  if (physics != MAPLE_ELASTIC && physics != MAPLE_THERMAL) return fail(MAPLE_EINVAL, "physics must be one of ELASTIC, THERMAL");
  if (!rho || !u || !v || u == v) return fail(MAPLE_EINVAL, "need rho and distinct u, v");
  Phys& P = c->ph[physics == MAPLE_ELASTIC ? 0 : 1];
  if (ncase < 1 || ncase > P.nc) return fail(MAPLE_EINVAL, "ncase out of range");
  hipStream_t s = (hipStream_t)stream;
  const int keep = P.nc;
  P.nc = ncase;
  apply_fine64(c, P, 0, rho, u, nullptr, v, nullptr, s);
  P.nc = keep;
  HIPCHK(hipGetLastError());
  return 0;
}

int maple_get_fields(maple_ctx* c, int physics, double* u, void* stream) {
  TRY(check_ctx(c));
  if (physics != MAPLE_ELASTIC && physics != MAPLE_THERMAL) return fail(MAPLE_EINVAL, "physics must be one of ELASTIC, THERMAL");
  Phys& P = c->ph[physics == MAPLE_ELASTIC ? 0 : 1];
  if (!u) return fail(MAPLE_EINVAL, "null output");
  HIPCHK(hipMemcpyAsync(u, P.u, sizeof(double) * P.nc * P.nd * c->lv[0].Ntot, hipMemcpyDeviceToDevice,
                        (hipStream_t)stream));
  return 0;
}

int maple_set_fields(maple_ctx* c, int physics, const double* u, void* stream) {
  TRY(check_ctx(c));
  if (physics != MAPLE_ELASTIC && physics != MAPLE_THERMAL) return fail(MAPLE_EINVAL, "physics must be one of ELASTIC, THERMAL");
  Phys& P = c->ph[physics == MAPLE_ELASTIC ? 0 : 1];
  if (!u) return fail(MAPLE_EINVAL, "null input");
  HIPCHK(hipMemcpyAsync(P.u, u, sizeof(double) * P.nc * P.nd * c->lv[0].Ntot, hipMemcpyDeviceToDevice,
                        (hipStream_t)stream));
  P.have_fields = true;
  return 0;
}

int maple_set_start(maple_ctx* c, int physics, const double* u, void* stream) {
  TRY(check_ctx(c));
  if (physics != MAPLE_ELASTIC && physics != MAPLE_THERMAL) return fail(MAPLE_EINVAL, "physics must be one of ELASTIC, THERMAL");
  Phys& P = c->ph[physics == MAPLE_ELASTIC ? 0 : 1];
  if (!u) return fail(MAPLE_EINVAL, "null input");
  HIPCHK(hipMemcpyAsync(P.ustart, u, sizeof(double) * P.nc * P.nd * c->lv[0].Ntot, hipMemcpyDeviceToDevice,
                        (hipStream_t)stream));
  P.have_start = true;
  return 0;
}

int maple_start_sources(maple_ctx* c, int physics, int* src_host) {
  if (!c || !src_host) return fail(MAPLE_EINVAL, "null argument");
  if (physics != MAPLE_ELASTIC && physics != MAPLE_THERMAL) return fail(MAPLE_EINVAL, "physics must be one of ELASTIC, THERMAL");
  const Phys& P = c->ph[physics == MAPLE_ELASTIC ? 0 : 1];
  for (size_t i = 0; i < P.hsc.size(); ++i) src_host[i] = P.hsc[i].src;
  return 0;
}

int maple_prolong(maple_ctx* c, int physics, int level, const double* coarse, double* fine, void* stream) {
  TRY(check_ctx(c));
  if (physics != MAPLE_ELASTIC && physics != MAPLE_THERMAL) return fail(MAPLE_EINVAL, "physics must be one of ELASTIC, THERMAL");
  if (level < 1 || level >= c->nlevels) return fail(MAPLE_EINVAL, "level must be in [1, nlevels)");
  if (!coarse || !fine) return fail(MAPLE_EINVAL, "null input");
  const Phys& P = c->ph[physics == MAPLE_ELASTIC ? 0 : 1];
  hipStream_t s = (hipStream_t)stream;
  vec_zero<double>((long long)P.nc * P.nd * c->lv[level - 1].Ntot, fine, s);
  prolong_add<double>(c->lv[level - 1].node, c->lv[level].node, P.nd, P.nc, coarse, fine, s);
  HIPCHK(hipGetLastError());
  return 0;
}

long long maple_num_nodes(const maple_ctx* c, int level) {
  if (!c || level < 0 || level >= c->nlevels) return 0;
  return c->lv[level].Ntot;
}

long long maple_num_elements(const maple_ctx* c) { return c ? c->lv[0].Ntot : 0; }

int maple_timing_enable(maple_ctx* c, int enable) {
  if (!c) return fail(MAPLE_EINVAL, "null context");
  c->timing = enable != 0;
  return 0;
}

int maple_timing_read(maple_ctx* c, int kind, long long* count, double* total_ms, long long* work) {
  TRY(check_ctx(c));
  if (kind < 0 || kind > 2 || !count || !total_ms) return fail(MAPLE_EINVAL, "bad timing query");
  TRY(set_device(c));
  Timer& t = c->timers[kind];
  double ms = 0.0;
  if (t.used > 0) HIPCHK(hipEventSynchronize(t.ev[t.used - 1]));
  for (size_t k = 0; k + 1 < t.used; k += 2) {
    float m = 0.0f;
    HIPCHK(hipEventElapsedTime(&m, t.ev[k], t.ev[k + 1]));
    ms += m;
  }
  *count = (long long)(t.used / 2);
  *total_ms = ms;
  t.used = 0;
  HIPCHK(hipDeviceSynchronize());
  unsigned long long w = 0;
  HIPCHK(hipMemcpy(&w, c->work_d + kind, sizeof(w), hipMemcpyDeviceToHost));
  HIPCHK(hipMemset(c->work_d + kind, 0, sizeof(w)));
  if (work) *work = (long long)w;
  return 0;
}

}  // extern "C"

// ---- multigrid internals ----------------------------------------------------

namespace {

int mg_args(maple_ctx* c, int physics, int prec, Phys** P) {
  TRY(check_ctx(c));
  if (physics != MAPLE_ELASTIC && physics != MAPLE_THERMAL) return fail(MAPLE_EINVAL, "physics must be one of ELASTIC, THERMAL");
  if (prec != MAPLE_PREC_FP64 && prec != MAPLE_PREC_MIXED) return fail(MAPLE_EINVAL, "unknown precision");
  if (prec == MAPLE_PREC_MIXED && c->prec != MAPLE_PREC_MIXED)
    return fail(MAPLE_EINVAL, "the FP32 hierarchy exists in MAPLE_PREC_MIXED contexts only");
  *P = &c->ph[physics == MAPLE_ELASTIC ? 0 : 1];
  return 0;
}

int mg_built(const Phys& P, int prec) {
  if (!P.built[prec == MAPLE_PREC_FP64 ? 0 : 1]) return fail(MAPLE_EINVAL, "no hierarchy of this precision yet (maple_mg_setup)");
  return 0;
}

// Launch tables over every cell and case (the last solve's convergence flags cleared).
void mg_full(maple_ctx* c, Phys& P, hipStream_t s) {
  vec_zero<int>((long long)c->ncells * P.nc, P.done_d, s);
  full_launches(c, P);
}

template <typename T>
int mg_apply_t(maple_ctx* c, Phys& P, VC<T>& V, int l, int mode, const double* u, const double* f, double* out,
               hipStream_t s) {
  const long long n = (long long)P.nc * P.nd * c->lv[l].Ntot;
  if constexpr (std::is_same<T, double>::value) {
    apply_t<double>(c, P, V, l, mode, u, f, out, s);
  } else {
    // FP32 staging in the (unused between solves) FP64 CG vectors, which hold twice as many floats
    float* u32 = (float*)P.p;
    float* f32 = (float*)P.q;
    float* o32 = (float*)P.r;
    vec_copy<float, double>(n, u, u32, s);
    if (f) vec_copy<float, double>(n, f, f32, s);
    apply_t<float>(c, P, V, l, mode, u32, f ? f32 : nullptr, o32, s);
    vec_copy<double, float>(n, o32, out, s);
  }
  HIPCHK(hipGetLastError());
  return 0;
}

}  // namespace

extern "C" {

int maple_mg_setup(maple_ctx* c, int physics, const double* rho, int prec, void* stream) {
  Phys* P;
  TRY(mg_args(c, physics, prec, &P));
  if (!rho) return fail(MAPLE_EINVAL, "rho is NULL");
  hipStream_t s = (hipStream_t)stream;
  if (prec == MAPLE_PREC_FP64) setup_fp64(c, *P, rho, s, true);
  else setup_fp32(c, *P, rho, s, true);
  HIPCHK(hipGetLastError());
  return 0;
}

int maple_mg_stencil(maple_ctx* c, int physics, int prec, int level, double* S, void* stream) {
  Phys* P;
  TRY(mg_args(c, physics, prec, &P));
  TRY(mg_built(*P, prec));
  if (level < 1 || level >= c->nlevels) return fail(MAPLE_EINVAL, "stencils exist on levels 1 .. nlevels-1");
  if (!S) return fail(MAPLE_EINVAL, "null output");
  hipStream_t s = (hipStream_t)stream;
  // all 27 blocks, the lower ones from the stored half's transposes
  if (prec == MAPLE_PREC_FP64) stencil_expand<double>(c->lv[level].node, P->nd, P->v64.S[level], S, s);
  else stencil_expand<float>(c->lv[level].node, P->nd, P->v32.S[level], S, s);
  HIPCHK(hipGetLastError());
  return 0;
}

int maple_mg_dinv(maple_ctx* c, int physics, int prec, int level, double* dinv, void* stream) {
  Phys* P;
  TRY(mg_args(c, physics, prec, &P));
  TRY(mg_built(*P, prec));
  if (level < 0 || level >= c->nlevels - 1) return fail(MAPLE_EINVAL, "smoother diagonals exist on levels 0 .. nlevels-2");
  if (!dinv) return fail(MAPLE_EINVAL, "null output");
  const long long n = (long long)P->nd * c->lv[level].Ntot;
  hipStream_t s = (hipStream_t)stream;
  if (prec == MAPLE_PREC_FP64) vec_copy<double, double>(n, P->v64.dinv[level], dinv, s);
  else vec_copy<double, float>(n, P->v32.dinv[level], dinv, s);
  HIPCHK(hipGetLastError());
  return 0;
}

int maple_mg_apply(maple_ctx* c, int physics, int prec, int level, int mode, const double* u, const double* f,
                   double* out, void* stream) {
  Phys* P;
  TRY(mg_args(c, physics, prec, &P));
  TRY(mg_built(*P, prec));
  if (level < 0 || level >= c->nlevels) return fail(MAPLE_EINVAL, "level out of range");
  if (mode != 0 && mode != 2 && mode != 3) return fail(MAPLE_EINVAL, "mode must be 0 (A u), 2 (f - A u) or 3 (Jacobi sweep)");
  if (mode == 3 && level == c->nlevels - 1) return fail(MAPLE_EINVAL, "the coarsest level has no smoother");
  if (!u || !out || (mode != 0 && !f)) return fail(MAPLE_EINVAL, "null argument");
  if (u == out || (f && f == out)) return fail(MAPLE_EINVAL, "out may not alias u or f");
  hipStream_t s = (hipStream_t)stream;
  mg_full(c, *P, s);
  if (prec == MAPLE_PREC_FP64) return mg_apply_t<double>(c, *P, P->v64, level, mode, u, mode ? f : nullptr, out, s);
  return mg_apply_t<float>(c, *P, P->v32, level, mode, u, mode ? f : nullptr, out, s);
}

int maple_mg_restrict(maple_ctx* c, int physics, int level, const double* fine, double* coarse, void* stream) {
  Phys* P;
  TRY(mg_args(c, physics, MAPLE_PREC_FP64, &P));
  if (level < 1 || level >= c->nlevels) return fail(MAPLE_EINVAL, "level must be in [1, nlevels)");
  if (!fine || !coarse) return fail(MAPLE_EINVAL, "null argument");
  hipStream_t s = (hipStream_t)stream;
  mg_full(c, *P, s);
  restrict_pt<double>(P->ln[level], P->ln[level - 1], P->nd, P->nc, fine, coarse, s);
  HIPCHK(hipGetLastError());
  return 0;
}

int maple_mg_vcycle(maple_ctx* c, int physics, int prec, const double* r, double* z, void* stream) {
  Phys* P;
  TRY(mg_args(c, physics, prec, &P));
  TRY(mg_built(*P, prec));
  if (!r || !z || r == z) return fail(MAPLE_EINVAL, "need distinct r and z");
  hipStream_t s = (hipStream_t)stream;
  mg_full(c, *P, s);
  const long long n = (long long)P->nc * P->nd * c->lv[0].Ntot;
  if (prec == MAPLE_PREC_FP64) {
    double* zz = vcycle_t<double>(c, *P, P->v64, 0, r, s);
    vec_copy<double, double>(n, zz, z, s);
  } else {
    vec_copy<float, double>(n, r, P->v32.B[0], s);
    float* zz = vcycle_t<float>(c, *P, P->v32, 0, P->v32.B[0], s);
    vec_copy<double, float>(n, zz, z, s);
  }
  HIPCHK(hipGetLastError());
  return 0;
}

int maple_mg_coarsest_inverse(maple_ctx* c, int physics, int cell, double* Ainv, void* stream) {
  Phys* P;
  TRY(mg_args(c, physics, MAPLE_PREC_FP64, &P));
  if (!P->built[0] && !P->built[1]) return fail(MAPLE_EINVAL, "no hierarchy yet (maple_mg_setup)");
  if (cell < 0 || cell >= c->ncells) return fail(MAPLE_EINVAL, "cell index out of range");
  if (!Ainv) return fail(MAPLE_EINVAL, "null output");
  const DenseCell& dc = P->hdc[cell];
  HIPCHK(hipMemcpyAsync(Ainv, P->A + dc.aoff, sizeof(double) * dc.n * dc.n, hipMemcpyDeviceToDevice,
                        (hipStream_t)stream));
  return 0;
}

int maple_mg_coarsest_order(const maple_ctx* c, int physics, int cell) {
  if (!c || cell < 0 || cell >= c->ncells || (physics != MAPLE_ELASTIC && physics != MAPLE_THERMAL)) return 0;
  return c->ph[physics == MAPLE_ELASTIC ? 0 : 1].hdc[cell].n;
}

}  // extern "C"

// ---- additive-manufacturing penalties ---------------------------------------------

extern "C" {

int maple_penalty_overhang(maple_ctx* c, const double* rho, double deg, int periodic_xy, double* P, double* dP,
                           void* stream) {
  TRY(check_ctx(c));
  if (!rho || !P) return fail(MAPLE_EINVAL, "null argument");
  const double od = std::min(89.0, std::max(1.0, deg));
  const int rad = std::max(1, (int)std::ceil(std::tan((90.0 - od) * M_PI / 180.0)));
  if (rad > 7) return fail(MAPLE_EINVAL, "overhang angle too shallow: support radius above 7 voxels");
  int rmax = 0;
  for (int i = 0; i < c->ncells; ++i) rmax = std::max(rmax, c->dims[3 * i]);
  if (rmax > 64) return fail(MAPLE_EINVAL, "the overhang penalty takes cells up to 64^3");
  if (periodic_xy && rmax > 0)
    for (int i = 0; i < c->ncells; ++i)
      if (c->dims[3 * i] < 2 * rad + 1) return fail(MAPLE_EINVAL, "cell smaller than the support window");
  overhang_penalty(c->lv[0].d, c->ncells, rmax, rho, rad, periodic_xy ? 1 : 0, c->oh_tape, P, dP, (hipStream_t)stream);
  HIPCHK(hipGetLastError());
  return 0;
}

int maple_set_cavity(maple_ctx* c, const maple_cavity_opts* o) {
  TRY(check_ctx(c));
  if (!o) return fail(MAPLE_EINVAL, "null options");
  if (o->min_channel_rad < 0 || o->min_channel_rad > 3) return fail(MAPLE_EINVAL, "min_channel_rad must lie in [0, 3]");
  if (o->rad < 0 || o->rad > 1) return fail(MAPLE_EINVAL, "the flood-fill radius must be 0 or 1");
  if (o->iters < 0 || o->iters > 1024) return fail(MAPLE_EINVAL, "iters must lie in [0, 1024]");
  if (o->seed_faces < MAPLE_SEED_ALL || o->seed_faces > MAPLE_SEED_XY) return fail(MAPLE_EINVAL, "unknown seed faces");
  if (!(o->beta > 0.0)) return fail(MAPLE_EINVAL, "beta must be positive");
  int t[3];
  for (int d = 0; d < 3; ++d) {
    t[d] = std::max(1, o->tiles[d]);
    if (t[d] > 8) return fail(MAPLE_EINVAL, "at most 8 tiles per axis");
  }
  const bool array = t[0] > 1 || t[1] > 1 || t[2] > 1;
  for (void* p : c->cav_allocs) (void)hipFree(p);
  c->cav_allocs.clear();
  c->cav_ready = false;
  CavityBuffers& B = c->cav;
  B = CavityBuffers();
  B.beta = o->beta;
  B.thr = o->void_thresh;
  B.mcr = o->min_channel_rad;
  B.rad = o->rad;
  B.periodic = array ? 0 : (o->periodic ? 1 : 0);
  B.faces = o->seed_faces;
  B.focus = array && o->focus_center && t[0] >= 3 && t[1] >= 3 && t[2] >= 3;
  std::vector<LevelCell> sc(c->ncells);
  std::vector<int> blk(c->ncells + 1, 0), reg(3 * c->ncells);
  long long off = 0;
  int side = 0;
  for (int i = 0; i < c->ncells; ++i) {
    LevelCell s;
    s.nx = t[0] * c->dims[3 * i];
    s.ny = t[1] * c->dims[3 * i + 1];
    s.nz = t[2] * c->dims[3 * i + 2];
    s.id = i;
    s.off = off;
    sc[i] = s;
    const long long n = (long long)s.nx * s.ny * s.nz;
    off += n;
    blk[i + 1] = blk[i] + (int)((n + 255) / 256);
    for (int d = 0; d < 3; ++d) reg[3 * i + d] = c->dims[3 * i + d];
    side = std::max(side, std::max(s.nx, std::max(s.ny, s.nz)));
    if (B.periodic && std::min(s.nx, std::min(s.ny, s.nz)) < 2 * std::max(B.mcr, B.rad) + 1)
      return fail(MAPLE_EINVAL, "cell smaller than the pooling window");
  }
  B.nsteps = o->iters > 0 ? o->iters : side;
  const long long G = off;
  auto al = [&](auto** p, size_t n) -> int {
    using T = typename std::remove_pointer<typename std::remove_reference<decltype(*p)>::type>::type;
    *p = nullptr;
    if (hipMalloc((void**)p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) {
      (void)hipGetLastError();
      return fail(MAPLE_ENOMEM, "hipMalloc of the cavity buffers failed (" + std::to_string(n * sizeof(T)) + " bytes)");
    }
    c->cav_allocs.push_back((void*)*p);
    return 0;
  };
  LevelCell* dsc = nullptr;
  int* dblk = nullptr;
  TRY(al(&dsc, c->ncells));
  TRY(al(&dblk, c->ncells + 1));
  TRY(al(&B.region, 3 * c->ncells));
  HIPCHK(hipMemcpy(dsc, sc.data(), sizeof(LevelCell) * sc.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dblk, blk.data(), sizeof(int) * blk.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(B.region, reg.data(), sizeof(int) * reg.size(), hipMemcpyHostToDevice));
  B.L.cells = dsc;
  B.L.blk_start = dblk;
  B.L.ncells = c->ncells;
  B.L.nblocks = blk.back();
  B.L.Ntot = G;
  for (double** p : {&B.voidv, &B.reach[0], &B.reach[1], &B.g[0], &B.g[1], &B.gvc, &B.gvoid}) TRY(al(p, (size_t)G));
  if (B.mcr > 0) {
    TRY(al(&B.vc, (size_t)G));
    TRY(al(&B.er, (size_t)G));
    TRY(al(&B.argE, (size_t)G));
    TRY(al(&B.argD, (size_t)G));
  }
  TRY(al(&B.partial, (size_t)B.L.nblocks));
  if (B.rad > 0) TRY(al(&B.tape, (size_t)G * B.nsteps));
  c->cav_ready = true;
  return 0;
}

int maple_penalty_cavity(maple_ctx* c, const double* rho, double* P, double* dP, void* stream) {
  TRY(check_ctx(c));
  if (!c->cav_ready) return fail(MAPLE_EINVAL, "configure the cavity penalty first (maple_set_cavity)");
  if (!rho || !P) return fail(MAPLE_EINVAL, "null argument");
  TRY(cavity_penalty(c->cav, c->lv[0].d, c->lv[0].node, rho, P, dP, (hipStream_t)stream));
  HIPCHK(hipGetLastError());
  return 0;
}

}  // extern "C"
