// Block preconditioned CG: the NC load cases of a cell share one Krylov space.
//
// The cases of one cell solve with the same matrix K(rho) and the same
// preconditioner, and the slow modes that hold CG back (soft members and
// islands of the design) are common to all of them.  Block CG searches the
// span of all NC preconditioned residuals at every step, so an outlying
// eigenvalue costs one step for the whole block instead of one per case.
// Per cell and iteration (P: NC search directions, orthonormal):
//   Q = K P;  alpha = (P^T Q)^-1 P^T R;  X += P alpha;  R -= Q alpha
//   Z = M R;  beta = -(P^T Q)^-1 Q^T Z;  V = Z + P beta;  P = V T
// with T from a Cholesky factorisation of the column-scaled V^T V that drops
// numerically dependent columns (breakdown-free block CG).  beta makes the new
// block K-conjugate to the last one whatever M is, so a preconditioner that is
// not exactly symmetric (the FP32 V-cycle) is allowed.  Small per-cell
// matrices live on the device; Gram matrices are reduced deterministically
// (block partials, then a fixed-order sum per cell).
#include "kernels.h"

namespace {

constexpr int KG = 36;   // one NC x NC Gram matrix (NC <= 6), row-major

// ci: the id of the block's cell (per-cell arrays); gid: the thread's node
__device__ __forceinline__ bool bnode(const Launch& L, int& ci, long long& gid, int* kc = nullptr) {
  const int k = find_cell(L.blk_start, L.ncells, lbid());
  const LevelCell c = L.cells[k];
  ci = c.id;
  if (kc) *kc = k;
  const long long n = (long long)(lbid() - L.blk_start[k]) * 256 + threadIdx.x;
  if (n >= (long long)c.nx * c.ny * c.nz) return false;
  gid = c.off + n;
  return true;
}

// partial[blk][g*NC*NC + i*NC + j] = sum over the block's nodes of A_i . B_g_j;
// with SUMS also partial[blk][NG*NC*NC + i*ND + c] = sum of component c of A_i
// (the translation content of A).
template <int NC, int ND, int NG, bool SUMS>
__global__ __launch_bounds__(256) void k_gram(Launch L, const double* __restrict__ A, const double* __restrict__ B0,
                                              const double* __restrict__ B1, double* __restrict__ partial) {
  constexpr int NV = NG * NC * NC + (SUMS ? NC * ND : 0);
  int ci;
  long long gid = 0;
  const bool ok = bnode(L, ci, gid);
  double acc[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) acc[k] = 0.0;
  if (ok && !cell_done(L.done, L.ndc, ci)) {
#pragma unroll
    for (int c = 0; c < ND; ++c) {
      double a[NC], b[NC];
#pragma unroll
      for (int i = 0; i < NC; ++i) a[i] = A[(long long)(i * ND + c) * L.Ntot + gid];
      if constexpr (SUMS) {
#pragma unroll
        for (int i = 0; i < NC; ++i) acc[NG * NC * NC + i * ND + c] = a[i];
      }
#pragma unroll
      for (int g = 0; g < NG; ++g) {
        const double* B = g == 0 ? B0 : B1;
#pragma unroll
        for (int j = 0; j < NC; ++j) b[j] = B[(long long)(j * ND + c) * L.Ntot + gid];
#pragma unroll
        for (int i = 0; i < NC; ++i)
#pragma unroll
          for (int j = 0; j < NC; ++j) acc[(g * NC + i) * NC + j] = fma(a[i], b[j], acc[(g * NC + i) * NC + j]);
      }
    }
  }
  block_sum<NV>(acc, partial + (long long)lbid() * (2 * KG));
}

// Reduce the k-th partial of every cell over its blocks: Gram entries (k < ngram*per)
// go to gram[(k/per)*36 + k%per], the component sums after them to gram[36 + ...].
__global__ __launch_bounds__(64) void k_gram_reduce(const LevelCell* __restrict__ cells,
                                                    const int* __restrict__ blk_start, const double* __restrict__ partial,
                                                    int per, int ngram, double* __restrict__ out, int stride,
                                                    int offset) {
  const int kcell = blockIdx.x, k = blockIdx.y, ci = cells[kcell].id;
  const int b0 = blk_start[kcell], b1 = blk_start[kcell + 1];
  double v = 0.0;
  for (int b = b0 + (int)threadIdx.x; b < b1; b += 64) v += partial[(long long)b * (2 * KG) + k];
  v = wave_sum(v);
  const int dst = k < ngram * per ? (k / per) * KG + k % per : KG + (k - ngram * per);
  if (threadIdx.x == 0) out[(long long)ci * stride + offset + dst] = v;
}

// Cholesky of the active part of a symmetric NC x NC matrix (inactive rows and
// columns become identity); lower factor in Lf.  Returns false on a nonpositive pivot.
template <int NC>
__device__ bool chol(const double* S, const int* act, double* Lf) {
  for (int i = 0; i < NC; ++i)
    for (int j = 0; j < NC; ++j) Lf[i * NC + j] = 0.0;
  bool ok = true;
  for (int j = 0; j < NC; ++j) {
    if (!act[j]) {
      Lf[j * NC + j] = 1.0;
      continue;
    }
    double d = 0.5 * (S[j * NC + j] + S[j * NC + j]);
    for (int k = 0; k < j; ++k) d -= Lf[j * NC + k] * Lf[j * NC + k];
    if (!(d > 0.0)) {
      ok = false;
      d = 1.0;
    }
    const double l = sqrt(d);
    Lf[j * NC + j] = l;
    for (int i = j + 1; i < NC; ++i) {
      if (!act[i]) continue;
      double v = 0.5 * (S[i * NC + j] + S[j * NC + i]);
      for (int k = 0; k < j; ++k) v -= Lf[i * NC + k] * Lf[j * NC + k];
      Lf[i * NC + j] = v / l;
    }
  }
  return ok;
}

// X = (L L^T)^-1 Y for NC right-hand sides (columns of Y), in place in Y.
template <int NC>
__device__ void chol_solve(const double* Lf, double* Y) {
  for (int col = 0; col < NC; ++col) {
    for (int i = 0; i < NC; ++i) {
      double v = Y[i * NC + col];
      for (int k = 0; k < i; ++k) v -= Lf[i * NC + k] * Y[k * NC + col];
      Y[i * NC + col] = v / Lf[i * NC + i];
    }
    for (int i = NC - 1; i >= 0; --i) {
      double v = Y[i * NC + col];
      for (int k = i + 1; k < NC; ++k) v -= Lf[k * NC + i] * Y[k * NC + col];
      Y[i * NC + col] = v / Lf[i * NC + i];
    }
  }
}

// alpha = (P^T Q)^-1 P^T R  (P^T Q at gram[0..36), P^T R at gram[36..72))
template <int NC>
__global__ void k_bcg_alpha(int ncells, BlockScalars* __restrict__ bs) {
  const int ci = blockIdx.x * blockDim.x + threadIdx.x;
  if (ci >= ncells) return;
  BlockScalars& b = bs[ci];
  if (b.done) return;
  double Lf[NC * NC], Y[NC * NC];
  chol<NC>(b.gram, b.active, Lf);
  for (int i = 0; i < NC; ++i)
    for (int j = 0; j < NC; ++j) Y[i * NC + j] = b.active[i] ? b.gram[KG + i * NC + j] : 0.0;
  chol_solve<NC>(Lf, Y);
  for (int k = 0; k < NC * NC; ++k) {
    b.alpha[k] = Y[k];
    b.lg[k] = Lf[k];
  }
}

// X += P alpha; R -= Q alpha; partial rr per case
template <int NC, int ND>
__global__ __launch_bounds__(256) void k_bcg_xr(Launch L, const BlockScalars* __restrict__ bs,
                                                const double* __restrict__ P, const double* __restrict__ Q,
                                                double* __restrict__ X, double* __restrict__ R,
                                                double* __restrict__ partial) {
  int ci;
  long long gid = 0;
  const bool ok = bnode(L, ci, gid);
  double rr[NC];
#pragma unroll
  for (int j = 0; j < NC; ++j) rr[j] = 0.0;
  if (ok && !bs[ci].done) {
    const double* al = bs[ci].alpha;
#pragma unroll
    for (int c = 0; c < ND; ++c) {
      double p[NC], q[NC];
#pragma unroll
      for (int i = 0; i < NC; ++i) {
        p[i] = P[(long long)(i * ND + c) * L.Ntot + gid];
        q[i] = Q[(long long)(i * ND + c) * L.Ntot + gid];
      }
#pragma unroll
      for (int j = 0; j < NC; ++j) {
        const long long idx = (long long)(j * ND + c) * L.Ntot + gid;
        double dx = 0.0, dr = 0.0;
#pragma unroll
        for (int i = 0; i < NC; ++i) {
          dx = fma(p[i], al[i * NC + j], dx);
          dr = fma(q[i], al[i * NC + j], dr);
        }
        X[idx] += dx;
        const double rn = R[idx] - dr;
        R[idx] = rn;
        rr[j] = fma(rn, rn, rr[j]);
      }
    }
  }
  double v[MAPLE_MAX_CASES];
#pragma unroll
  for (int j = 0; j < MAPLE_MAX_CASES; ++j) v[j] = j < NC ? rr[j] : 0.0;
  block_sum<MAPLE_MAX_CASES>(v, partial + (long long)lbid() * MAPLE_MAX_CASES);
}

// A cell is done when every case meets its target; counts iterations.
__global__ void k_bcg_check(int ncells, int nc, BlockScalars* __restrict__ bs, CaseScalars* __restrict__ sc,
                            int* __restrict__ done_flags) {
  const int ci = blockIdx.x * blockDim.x + threadIdx.x;
  if (ci >= ncells) return;
  BlockScalars& b = bs[ci];
  if (b.done) return;
  bool all = true;
  for (int j = 0; j < nc; ++j) {
    CaseScalars& s = sc[ci * nc + j];
    s.iters += 1;
    all = all && s.rr <= s.tgt2;
  }
  if (all) {
    b.done = 1;
    for (int j = 0; j < nc; ++j) {
      sc[ci * nc + j].done = 1;
      done_flags[ci * nc + j] = 1;
    }
  }
}

// beta = -(P^T Q)^-1 Q^T Z  (H = Z^T Q at gram[0..36))
template <int NC>
__global__ void k_bcg_beta(int ncells, BlockScalars* __restrict__ bs) {
  const int ci = blockIdx.x * blockDim.x + threadIdx.x;
  if (ci >= ncells) return;
  BlockScalars& b = bs[ci];
  if (b.done) return;
  double Y[NC * NC];
  for (int i = 0; i < NC; ++i)
    for (int j = 0; j < NC; ++j) Y[i * NC + j] = b.active[i] ? -b.gram[j * NC + i] : 0.0;   // (Q^T Z)_ij = H_ji
  chol_solve<NC>(b.lg, Y);
  for (int k = 0; k < NC * NC; ++k) b.beta[k] = Y[k];
}

// V = Z + P beta, written over P; partial V^T V
template <int NC, int ND>
__global__ __launch_bounds__(256) void k_bcg_v(Launch L, const BlockScalars* __restrict__ bs, int first,
                                               const double* __restrict__ Z, double* __restrict__ P,
                                               double* __restrict__ partial) {
  int ci, kc;
  long long gid = 0;
  const bool ok = bnode(L, ci, gid, &kc);
  double acc[NC * NC];
#pragma unroll
  for (int k = 0; k < NC * NC; ++k) acc[k] = 0.0;
  if (ok && !bs[ci].done) {
    const double* be = bs[ci].beta;
    const LevelCell cc = L.cells[kc];
    const double inv_n = 1.0 / ((double)cc.nx * cc.ny * cc.nz);
#pragma unroll
    for (int c = 0; c < ND; ++c) {
      double p[NC], v[NC];
#pragma unroll
      for (int i = 0; i < NC; ++i) p[i] = first ? 0.0 : P[(long long)(i * ND + c) * L.Ntot + gid];
#pragma unroll
      for (int j = 0; j < NC; ++j) {
        const long long idx = (long long)(j * ND + c) * L.Ntot + gid;
        // rigid translations are K's null space: a search direction along one
        // would be amplified by (P^T K P)^-1 without changing the residual
        double s = Z[idx] - bs[ci].gram[KG + j * ND + c] * inv_n;
        if (!first) {
#pragma unroll
          for (int i = 0; i < NC; ++i) s = fma(p[i], be[i * NC + j], s);
        }
        v[j] = s;
        P[idx] = s;
      }
#pragma unroll
      for (int i = 0; i < NC; ++i)
#pragma unroll
        for (int j = 0; j < NC; ++j) acc[i * NC + j] = fma(v[i], v[j], acc[i * NC + j]);
    }
  }
  block_sum<NC * NC>(acc, partial + (long long)lbid() * (2 * KG));
}

// T with V T orthonormal on the independent columns of V (gram[0..36) = V^T V).
template <int NC>
__global__ void k_bcg_orth(int ncells, BlockScalars* __restrict__ bs) {
  const int ci = blockIdx.x * blockDim.x + threadIdx.x;
  if (ci >= ncells) return;
  BlockScalars& b = bs[ci];
  if (b.done) return;
  double sc[NC], W[NC * NC], Lf[NC * NC];
  for (int j = 0; j < NC; ++j) {
    const double d = b.gram[j * NC + j];
    sc[j] = d > 0.0 ? 1.0 / sqrt(d) : 0.0;
  }
  for (int i = 0; i < NC; ++i)
    for (int j = 0; j < NC; ++j) W[i * NC + j] = 0.5 * (b.gram[i * NC + j] + b.gram[j * NC + i]) * sc[i] * sc[j];
  // Cholesky with column dropping: a column whose part outside the span of the
  // previous ones is below 1e-10 of its length is dependent
  for (int i = 0; i < NC * NC; ++i) Lf[i] = 0.0;
  for (int j = 0; j < NC; ++j) {
    int act = sc[j] > 0.0;
    double d = W[j * NC + j];
    for (int k = 0; k < j; ++k) d -= Lf[j * NC + k] * Lf[j * NC + k];
    if (!(d > 1e-10)) act = 0;
    b.active[j] = act;
    if (!act) {   // identity row and column: the dropped direction takes no part
      for (int k = 0; k < j; ++k) Lf[j * NC + k] = 0.0;
      Lf[j * NC + j] = 1.0;
      continue;
    }
    const double l = sqrt(d);
    Lf[j * NC + j] = l;
    for (int i = j + 1; i < NC; ++i) {
      double v = W[i * NC + j];
      for (int k = 0; k < j; ++k) v -= Lf[i * NC + k] * Lf[j * NC + k];
      Lf[i * NC + j] = v / l;
    }
  }
  // T = S L^-T restricted to the active columns: solve L^T T' = I (upper triangular)
  double Ti[NC * NC];
  for (int col = 0; col < NC; ++col) {
    for (int i = NC - 1; i >= 0; --i) {
      double v = (i == col) ? 1.0 : 0.0;
      for (int k = i + 1; k < NC; ++k) v -= Lf[k * NC + i] * Ti[k * NC + col];
      Ti[i * NC + col] = v / Lf[i * NC + i];
    }
  }
  for (int i = 0; i < NC; ++i)
    for (int j = 0; j < NC; ++j)
      b.T[i * NC + j] = (b.active[i] && b.active[j]) ? sc[i] * Ti[i * NC + j] : 0.0;
}

// P = V T (in place)
template <int NC, int ND>
__global__ __launch_bounds__(256) void k_bcg_p(Launch L, const BlockScalars* __restrict__ bs, double* __restrict__ P) {
  int ci;
  long long gid = 0;
  if (!bnode(L, ci, gid)) return;
  if (bs[ci].done) return;
  const double* T = bs[ci].T;
#pragma unroll
  for (int c = 0; c < ND; ++c) {
    double v[NC];
#pragma unroll
    for (int i = 0; i < NC; ++i) v[i] = P[(long long)(i * ND + c) * L.Ntot + gid];
#pragma unroll
    for (int j = 0; j < NC; ++j) {
      double s = 0.0;
#pragma unroll
      for (int i = 0; i < NC; ++i) s = fma(v[i], T[i * NC + j], s);
      P[(long long)(j * ND + c) * L.Ntot + gid] = s;
    }
  }
}

// A cell whose cases all start converged is done; otherwise every case of it
// iterates (block directions mix the cases, so no case of a live cell may be skipped).
__global__ void k_bcg_init(int ncells, int nc, BlockScalars* __restrict__ bs, CaseScalars* __restrict__ sc,
                           int* __restrict__ done_flags) {
  const int ci = blockIdx.x * blockDim.x + threadIdx.x;
  if (ci >= ncells) return;
  BlockScalars& b = bs[ci];
  bool all = true;
  for (int j = 0; j < nc; ++j) all = all && sc[ci * nc + j].done;
  b.done = all ? 1 : 0;
  for (int j = 0; j < nc; ++j) {
    sc[ci * nc + j].done = all ? 1 : 0;
    done_flags[ci * nc + j] = all ? 1 : 0;
  }
  for (int j = 0; j < MAPLE_MAX_CASES; ++j) b.active[j] = 0;
}

// ng = 2: A^T B0 and A^T B1; ng = 1: A^T B0 and the component sums of A.
template <int NC, int ND>
void gram_t(const Launch& L, int ng, const double* A, const double* B0, const double* B1, double* partial, double* out,
            int offset, hipStream_t s) {
  int nv;
  if (ng == 1) {
    k_gram<NC, ND, 1, true><<<L.nblocks, 256, 0, s>>>(L, A, B0, B1, partial);
    nv = NC * NC + NC * ND;
  } else {
    k_gram<NC, ND, 2, false><<<L.nblocks, 256, 0, s>>>(L, A, B0, B1, partial);
    nv = 2 * NC * NC;
  }
  k_gram_reduce<<<dim3(L.ncells, nv), 64, 0, s>>>(L.cells, L.blk_start, partial, NC * NC, ng, out,
                                                  (int)(sizeof(BlockScalars) / sizeof(double)), offset);
}

}  // namespace

static_assert(sizeof(BlockScalars) % sizeof(double) == 0, "BlockScalars must be a whole number of doubles");

#define BCG_DISPATCH(NCV, NDV, CALL) \
  if (nc == NCV && nd == NDV) {      \
    CALL;                            \
    return;                          \
  }

void bcg_gram(const Launch& L, int nd, int nc, int ng, const double* A, const double* B0, const double* B1,
              double* partial, BlockScalars* bs, hipStream_t s) {
  double* out = reinterpret_cast<double*>(bs);
  const int off = (int)(offsetof(BlockScalars, gram) / sizeof(double));
  BCG_DISPATCH(6, 3, (gram_t<6, 3>(L, ng, A, B0, B1, partial, out, off, s)));
  BCG_DISPATCH(3, 1, (gram_t<3, 1>(L, ng, A, B0, B1, partial, out, off, s)));
}

void bcg_alpha(int ncells, int nc, BlockScalars* bs, hipStream_t s) {
  const int nb = (ncells + 63) / 64;
  if (nc == 6) k_bcg_alpha<6><<<nb, 64, 0, s>>>(ncells, bs);
  else k_bcg_alpha<3><<<nb, 64, 0, s>>>(ncells, bs);
}

void bcg_update_xr(const Launch& L, int nd, int nc, const BlockScalars* bs, const double* P, const double* Q,
                   double* X, double* R, double* partial, hipStream_t s) {
  BCG_DISPATCH(6, 3, (k_bcg_xr<6, 3><<<L.nblocks, 256, 0, s>>>(L, bs, P, Q, X, R, partial)));
  BCG_DISPATCH(3, 1, (k_bcg_xr<3, 1><<<L.nblocks, 256, 0, s>>>(L, bs, P, Q, X, R, partial)));
}

void bcg_check(int ncells, int nc, BlockScalars* bs, CaseScalars* sc, int* done_flags, hipStream_t s) {
  k_bcg_check<<<(ncells + 63) / 64, 64, 0, s>>>(ncells, nc, bs, sc, done_flags);
}

void bcg_beta(int ncells, int nc, BlockScalars* bs, hipStream_t s) {
  const int nb = (ncells + 63) / 64;
  if (nc == 6) k_bcg_beta<6><<<nb, 64, 0, s>>>(ncells, bs);
  else k_bcg_beta<3><<<nb, 64, 0, s>>>(ncells, bs);
}

void bcg_new_direction(const Launch& L, int nd, int nc, bool first, BlockScalars* bs, const double* Z, double* P,
                       double* partial, hipStream_t s) {
  double* out = reinterpret_cast<double*>(bs);
  const int off = (int)(offsetof(BlockScalars, gram) / sizeof(double));
  const int stride = (int)(sizeof(BlockScalars) / sizeof(double));
  const int nb = (L.ncells + 63) / 64;
  if (nc == 6 && nd == 3) {
    k_bcg_v<6, 3><<<L.nblocks, 256, 0, s>>>(L, bs, first ? 1 : 0, Z, P, partial);
    k_gram_reduce<<<dim3(L.ncells, 36), 64, 0, s>>>(L.cells, L.blk_start, partial, 36, 1, out, stride, off);
    k_bcg_orth<6><<<nb, 64, 0, s>>>(L.ncells, bs);
    k_bcg_p<6, 3><<<L.nblocks, 256, 0, s>>>(L, bs, P);
  } else {
    k_bcg_v<3, 1><<<L.nblocks, 256, 0, s>>>(L, bs, first ? 1 : 0, Z, P, partial);
    k_gram_reduce<<<dim3(L.ncells, 9), 64, 0, s>>>(L.cells, L.blk_start, partial, 9, 1, out, stride, off);
    k_bcg_orth<3><<<nb, 64, 0, s>>>(L.ncells, bs);
    k_bcg_p<3, 1><<<L.nblocks, 256, 0, s>>>(L, bs, P);
  }
}

void bcg_init(int ncells, int nc, BlockScalars* bs, CaseScalars* sc, int* done_flags, hipStream_t s) {
  k_bcg_init<<<(ncells + 63) / 64, 64, 0, s>>>(ncells, nc, bs, sc, done_flags);
}
