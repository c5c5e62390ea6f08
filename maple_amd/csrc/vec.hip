// Vector and scalar kernels of the multi-right-hand-side preconditioned CG.
//
// All load cases of all cells advance together; every (cell, case) pair keeps
// its own CG scalars on the device (CaseScalars), so no value goes to the host
// inside an iteration.  A converged pair gets alpha = beta = 0 and stays fixed
// while the others finish.  Dot products are reduced deterministically:
// per-block partials in a fixed order, then one wavefront per (cell, case).
#include "kernels.h"

__global__ __launch_bounds__(64) void k_reduce_partials(const LevelCell* __restrict__ cells,
                                                        const int* __restrict__ blk_start, int ncase,
                                                        const double* __restrict__ partial,
                                                        CaseScalars* __restrict__ sc, int which) {
  const int k = blockIdx.x, j = blockIdx.y, ci = cells[k].id;
  const int b0 = blk_start[k], b1 = blk_start[k + 1];
  double v = 0.0;
  for (int b = b0 + (int)threadIdx.x; b < b1; b += 64) v += partial[(long long)b * MAPLE_MAX_CASES + j];
  v = wave_sum(v);
  if (threadIdx.x == 0) {
    CaseScalars& s = sc[ci * ncase + j];
    if (which == SC_FF) s.ff = v;
    else if (which == SC_RZ_NEW) s.rz_new = v;
    else if (which == SC_PQ) s.pq = v;
    else if (which == SC_RZO) s.rzo = v;
    else if (which == SC_RRS) s.rrs = v;
    else if (which == SC_RR_LIVE) { if (!s.done) s.rr = v; }
    else s.rr = v;
  }
}

void reduce_partials(const Launch& L, int ncase, const double* partial, CaseScalars* sc, int which, hipStream_t s) {
  k_reduce_partials<<<dim3(L.ncells, ncase), 64, 0, s>>>(L.cells, L.blk_start, ncase, partial, sc, which);
}

// ci: the id of the block's cell (per-cell arrays); gid: the thread's node
__device__ __forceinline__ bool vnode(const Launch& L, int& ci, long long& gid) {
  const int k = find_cell(L.blk_start, L.ncells, lbid());
  const LevelCell c = L.cells[k];
  ci = c.id;
  const long long n = (long long)(lbid() - L.blk_start[k]) * 256 + threadIdx.x;
  if (n >= (long long)c.nx * c.ny * c.nz) return false;
  gid = c.off + n;
  return true;
}

// The vector kernels come in two forms: ND and NC (components per node, load
// cases) as compile-time constants for the two physics (3 x 6 elastic, 1 x 3
// thermal), fully unrolled so that every load of a thread is in flight at once
// and a dot product needs one block reduction for all its cases; and a generic
// form for any other shape.
template <int ND, int NC, typename T>
__global__ __launch_bounds__(256) void k_scale_dinv_t(Launch L, T omega, const T* __restrict__ dinv,
                                                      const T* __restrict__ b, T* __restrict__ x) {
  int ci; long long gid;
  if (!vnode(L, ci, gid)) return;
  if (cell_done(L.done, L.ndc, ci)) return;
#pragma unroll
  for (int i = 0; i < ND; ++i) {
    const T w = omega * dinv[(long long)i * L.Ntot + gid];
#pragma unroll
    for (int j = 0; j < NC; ++j) {
      const long long idx = (long long)(j * ND + i) * L.Ntot + gid;
      x[idx] = w * b[idx];
    }
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_scale_dinv(Launch L, int nd, int ncase, T omega,
                                                    const T* __restrict__ dinv, const T* __restrict__ b,
                                                    T* __restrict__ x) {
  int ci; long long gid;
  if (!vnode(L, ci, gid)) return;
  if (cell_done(L.done, L.ndc, ci)) return;
  for (int i = 0; i < nd; ++i) {
    const T w = omega * dinv[(long long)i * L.Ntot + gid];
    for (int j = 0; j < ncase; ++j) {
      const long long idx = (long long)(j * nd + i) * L.Ntot + gid;
      x[idx] = w * b[idx];
    }
  }
}

template <typename T>
void vec_scale_dinv(const Launch& L, int nd, int ncase, T omega, const T* dinv, const T* b, T* x, hipStream_t s) {
  if (nd == 3 && ncase == 6) k_scale_dinv_t<3, 6, T><<<L.nblocks, 256, 0, s>>>(L, omega, dinv, b, x);
  else if (nd == 1 && ncase == 3) k_scale_dinv_t<1, 3, T><<<L.nblocks, 256, 0, s>>>(L, omega, dinv, b, x);
  else k_scale_dinv<T><<<L.nblocks, 256, 0, s>>>(L, nd, ncase, omega, dinv, b, x);
}
template void vec_scale_dinv<double>(const Launch&, int, int, double, const double*, const double*, double*, hipStream_t);
template void vec_scale_dinv<float>(const Launch&, int, int, float, const float*, const float*, float*, hipStream_t);

// A launch with convergence flags (L.done) skips converged (cell, case) pairs: their
// partial is zero and their vectors are not read.
template <int ND, int NC, typename Ta, typename Tb>
__global__ __launch_bounds__(256) void k_dot_t(Launch L, const Ta* __restrict__ a, const Tb* __restrict__ b,
                                               double* __restrict__ partial) {
  int ci; long long gid = 0;
  const bool ok = vnode(L, ci, gid);
  double v[NC];
#pragma unroll
  for (int j = 0; j < NC; ++j) {
    v[j] = 0.0;
    if (ok && !case_done(L.done, L.ndc, ci, j))
#pragma unroll
      for (int i = 0; i < ND; ++i) {
        const long long idx = (long long)(j * ND + i) * L.Ntot + gid;
        v[j] += (double)a[idx] * (double)b[idx];
      }
  }
  block_sum<NC>(v, partial + (long long)lbid() * MAPLE_MAX_CASES);
}

template <typename Ta, typename Tb>
__global__ __launch_bounds__(256) void k_dot(Launch L, int nd, int ncase, const Ta* __restrict__ a,
                                             const Tb* __restrict__ b, double* __restrict__ partial) {
  int ci; long long gid = 0;
  const bool ok = vnode(L, ci, gid);
  for (int j = 0; j < ncase; ++j) {
    double v[1] = {0.0};
    if (ok && !case_done(L.done, L.ndc, ci, j))
      for (int i = 0; i < nd; ++i) {
        const long long idx = (long long)(j * nd + i) * L.Ntot + gid;
        v[0] += (double)a[idx] * (double)b[idx];
      }
    block_sum<1>(v, partial + (long long)lbid() * MAPLE_MAX_CASES + j);
  }
}

template <typename Ta, typename Tb>
void vec_dot(const Launch& L, int nd, int ncase, const Ta* a, const Tb* b, double* partial, hipStream_t s) {
  if (nd == 3 && ncase == 6) k_dot_t<3, 6, Ta, Tb><<<L.nblocks, 256, 0, s>>>(L, a, b, partial);
  else if (nd == 1 && ncase == 3) k_dot_t<1, 3, Ta, Tb><<<L.nblocks, 256, 0, s>>>(L, a, b, partial);
  else k_dot<Ta, Tb><<<L.nblocks, 256, 0, s>>>(L, nd, ncase, a, b, partial);
}
template void vec_dot<double, double>(const Launch&, int, int, const double*, const double*, double*, hipStream_t);
template void vec_dot<double, float>(const Launch&, int, int, const double*, const float*, double*, hipStream_t);
template void vec_dot<float, float>(const Launch&, int, int, const float*, const float*, double*, hipStream_t);

// (r, z) per case and the sums of z per (case, component) in one pass
template <int ND, int NC, typename Tz>
__global__ __launch_bounds__(256) void k_dot_rz_t(Launch L, const double* __restrict__ r, const Tz* __restrict__ z,
                                                  double* __restrict__ partial, double* __restrict__ zpart) {
  int ci; long long gid = 0;
  const bool ok = vnode(L, ci, gid);
  double v[NC], w[NC * ND];
#pragma unroll
  for (int j = 0; j < NC; ++j) {
    v[j] = 0.0;
#pragma unroll
    for (int i = 0; i < ND; ++i) w[j * ND + i] = 0.0;
    if (ok && !case_done(L.done, L.ndc, ci, j))
#pragma unroll
      for (int i = 0; i < ND; ++i) {
        const long long idx = (long long)(j * ND + i) * L.Ntot + gid;
        const double zv = (double)z[idx];
        v[j] += r[idx] * zv;
        w[j * ND + i] = zv;
      }
  }
  block_sum<NC>(v, partial + (long long)lbid() * MAPLE_MAX_CASES);
  block_sum<NC * ND>(w, zpart + (long long)lbid() * 18);
}

template <typename Tz>
__global__ __launch_bounds__(256) void k_dot_rz(Launch L, int nd, int ncase, const double* __restrict__ r,
                                                const Tz* __restrict__ z, double* __restrict__ partial,
                                                double* __restrict__ zpart) {
  int ci; long long gid = 0;
  const bool ok = vnode(L, ci, gid);
  for (int j = 0; j < ncase; ++j) {
    double v[1] = {0.0};
    double w[3] = {0.0, 0.0, 0.0};
    if (ok && !case_done(L.done, L.ndc, ci, j))
      for (int i = 0; i < nd; ++i) {
        const long long idx = (long long)(j * nd + i) * L.Ntot + gid;
        const double zv = (double)z[idx];
        v[0] += r[idx] * zv;
        w[i] = zv;
      }
    block_sum<1>(v, partial + (long long)lbid() * MAPLE_MAX_CASES + j);
    block_sum<3>(w, zpart + (long long)lbid() * 18 + j * 3);
  }
}

template <typename Tz>
void pcg_dot_rz(const Launch& L, int nd, int ncase, const double* r, const Tz* z, double* partial, double* zpart,
                hipStream_t s) {
  if (nd == 3 && ncase == 6) k_dot_rz_t<3, 6, Tz><<<L.nblocks, 256, 0, s>>>(L, r, z, partial, zpart);
  else if (nd == 1 && ncase == 3) k_dot_rz_t<1, 3, Tz><<<L.nblocks, 256, 0, s>>>(L, r, z, partial, zpart);
  else k_dot_rz<Tz><<<L.nblocks, 256, 0, s>>>(L, nd, ncase, r, z, partial, zpart);
}
template void pcg_dot_rz<double>(const Launch&, int, int, const double*, const double*, double*, double*, hipStream_t);
template void pcg_dot_rz<float>(const Launch&, int, int, const double*, const float*, double*, double*, hipStream_t);

// one wavefront per listed (cell, case); the generic dot kernel stores its sums with a
// stride of 3 per case, the unrolled one with nd
__global__ __launch_bounds__(64) void k_reduce_zmean(const LevelCell* __restrict__ cells,
                                                     const int* __restrict__ blk_start, int nd, int ncase,
                                                     int stride, const double* __restrict__ zpart,
                                                     CaseScalars* __restrict__ sc) {
  const int k = blockIdx.x, j = blockIdx.y;
  const LevelCell c = cells[k];
  const int b0 = blk_start[k], b1 = blk_start[k + 1];
  const double inv = 1.0 / ((double)c.nx * c.ny * c.nz);
  for (int i = 0; i < nd; ++i) {
    double v = 0.0;
    for (int b = b0 + (int)threadIdx.x; b < b1; b += 64) v += zpart[(long long)b * 18 + j * stride + i];
    v = wave_sum(v);
    if (threadIdx.x == 0) sc[c.id * ncase + j].zmean[i] = v * inv;
  }
}

void pcg_reduce_zmean(const Launch& L, int nd, int ncase, const double* zpart, CaseScalars* sc, hipStream_t s) {
  const bool unrolled = (nd == 3 && ncase == 6) || (nd == 1 && ncase == 3);
  k_reduce_zmean<<<dim3(L.ncells, ncase), 64, 0, s>>>(L.cells, L.blk_start, nd, ncase, unrolled ? nd : 3, zpart, sc);
}

template <int ND, int NC>
__global__ __launch_bounds__(256) void k_update_ur_t(Launch L, const CaseScalars* __restrict__ sc,
                                                     const double* __restrict__ p, const double* __restrict__ q,
                                                     double* __restrict__ u, double* __restrict__ r,
                                                     double* __restrict__ partial) {
  int ci; long long gid = 0;
  const bool ok = vnode(L, ci, gid);
  double v[NC];
#pragma unroll
  for (int j = 0; j < NC; ++j) {
    v[j] = 0.0;
    const double al = sc[ci * NC + j].alpha;
    // a stopped case keeps u and r untouched (its p may be stale); a live one with
    // alpha = 0 (a restart) keeps them too but still sums its residual norm
    if (ok && !case_done(L.done, L.ndc, ci, j)) {
#pragma unroll
      for (int i = 0; i < ND; ++i) {
        const long long idx = (long long)(j * ND + i) * L.Ntot + gid;
        u[idx] = fma(al, p[idx], u[idx]);
        const double rn = fma(-al, q[idx], r[idx]);
        r[idx] = rn;
        v[j] += rn * rn;
      }
    }
  }
  block_sum<NC>(v, partial + (long long)lbid() * MAPLE_MAX_CASES);
}

__global__ __launch_bounds__(256) void k_update_ur(Launch L, int nd, int ncase, const CaseScalars* __restrict__ sc,
                                                   const double* __restrict__ p, const double* __restrict__ q,
                                                   double* __restrict__ u, double* __restrict__ r,
                                                   double* __restrict__ partial) {
  int ci; long long gid = 0;
  const bool ok = vnode(L, ci, gid);
  for (int j = 0; j < ncase; ++j) {
    double v[1] = {0.0};
    const double al = ok ? sc[ci * ncase + j].alpha : 0.0;
    if (ok && !case_done(L.done, L.ndc, ci, j)) {
      for (int i = 0; i < nd; ++i) {
        const long long idx = (long long)(j * nd + i) * L.Ntot + gid;
        u[idx] = fma(al, p[idx], u[idx]);
        const double rn = fma(-al, q[idx], r[idx]);
        r[idx] = rn;
        v[0] += rn * rn;
      }
    }
    block_sum<1>(v, partial + (long long)lbid() * MAPLE_MAX_CASES + j);
  }
}

// The mixed form's update: u += alpha p, r -= alpha q as k_update_ur, and in the same
// pass the FP32 copy of the new r (the FP32 preconditioner's input) and, when the
// previous preconditioned residual was FP32 (zold != nullptr), the flexible-CG dot
// (r_new, z_old) into partial2.  Saves the separate copy and dot passes.
template <int ND, int NC>
__global__ __launch_bounds__(256) void k_update_ur_mixed(Launch L, const CaseScalars* __restrict__ sc,
                                                         const double* __restrict__ p, const double* __restrict__ q,
                                                         double* __restrict__ u, double* __restrict__ r,
                                                         float* __restrict__ r32, const float* __restrict__ zold,
                                                         double* __restrict__ partial, double* __restrict__ partial2) {
  int ci; long long gid = 0;
  const bool ok = vnode(L, ci, gid);
  double v[NC], w[NC];
#pragma unroll
  for (int j = 0; j < NC; ++j) {
    v[j] = w[j] = 0.0;
    const double al = sc[ci * NC + j].alpha;
    if (ok && !case_done(L.done, L.ndc, ci, j)) {   // a stopped case keeps u, r and r32 untouched
#pragma unroll
      for (int i = 0; i < ND; ++i) {
        const long long idx = (long long)(j * ND + i) * L.Ntot + gid;
        u[idx] = fma(al, p[idx], u[idx]);
        const double rn = fma(-al, q[idx], r[idx]);
        r[idx] = rn;
        r32[idx] = (float)rn;
        v[j] += rn * rn;
        if (zold != nullptr) w[j] += rn * (double)zold[idx];
      }
    }
  }
  block_sum<NC>(v, partial + (long long)lbid() * MAPLE_MAX_CASES);
  if (zold != nullptr) block_sum<NC>(w, partial2 + (long long)lbid() * MAPLE_MAX_CASES);
}

int pcg_update_ur_mixed(const Launch& L, int nd, int ncase, const CaseScalars* sc, const double* p, const double* q,
                        double* u, double* r, float* r32, const float* zold, double* partial, double* partial2,
                        hipStream_t s) {
  if (nd == 3 && ncase == 6)
    k_update_ur_mixed<3, 6><<<L.nblocks, 256, 0, s>>>(L, sc, p, q, u, r, r32, zold, partial, partial2);
  else if (nd == 1 && ncase == 3)
    k_update_ur_mixed<1, 3><<<L.nblocks, 256, 0, s>>>(L, sc, p, q, u, r, r32, zold, partial, partial2);
  else
    return -1;
  return 0;
}

void pcg_update_ur(const Launch& L, int nd, int ncase, const CaseScalars* sc, const double* p, const double* q,
                   double* u, double* r, double* partial, hipStream_t s) {
  if (nd == 3 && ncase == 6) k_update_ur_t<3, 6><<<L.nblocks, 256, 0, s>>>(L, sc, p, q, u, r, partial);
  else if (nd == 1 && ncase == 3) k_update_ur_t<1, 3><<<L.nblocks, 256, 0, s>>>(L, sc, p, q, u, r, partial);
  else k_update_ur<<<L.nblocks, 256, 0, s>>>(L, nd, ncase, sc, p, q, u, r, partial);
}

template <int ND, int NC, typename Tz>
__global__ __launch_bounds__(256) void k_update_p_t(Launch L, const CaseScalars* __restrict__ sc,
                                                    const Tz* __restrict__ z, double* __restrict__ p) {
  int ci; long long gid;
  if (!vnode(L, ci, gid)) return;
#pragma unroll
  for (int j = 0; j < NC; ++j) {
    const CaseScalars& cs = sc[ci * NC + j];
    if (cs.done) continue;   // a converged case's p is never used again
    const double be = cs.beta;
#pragma unroll
    for (int i = 0; i < ND; ++i) {
      const long long idx = (long long)(j * ND + i) * L.Ntot + gid;
      const double zc = (double)z[idx] - cs.zmean[i];
      p[idx] = be == 0.0 ? zc : fma(be, p[idx], zc);
    }
  }
}

__global__ __launch_bounds__(256) void k_update_p(Launch L, int nd, int ncase, const CaseScalars* __restrict__ sc,
                                                  const double* __restrict__ z, double* __restrict__ p) {
  int ci; long long gid;
  if (!vnode(L, ci, gid)) return;
  for (int j = 0; j < ncase; ++j) {
    const CaseScalars& cs = sc[ci * ncase + j];
    if (cs.done) continue;
    const double be = cs.beta;
    for (int i = 0; i < nd; ++i) {
      const long long idx = (long long)(j * nd + i) * L.Ntot + gid;
      const double zc = z[idx] - cs.zmean[i];
      p[idx] = be == 0.0 ? zc : fma(be, p[idx], zc);
    }
  }
}

void pcg_update_p(const Launch& L, int nd, int ncase, const CaseScalars* sc, const double* z, double* p,
                  hipStream_t s) {
  if (nd == 3 && ncase == 6) k_update_p_t<3, 6, double><<<L.nblocks, 256, 0, s>>>(L, sc, z, p);
  else if (nd == 1 && ncase == 3) k_update_p_t<1, 3, double><<<L.nblocks, 256, 0, s>>>(L, sc, z, p);
  else k_update_p<<<L.nblocks, 256, 0, s>>>(L, nd, ncase, sc, z, p);
}

// p = z + beta p with the FP32 preconditioned residual of the mixed form
int pcg_update_p32(const Launch& L, int nd, int ncase, const CaseScalars* sc, const float* z, double* p,
                   hipStream_t s) {
  if (nd == 3 && ncase == 6) k_update_p_t<3, 6, float><<<L.nblocks, 256, 0, s>>>(L, sc, z, p);
  else if (nd == 1 && ncase == 3) k_update_p_t<1, 3, float><<<L.nblocks, 256, 0, s>>>(L, sc, z, p);
  else return -1;
  return 0;
}

__global__ void k_pick_src(int n, int have_prev, int have_start, int start_src, CaseScalars* __restrict__ sc) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  CaseScalars& c = sc[t];
  int src = 0;
  double best = c.ff;
  if (have_prev && c.rr < best) {
    src = 1;
    best = c.rr;
  }
  if (have_start && c.rrs < best) src = start_src;
  c.src = src;
}

__global__ __launch_bounds__(256) void k_apply_start(Launch L, int nd, int ncase, const CaseScalars* __restrict__ sc,
                                                     double* __restrict__ u, double* __restrict__ r,
                                                     const double* __restrict__ us, const double* __restrict__ rs,
                                                     const double* __restrict__ f) {
  int ci; long long gid;
  if (!vnode(L, ci, gid)) return;
  for (int j = 0; j < ncase; ++j) {
    const int src = sc[ci * ncase + j].src;
    if (src == 1) continue;
    for (int i = 0; i < nd; ++i) {
      const long long idx = (long long)(j * nd + i) * L.Ntot + gid;
      u[idx] = src >= 2 ? us[idx] : 0.0;
      r[idx] = src >= 2 ? rs[idx] : f[idx];
    }
  }
}

void pcg_pick_start(const Launch& L, int ncells, int nd, int ncase, bool have_prev, bool have_start, CaseScalars* sc,
                    double* u, double* r, const double* us, const double* rs, const double* f, hipStream_t s,
                    int start_src) {
  const int n = ncells * ncase;
  k_pick_src<<<(n + 63) / 64, 64, 0, s>>>(n, have_prev ? 1 : 0, have_start ? 1 : 0, start_src, sc);
  k_apply_start<<<L.nblocks, 256, 0, s>>>(L, nd, ncase, sc, u, r, us, rs, f);
}

// ff and rr must already hold ||f||^2 and ||r0||^2.
__global__ void k_pcg_init(int n, int ncase, double rtol, const double* __restrict__ atol2,
                           CaseScalars* __restrict__ sc, int* __restrict__ done_flags) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  CaseScalars& s = sc[t];
  s.tgt2 = fmax(rtol * rtol * s.ff, atol2[t / ncase]);
  s.done = s.rr <= s.tgt2 ? 1 : 0;
  s.iters = 0;
  s.brk = 0;
  s.nrst = 0;
  s.rr_min = s.rr;
  s.it_min = 0;
  s.alpha = s.beta = 0.0;
  s.rz = s.rz_new;
  done_flags[t] = s.done;
}

void pcg_init_scalars(int ncells, int ncase, double rtol, const double* atol2, CaseScalars* sc, int* done_flags,
                      hipStream_t s) {
  const int n = ncells * ncase;
  k_pcg_init<<<(n + 63) / 64, 64, 0, s>>>(n, ncase, rtol, atol2, sc, done_flags);
}

__global__ void k_pcg_alpha(int n, CaseScalars* __restrict__ sc, int* __restrict__ done_flags,
                            int* __restrict__ restart_flag) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  CaseScalars& s = sc[t];
  if (s.done) {
    s.alpha = 0.0;
    return;
  }
  if (s.pq > 0.0) {
    s.alpha = s.rz / s.pq;
    s.nrst = 0;
    return;
  }
  // p^T K p <= 0 (or NaN) on a live case: the preconditioned direction is not a descent
  // direction of this operator (a preconditioner that is not SPD there, or a direction
  // lost in the operator's null space).  No step; the next direction is p = z.
  s.alpha = 0.0;
  s.nrst += 1;
  if (s.nrst >= MAPLE_CG_RESTARTS) {
    s.brk = 1;
    s.done = 1;
    done_flags[t] = 1;
  } else {
    *restart_flag = 1;
  }
}

void pcg_alpha(int ncells, int ncase, CaseScalars* sc, int* done_flags, int* restart_flag, hipStream_t s) {
  const int n = ncells * ncase;
  k_pcg_alpha<<<(n + 63) / 64, 64, 0, s>>>(n, sc, done_flags, restart_flag);
}

__global__ void k_pcg_check(int n, CaseScalars* __restrict__ sc, int* __restrict__ done_flags) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  CaseScalars& s = sc[t];
  if (!s.done) {
    s.iters += 1;
    if (s.rr <= s.tgt2 || !(s.rr == s.rr)) {
      s.done = 1;
    } else if (s.rr < s.rr_min) {
      s.rr_min = s.rr;
      s.it_min = s.iters;
    } else if (s.iters - s.it_min >= MAPLE_CG_STALL || s.rr > 1e12 * s.rr_min) {
      // below the attainable accuracy (the recursive residual no longer falls) or
      // diverging: stop with the fields reached, before rounding turns them into NaN
      s.done = 1;
      s.brk = 2;
    }
  }
  done_flags[t] = s.done;
}

void pcg_check(int ncells, int ncase, CaseScalars* sc, int* done_flags, hipStream_t s) {
  const int n = ncells * ncase;
  k_pcg_check<<<(n + 63) / 64, 64, 0, s>>>(n, sc, done_flags);
}

__global__ void k_pcg_beta(int n, int flexible, CaseScalars* __restrict__ sc) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  CaseScalars& s = sc[t];
  const double num = flexible ? s.rz_new - s.rzo : s.rz_new;
  // a restarted case (no step this iteration) starts over from p = z
  s.beta = (!s.done && s.nrst == 0 && s.rz != 0.0) ? num / s.rz : 0.0;
  s.rz = s.rz_new;
}

void pcg_beta(int ncells, int ncase, bool flexible, CaseScalars* sc, hipStream_t s) {
  const int n = ncells * ncase;
  k_pcg_beta<<<(n + 63) / 64, 64, 0, s>>>(n, flexible ? 1 : 0, sc);
}

template <typename Td, typename Ts>
__global__ void k_copy(long long n, const Ts* __restrict__ src, Td* __restrict__ dst) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = (Td)src[i];
}

template <typename Td, typename Ts>
void vec_copy(long long n, const Ts* src, Td* dst, hipStream_t s) {
  if (n > 0) k_copy<Td, Ts><<<(unsigned)((n + 255) / 256), 256, 0, s>>>(n, src, dst);
}
template void vec_copy<double, double>(long long, const double*, double*, hipStream_t);
template void vec_copy<float, double>(long long, const double*, float*, hipStream_t);
template void vec_copy<double, float>(long long, const float*, double*, hipStream_t);

// dst = src over the (cell, case) pairs still iterating (all of them for a launch
// without convergence flags)
template <int ND, int NC, typename Td, typename Ts>
__global__ __launch_bounds__(256) void k_copy_cases_t(Launch L, const Ts* __restrict__ src, Td* __restrict__ dst) {
  int ci; long long gid;
  if (!vnode(L, ci, gid)) return;
#pragma unroll
  for (int j = 0; j < NC; ++j) {
    if (case_done(L.done, L.ndc, ci, j)) continue;
#pragma unroll
    for (int i = 0; i < ND; ++i) {
      const long long idx = (long long)(j * ND + i) * L.Ntot + gid;
      dst[idx] = (Td)src[idx];
    }
  }
}

template <typename Td, typename Ts>
__global__ __launch_bounds__(256) void k_copy_cases(Launch L, int nd, int ncase, const Ts* __restrict__ src,
                                                    Td* __restrict__ dst) {
  int ci; long long gid;
  if (!vnode(L, ci, gid)) return;
  for (int j = 0; j < ncase; ++j) {
    if (case_done(L.done, L.ndc, ci, j)) continue;
    for (int i = 0; i < nd; ++i) {
      const long long idx = (long long)(j * nd + i) * L.Ntot + gid;
      dst[idx] = (Td)src[idx];
    }
  }
}

template <typename Td, typename Ts>
void vec_copy_cases(const Launch& L, int nd, int ncase, const Ts* src, Td* dst, hipStream_t s) {
  if (nd == 3 && ncase == 6) k_copy_cases_t<3, 6, Td, Ts><<<L.nblocks, 256, 0, s>>>(L, src, dst);
  else if (nd == 1 && ncase == 3) k_copy_cases_t<1, 3, Td, Ts><<<L.nblocks, 256, 0, s>>>(L, src, dst);
  else k_copy_cases<Td, Ts><<<L.nblocks, 256, 0, s>>>(L, nd, ncase, src, dst);
}
template void vec_copy_cases<float, double>(const Launch&, int, int, const double*, float*, hipStream_t);
template void vec_copy_cases<double, float>(const Launch&, int, int, const float*, double*, hipStream_t);

template <typename T>
__global__ void k_zero(long long n, T* __restrict__ dst) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = T(0);
}

template <typename T>
void vec_zero(long long n, T* dst, hipStream_t s) {
  if (n > 0) k_zero<T><<<(unsigned)((n + 255) / 256), 256, 0, s>>>(n, dst);
}
template void vec_zero<double>(long long, double*, hipStream_t);
template void vec_zero<float>(long long, float*, hipStream_t);
template void vec_zero<int>(long long, int*, hipStream_t);

// ---- start projected on earlier solutions ------------------------------------
// Per (cell, case) j, with basis vectors h_a (a < k) and their products g_a = K h_a:
// the packed upper triangle of G_ab = (h_a, g_b) (a <= b, rows in order) and
// b_a = (h_a, f), MAPLE_PROJ_NG values per block and case.  Slots a >= k read nothing
// and sum zero.  The cases run one after the other so a thread holds one case's sums.
template <int ND>
__global__ __launch_bounds__(256) void k_proj_gram(Launch L, int ncase, int k, ProjBasis B,
                                                   const double* __restrict__ f, double* __restrict__ partial) {
  int ci; long long gid = 0;
  const bool ok = vnode(L, ci, gid);
  for (int j = 0; j < ncase; ++j) {
    double v[MAPLE_PROJ_NG];
#pragma unroll
    for (int t = 0; t < MAPLE_PROJ_NG; ++t) v[t] = 0.0;
    if (ok) {
#pragma unroll
      for (int i = 0; i < ND; ++i) {
        const long long idx = (long long)(j * ND + i) * L.Ntot + gid;
        double hv[MAPLE_PROJ_MAX], gv[MAPLE_PROJ_MAX];
#pragma unroll
        for (int a = 0; a < MAPLE_PROJ_MAX; ++a) {
          hv[a] = a < k ? B.h[a][idx] : 0.0;
          gv[a] = a < k ? B.g[a][idx] : 0.0;
        }
        const double fv = f[idx];
        int t = 0;
#pragma unroll
        for (int a = 0; a < MAPLE_PROJ_MAX; ++a)
#pragma unroll
          for (int b = a; b < MAPLE_PROJ_MAX; ++b) v[t++] += hv[a] * gv[b];
#pragma unroll
        for (int a = 0; a < MAPLE_PROJ_MAX; ++a) v[t++] += hv[a] * fv;
      }
    }
    block_sum<MAPLE_PROJ_NG>(v, partial + ((long long)lbid() * ncase + j) * MAPLE_PROJ_NG);
  }
}

__device__ __forceinline__ int proj_packed(int a, int b) {   // a <= b
  return a * MAPLE_PROJ_MAX - a * (a - 1) / 2 + (b - a);
}

// One wavefront per (cell, case): sum the cell's block partials, then (lane 0) the
// coefficients c minimising the energy norm of the error of sum_a c_a h_a, i.e. G c = b
// on the basis vectors that are not (numerically) combinations of earlier ones: a
// Cholesky factorisation in basis order that drops a vector whose new pivot is below
// tol times its own G_aa (a zero or nearly dependent vector gets c_a = 0).
__global__ __launch_bounds__(64) void k_proj_coef(const LevelCell* __restrict__ cells,
                                                  const int* __restrict__ blk_start, int ncase, int k,
                                                  const double* __restrict__ partial, double tol,
                                                  double* __restrict__ coef) {
  const int kk = blockIdx.x, j = blockIdx.y, ci = cells[kk].id;
  const int b0 = blk_start[kk], b1 = blk_start[kk + 1];
  double s[MAPLE_PROJ_NG];
#pragma unroll
  for (int t = 0; t < MAPLE_PROJ_NG; ++t) {
    double v = 0.0;
    for (int b = b0 + (int)threadIdx.x; b < b1; b += 64) v += partial[((long long)b * ncase + j) * MAPLE_PROJ_NG + t];
    s[t] = wave_sum(v);
  }
  if (threadIdx.x != 0) return;
  const double* rhs = s + MAPLE_PROJ_MAX * (MAPLE_PROJ_MAX + 1) / 2;
  double Lm[MAPLE_PROJ_MAX][MAPLE_PROJ_MAX];
  int keep[MAPLE_PROJ_MAX];
  int nk = 0;
  for (int a = 0; a < k; ++a) {
    const double gaa = s[proj_packed(a, a)];
    if (!(gaa > 0.0)) continue;
    double l[MAPLE_PROJ_MAX];
    double d = gaa;
    for (int t = 0; t < nk; ++t) {
      const int b = keep[t];
      double v = s[b < a ? proj_packed(b, a) : proj_packed(a, b)];
      for (int q = 0; q < t; ++q) v -= l[q] * Lm[t][q];
      l[t] = v / Lm[t][t];
      d -= l[t] * l[t];
    }
    if (!(d > tol * gaa)) continue;
    for (int t = 0; t < nk; ++t) Lm[nk][t] = l[t];
    Lm[nk][nk] = sqrt(d);
    keep[nk++] = a;
  }
  double y[MAPLE_PROJ_MAX], x[MAPLE_PROJ_MAX];
  for (int t = 0; t < nk; ++t) {
    double v = rhs[keep[t]];
    for (int q = 0; q < t; ++q) v -= Lm[t][q] * y[q];
    y[t] = v / Lm[t][t];
  }
  for (int t = nk - 1; t >= 0; --t) {
    double v = y[t];
    for (int q = t + 1; q < nk; ++q) v -= Lm[q][t] * x[q];
    x[t] = v / Lm[t][t];
  }
  double* out = coef + ((long long)ci * ncase + j) * MAPLE_PROJ_MAX;
  for (int a = 0; a < MAPLE_PROJ_MAX; ++a) out[a] = 0.0;
  for (int t = 0; t < nk; ++t) out[keep[t]] = x[t];
}

// us = sum_a c_a h_a, rs = f - sum_a c_a g_a, partial (rs, rs) per case.  us may be one
// of the basis vectors (each thread reads its node's values before writing them).  A
// vector left out (c_a = 0) is not read: the fields of a failed solve may hold NaN.
template <int ND>
__global__ __launch_bounds__(256) void k_proj_combine(Launch L, int ncase, int k, ProjBasis B,
                                                      const double* __restrict__ coef, const double* __restrict__ f,
                                                      double* us, double* __restrict__ rs,
                                                      double* __restrict__ partial) {
  int ci; long long gid = 0;
  const bool ok = vnode(L, ci, gid);
  double v[MAPLE_MAX_CASES];
#pragma unroll
  for (int j = 0; j < MAPLE_MAX_CASES; ++j) v[j] = 0.0;
  if (ok) {
    for (int j = 0; j < ncase; ++j) {
      const double* cj = coef + ((long long)ci * ncase + j) * MAPLE_PROJ_MAX;
      double c[MAPLE_PROJ_MAX];
#pragma unroll
      for (int a = 0; a < MAPLE_PROJ_MAX; ++a) c[a] = a < k ? cj[a] : 0.0;
#pragma unroll
      for (int i = 0; i < ND; ++i) {
        const long long idx = (long long)(j * ND + i) * L.Ntot + gid;
        double u = 0.0, r = f[idx];
#pragma unroll
        for (int a = 0; a < MAPLE_PROJ_MAX; ++a)
          if (a < k && c[a] != 0.0) {
            u += c[a] * B.h[a][idx];
            r -= c[a] * B.g[a][idx];
          }
        us[idx] = u;
        rs[idx] = r;
        v[j] += r * r;
      }
    }
  }
  block_sum<MAPLE_MAX_CASES>(v, partial + (long long)lbid() * MAPLE_MAX_CASES);
}

int proj_start(const Launch& L, int nd, int ncase, int k, const ProjBasis& B, const double* f, double* gpart,
               double* coef, double* us, double* rs, double* partial, hipStream_t s) {
  if (k < 1 || k > MAPLE_PROJ_MAX || ncase > MAPLE_MAX_CASES || (nd != 1 && nd != 3)) return -1;
  if (nd == 3) k_proj_gram<3><<<L.nblocks, 256, 0, s>>>(L, ncase, k, B, f, gpart);
  else k_proj_gram<1><<<L.nblocks, 256, 0, s>>>(L, ncase, k, B, f, gpart);
  k_proj_coef<<<dim3(L.ncells, ncase), 64, 0, s>>>(L.cells, L.blk_start, ncase, k, gpart, MAPLE_PROJ_TOL, coef);
  if (nd == 3) k_proj_combine<3><<<L.nblocks, 256, 0, s>>>(L, ncase, k, B, coef, f, us, rs, partial);
  else k_proj_combine<1><<<L.nblocks, 256, 0, s>>>(L, ncase, k, B, coef, f, us, rs, partial);
  return 0;
}
