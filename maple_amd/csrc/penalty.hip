// Additive-manufacturing penalties and their adjoints: the overhang penalty
// (printability_penalty_overhang, printability.py:91-134) and the trapped-void
// penalties of one cell (cavity_penalty_closed_void, printability.py:139-223) or of
// a finite array of tiled cells (cavity_penalty_periodic_array, printability.py:229-323),
// with the soft opening of printability.py:8-71.
//
// Gradients follow the subgradients the reference's autograd takes: a max-pool
// window sends its gradient to its first maximum in (x, y, z) raster order (a
// window whose first maximum is a padding position drops it), min-pool to its
// first minimum, minimum(a, b) halves a tie, relu passes where its argument is > 0.
//
// Overhang: the support recursion runs layer by layer in z, one workgroup per cell
// with the current support layer in LDS; the forward pass keeps every layer's
// support (one value per element) for the adjoint, which runs the layers back.
// Trapped voids: every step of the flood fill is one launch over all cells'
// supercells; it records per voxel one byte (the window's first maximum, 0..26, and
// the branch of the minimum), and the adjoint gathers each voxel's gradient from
// the neighbours whose window chose it (no atomics, deterministic).
#include "../../include/maple_hip.h"
#include "kernels.h"

namespace {

__device__ __forceinline__ int wrapm(int i, int n) { return ((i % n) + n) % n; }

__device__ __forceinline__ double sigm(double t) { return 1.0 / (1.0 + exp(-t)); }

// voxel (x, y, z) of the supercell holding logical block b of a launch
__device__ __forceinline__ bool vox(const Launch& L, int& k, LevelCell& c, int& x, int& y, int& z, long long& n) {
  k = find_cell(L.blk_start, L.ncells, lbid());
  c = L.cells[k];
  n = (long long)(lbid() - L.blk_start[k]) * 256 + threadIdx.x;
  if (n >= (long long)c.nx * c.ny * c.nz) return false;
  z = (int)(n % c.nz);
  const long long t = n / c.nz;
  y = (int)(t % c.ny);
  x = (int)(t / c.ny);
  return true;
}

}  // namespace

// ---- overhang --------------------------------------------------------------------

// One workgroup per cell (r <= 64): sup_0 = rho[:, :, 0]; for z >= 1: s = maxpool2d(sup),
// P += mean(relu(rho_z - s)), sup_z = minimum(rho_z, s).  tape[(x r + y) r + z] = sup_z.
// The current layer sits in LDS; each thread forms its columns' next values in registers.
__global__ __launch_bounds__(1024) void k_overhang_fwd(const LevelCell* __restrict__ cells,
                                                       const double* __restrict__ rho, int rad, int periodic,
                                                       double* __restrict__ tape, double* __restrict__ P) {
  extern __shared__ double sh[];
  constexpr int PER = 4;   // columns per thread (r <= 64)
  const LevelCell c = cells[blockIdx.x];
  const int r = c.nx, n2 = r * r;
  double* A = sh;
  const double* rg = rho + c.off;
  double* tp = tape + c.off;
  for (int i = threadIdx.x; i < n2; i += blockDim.x) {
    A[i] = rg[(long long)i * r];
    tp[(long long)i * r] = A[i];
  }
  double acc = 0.0;
  for (int z = 1; z < r; ++z) {
    __syncthreads();
    double nv[PER];
    for (int q = 0; q < PER; ++q) {
      const int i = threadIdx.x + q * blockDim.x;
      nv[q] = 0.0;
      if (i >= n2) break;
      const int x = i / r, y = i - (i / r) * r;
      double s = -INFINITY;
      for (int dx = -rad; dx <= rad; ++dx)
        for (int dy = -rad; dy <= rad; ++dy) {
          const int xx = x + dx, yy = y + dy;
          double v;
          if (periodic) v = A[wrapm(xx, r) * r + wrapm(yy, r)];
          else v = (xx < 0 || xx >= r || yy < 0 || yy >= r) ? 0.0 : A[xx * r + yy];
          if (v > s) s = v;
        }
      const double rz = rg[(long long)i * r + z];
      acc += fmax(rz - s, 0.0);
      nv[q] = rz < s ? rz : s;
      tp[(long long)i * r + z] = nv[q];
    }
    __syncthreads();
    for (int q = 0; q < PER; ++q) {
      const int i = threadIdx.x + q * blockDim.x;
      if (i >= n2) break;
      A[i] = nv[q];
    }
  }
  double v[1] = {acc};
  __syncthreads();
  block_sum<1>(v, P + blockIdx.x);
  if (threadIdx.x == 0) P[blockIdx.x] = P[blockIdx.x] / (double)n2 / (double)(r > 1 ? r - 1 : 1);
}

// Adjoint: dP/drho of every element of the cell (the forward's tape read back).
__global__ __launch_bounds__(1024) void k_overhang_bwd(const LevelCell* __restrict__ cells,
                                                       const double* __restrict__ rho, int rad, int periodic,
                                                       const double* __restrict__ tape, double* __restrict__ grad) {
  extern __shared__ double sh[];
  const LevelCell c = cells[blockIdx.x];
  const int r = c.nx, n2 = r * r, K = 2 * rad + 1;
  double* gs = sh;                               // gradient of each window's maximum
  unsigned char* arg = (unsigned char*)(sh + n2);  // each window's first maximum (255: padding)
  const double* rg = rho + c.off;
  const double* tp = tape + c.off;
  double* gr = grad + c.off;
  const double cw = 1.0 / (double)(r > 1 ? r - 1 : 1) / (double)n2;
  constexpr int PER = 4;                         // columns per thread (r <= 64)
  double gsup[PER];
  for (int q = 0; q < PER; ++q) gsup[q] = 0.0;
  for (int z = r - 1; z >= 1; --z) {
    for (int q = 0; q < PER; ++q) {
      const int i = threadIdx.x + q * blockDim.x;
      if (i >= n2) break;
      const int x = i / r, y = i - (i / r) * r;
      double s = -INFINITY;
      int a = 255;
      for (int dx = -rad; dx <= rad; ++dx)
        for (int dy = -rad; dy <= rad; ++dy) {
          const int xx = x + dx, yy = y + dy;
          const bool pad = !periodic && (xx < 0 || xx >= r || yy < 0 || yy >= r);
          const double v = pad ? 0.0 : tp[((long long)wrapm(xx, r) * r + wrapm(yy, r)) * r + (z - 1)];
          if (v > s) {
            s = v;
            a = pad ? 255 : (dx + rad) * K + (dy + rad);
          }
        }
      const double rz = rg[(long long)i * r + z];
      const double pos = rz - s > 0.0 ? cw : 0.0;
      const double wr = rz < s ? 1.0 : (rz == s ? 0.5 : 0.0);
      gr[(long long)i * r + z] = pos + gsup[q] * wr;
      gs[i] = -pos + gsup[q] * (1.0 - wr);
      arg[i] = (unsigned char)a;
    }
    __syncthreads();
    for (int q = 0; q < PER; ++q) {
      const int i = threadIdx.x + q * blockDim.x;
      if (i >= n2) break;
      const int x = i / r, y = i - (i / r) * r;
      double g = 0.0;
      for (int dx = -rad; dx <= rad; ++dx)
        for (int dy = -rad; dy <= rad; ++dy) {
          int vx = x - dx, vy = y - dy;   // the window of v = u - off sees u at offset off
          if (periodic) {
            vx = wrapm(vx, r);
            vy = wrapm(vy, r);
          } else if (vx < 0 || vx >= r || vy < 0 || vy >= r) {
            continue;
          }
          const int v = vx * r + vy;
          if (arg[v] == (dx + rad) * K + (dy + rad)) g += gs[v];
        }
      gsup[q] = g;
    }
    __syncthreads();
  }
  for (int q = 0; q < PER; ++q) {
    const int i = threadIdx.x + q * blockDim.x;
    if (i >= n2) break;
    gr[(long long)i * r] = gsup[q];
  }
}

void overhang_penalty(const LevelCell* cells, int ncells, int rmax, const double* rho, int rad, int periodic,
                      double* tape, double* P, double* grad, hipStream_t s) {
  const size_t lds = (size_t)rmax * rmax * sizeof(double);
  k_overhang_fwd<<<ncells, 1024, lds, s>>>(cells, rho, rad, periodic, tape, P);
  if (grad) {
    const size_t lds2 = (size_t)rmax * rmax * (sizeof(double) + 1) + 16;
    k_overhang_bwd<<<ncells, 1024, lds2, s>>>(cells, rho, rad, periodic, tape, grad);
  }
}

// ---- trapped voids -----------------------------------------------------------------

// void = sigmoid(beta (thr - rho)) of the unit cell, tiled over the supercell
__global__ __launch_bounds__(256) void k_cav_void(Launch Ls, const LevelCell* __restrict__ uc,
                                                  const double* __restrict__ rho, double beta, double thr,
                                                  double* __restrict__ voidv) {
  int k, x, y, z; LevelCell c; long long n;
  if (!vox(Ls, k, c, x, y, z, n)) return;
  const LevelCell u = uc[c.id];
  const double rv = rho[u.off + ((long long)(x % u.nx) * u.ny + (y % u.ny)) * u.nz + (z % u.nz)];
  voidv[c.off + n] = sigm(beta * (thr - rv));
}

// out = window max (MIN: min) over (2 rad + 1)^3, first in raster order; padding with
// pad (not periodic) or the wrapped voxel; arg = the window offset index (-1: padding)
template <bool MIN>
__global__ __launch_bounds__(256) void k_cav_pool(Launch Ls, int rad, int periodic, double pad,
                                                  const double* __restrict__ in, double* __restrict__ out,
                                                  short* __restrict__ arg) {
  int k, x, y, z; LevelCell c; long long n;
  if (!vox(Ls, k, c, x, y, z, n)) return;
  const double* a = in + c.off;
  const int K = 2 * rad + 1;
  double best = MIN ? INFINITY : -INFINITY;
  int ba = -1;
  for (int dx = -rad; dx <= rad; ++dx)
    for (int dy = -rad; dy <= rad; ++dy)
      for (int dz = -rad; dz <= rad; ++dz) {
        int xx = x + dx, yy = y + dy, zz = z + dz;
        bool pd = false;
        if (periodic) {
          xx = wrapm(xx, c.nx); yy = wrapm(yy, c.ny); zz = wrapm(zz, c.nz);
        } else {
          pd = xx < 0 || xx >= c.nx || yy < 0 || yy >= c.ny || zz < 0 || zz >= c.nz;
        }
        const double v = pd ? pad : a[((long long)xx * c.ny + yy) * c.nz + zz];
        if (MIN ? v < best : v > best) {
          best = v;
          ba = pd ? -1 : ((dx + rad) * K + (dy + rad)) * K + (dz + rad);
        }
      }
  out[c.off + n] = best;
  arg[c.off + n] = (short)ba;
}

__device__ __forceinline__ double seed_of(const LevelCell& c, int x, int y, int z, int faces) {
  const bool fx = x == 0 || x == c.nx - 1, fy = y == 0 || y == c.ny - 1, fz = z == 0 || z == c.nz - 1;
  switch (faces) {
    case MAPLE_SEED_ALL: return (fx || fy || fz) ? 1.0 : 0.0;
    case MAPLE_SEED_X: return fx ? 1.0 : 0.0;
    case MAPLE_SEED_Y: return fy ? 1.0 : 0.0;
    default: return fz ? 1.0 : 0.0;   // z and xy (drainage along z)
  }
}

__global__ __launch_bounds__(256) void k_cav_seed(Launch Ls, int faces, const double* __restrict__ vc,
                                                  double* __restrict__ reach) {
  int k, x, y, z; LevelCell c; long long n;
  if (!vox(Ls, k, c, x, y, z, n)) return;
  reach[c.off + n] = vc[c.off + n] * seed_of(c, x, y, z, faces);
}

// one flood step (rad = 1): next = minimum(vc, maxpool(prev)); tape byte = first
// maximum's offset (0..26; 31: padding) | branch << 5 (0: vc smaller, 1: dil smaller, 2: tie)
__global__ __launch_bounds__(256) void k_cav_flood(Launch Ls, int periodic, const double* __restrict__ vc,
                                                   const double* __restrict__ prev, double* __restrict__ next,
                                                   unsigned char* __restrict__ tape) {
  int k, x, y, z; LevelCell c; long long n;
  if (!vox(Ls, k, c, x, y, z, n)) return;
  const double* a = prev + c.off;
  double best = -INFINITY;
  int ba = 31;
#pragma unroll
  for (int dx = -1; dx <= 1; ++dx)
#pragma unroll
    for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
      for (int dz = -1; dz <= 1; ++dz) {
        int xx = x + dx, yy = y + dy, zz = z + dz;
        bool pd = false;
        if (periodic) {
          xx = wrap(xx, c.nx); yy = wrap(yy, c.ny); zz = wrap(zz, c.nz);
        } else {
          pd = xx < 0 || xx >= c.nx || yy < 0 || yy >= c.ny || zz < 0 || zz >= c.nz;
        }
        const double v = pd ? 0.0 : a[((long long)xx * c.ny + yy) * c.nz + zz];
        if (v > best) {
          best = v;
          ba = pd ? 31 : (dx + 1) * 9 + (dy + 1) * 3 + (dz + 1);
        }
      }
  const double w = vc[c.off + n];
  const int br = w < best ? 0 : (best < w ? 1 : 2);
  next[c.off + n] = br == 0 ? w : best;
  tape[c.off + n] = (unsigned char)(ba | (br << 5));
}

// trapped = relu(void - reach) over the counted region: per-block partial sums; the
// gradient seeds gvoid = [trapped > 0] / count, greach = -gvoid
__global__ __launch_bounds__(256) void k_cav_trapped(Launch Ls, int focus, const int* __restrict__ reg,
                                                     const double* __restrict__ voidv, const double* __restrict__ reach,
                                                     double* __restrict__ partial, double* __restrict__ gvoid,
                                                     double* __restrict__ greach, double* __restrict__ gvc) {
  int k, x, y, z; LevelCell c; long long n = 0;
  const bool ok = vox(Ls, k, c, x, y, z, n);
  double t = 0.0;
  if (ok) {
    // counted region: the centre tile (focus) or the whole supercell; reg[3 cell + d] = tile side
    bool in = true;
    long long cnt = (long long)c.nx * c.ny * c.nz;
    if (focus) {
      const int rx = reg[3 * c.id], ry = reg[3 * c.id + 1], rz = reg[3 * c.id + 2];
      const int tx = c.nx / rx, ty = c.ny / ry, tz = c.nz / rz;
      in = x / rx == tx / 2 && y / ry == ty / 2 && z / rz == tz / 2;
      cnt = (long long)rx * ry * rz;
    }
    const long long i = c.off + n;
    const double d = voidv[i] - reach[i];
    const double g = (in && d > 0.0) ? 1.0 / (double)cnt : 0.0;
    if (in) t = fmax(d, 0.0);
    gvoid[i] = g;
    greach[i] = -g;
    gvc[i] = 0.0;
  }
  double v[1] = {t};
  block_sum<1>(v, partial + lbid());
}

// P[cell] = sum of the cell's block partials / count
__global__ __launch_bounds__(64) void k_cav_reduce(Launch Ls, int focus, const int* __restrict__ reg,
                                                   const double* __restrict__ partial, double* __restrict__ P) {
  const int k = blockIdx.x;
  const LevelCell c = Ls.cells[k];
  double s = 0.0;
  for (int b = Ls.blk_start[k] + (int)threadIdx.x; b < Ls.blk_start[k + 1]; b += 64) s += partial[b];
  s = wave_sum(s);
  const long long cnt = focus ? (long long)reg[3 * c.id] * reg[3 * c.id + 1] * reg[3 * c.id + 2]
                              : (long long)c.nx * c.ny * c.nz;
  if (threadIdx.x == 0) P[c.id] = s / (double)cnt;
}

// adjoint of one flood step: gprev(u) = sum over the windows v that chose u of gcur(v) wdil(v);
// gvc(u) += gcur(u) wvc(u)
__global__ __launch_bounds__(256) void k_cav_flood_bwd(Launch Ls, int periodic, const unsigned char* __restrict__ tape,
                                                       const double* __restrict__ gcur, double* __restrict__ gprev,
                                                       double* __restrict__ gvc) {
  int k, x, y, z; LevelCell c; long long n;
  if (!vox(Ls, k, c, x, y, z, n)) return;
  const unsigned char* tp = tape + c.off;
  const double* g = gcur + c.off;
  double acc = 0.0;
#pragma unroll
  for (int dx = -1; dx <= 1; ++dx)
#pragma unroll
    for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
      for (int dz = -1; dz <= 1; ++dz) {
        int vx = x - dx, vy = y - dy, vz = z - dz;
        if (periodic) {
          vx = wrap(vx, c.nx); vy = wrap(vy, c.ny); vz = wrap(vz, c.nz);
        } else if (vx < 0 || vx >= c.nx || vy < 0 || vy >= c.ny || vz < 0 || vz >= c.nz) {
          continue;
        }
        const long long v = ((long long)vx * c.ny + vy) * c.nz + vz;
        const int b = tp[v];
        if ((b & 31) == (dx + 1) * 9 + (dy + 1) * 3 + (dz + 1)) {
          const int br = b >> 5;
          acc += g[v] * (br == 1 ? 1.0 : (br == 2 ? 0.5 : 0.0));
        }
      }
  const long long i = c.off + n;
  gprev[i] = acc;
  const int br = tape[i] >> 5;
  gvc[i] += gcur[i] * (br == 0 ? 1.0 : (br == 2 ? 0.5 : 0.0));
}

// gvc += greach0 * seed
__global__ __launch_bounds__(256) void k_cav_seed_bwd(Launch Ls, int faces, const double* __restrict__ greach,
                                                      double* __restrict__ gvc) {
  int k, x, y, z; LevelCell c; long long n;
  if (!vox(Ls, k, c, x, y, z, n)) return;
  gvc[c.off + n] += greach[c.off + n] * seed_of(c, x, y, z, faces);
}

// adjoint of a window pool: out(u) (+)= sum over the windows v that chose u of gin(v)
__global__ __launch_bounds__(256) void k_cav_pool_bwd(Launch Ls, int rad, int periodic, const short* __restrict__ arg,
                                                      const double* __restrict__ gin, double* __restrict__ out,
                                                      int accumulate) {
  int k, x, y, z; LevelCell c; long long n;
  if (!vox(Ls, k, c, x, y, z, n)) return;
  const int K = 2 * rad + 1;
  const short* ag = arg + c.off;
  const double* g = gin + c.off;
  double acc = 0.0;
  for (int dx = -rad; dx <= rad; ++dx)
    for (int dy = -rad; dy <= rad; ++dy)
      for (int dz = -rad; dz <= rad; ++dz) {
        int vx = x - dx, vy = y - dy, vz = z - dz;
        if (periodic) {
          vx = wrapm(vx, c.nx); vy = wrapm(vy, c.ny); vz = wrapm(vz, c.nz);
        } else if (vx < 0 || vx >= c.nx || vy < 0 || vy >= c.ny || vz < 0 || vz >= c.nz) {
          continue;
        }
        const long long v = ((long long)vx * c.ny + vy) * c.nz + vz;
        if (ag[v] == ((dx + rad) * K + (dy + rad)) * K + (dz + rad)) acc += g[v];
      }
  if (accumulate) out[c.off + n] += acc;
  else out[c.off + n] = acc;
}

// gvoid += gvc (no opening)
__global__ __launch_bounds__(256) void k_cav_add(long long n, const double* __restrict__ a, double* __restrict__ b) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) b[i] += a[i];
}

// dP/drho of each unit-cell element: the tiles' gvoid summed, times dvoid/drho
__global__ __launch_bounds__(256) void k_cav_grad_uc(Launch Lu, const LevelCell* __restrict__ sc,
                                                     const double* __restrict__ rho, double beta, double thr,
                                                     const double* __restrict__ gvoid, double* __restrict__ grad) {
  int k, x, y, z; LevelCell u; long long n;
  if (!vox(Lu, k, u, x, y, z, n)) return;
  const LevelCell s = sc[u.id];
  const int tx = s.nx / u.nx, ty = s.ny / u.ny, tz = s.nz / u.nz;
  double g = 0.0;
  for (int a = 0; a < tx; ++a)
    for (int b = 0; b < ty; ++b)
      for (int cc = 0; cc < tz; ++cc)
        g += gvoid[s.off + ((long long)(a * u.nx + x) * s.ny + (b * u.ny + y)) * s.nz + (cc * u.nz + z)];
  const double v = sigm(beta * (thr - rho[u.off + n]));
  grad[u.off + n] = g * (-beta) * v * (1.0 - v);
}

int cavity_penalty(const CavityBuffers& B, const LevelCell* uc, const Launch& Lu, const double* rho, double* P,
                   double* grad, hipStream_t s) {
  const Launch& Ls = B.L;
  const long long G = Ls.Ntot;
  const int nb = Ls.nblocks;
  k_cav_void<<<nb, 256, 0, s>>>(Ls, uc, rho, B.beta, B.thr, B.voidv);
  const double* vc = B.voidv;
  if (B.mcr > 0) {
    k_cav_pool<true><<<nb, 256, 0, s>>>(Ls, B.mcr, B.periodic, 1.0, B.voidv, B.er, B.argE);
    k_cav_pool<false><<<nb, 256, 0, s>>>(Ls, B.mcr, B.periodic, 0.0, B.er, B.vc, B.argD);
    vc = B.vc;
  }
  k_cav_seed<<<nb, 256, 0, s>>>(Ls, B.faces, vc, B.reach[0]);
  const int T = B.rad > 0 ? B.nsteps : 0;
  for (int t = 0; t < T; ++t)
    k_cav_flood<<<nb, 256, 0, s>>>(Ls, B.periodic, vc, B.reach[t & 1], B.reach[(t + 1) & 1], B.tape + (long long)t * G);
  const double* reach = B.reach[T & 1];
  k_cav_trapped<<<nb, 256, 0, s>>>(Ls, B.focus, B.region, B.voidv, reach, B.partial, B.gvoid, B.g[0], B.gvc);
  k_cav_reduce<<<Ls.ncells, 64, 0, s>>>(Ls, B.focus, B.region, B.partial, P);
  if (!grad) return 0;
  for (int t = T - 1; t >= 0; --t)
    k_cav_flood_bwd<<<nb, 256, 0, s>>>(Ls, B.periodic, B.tape + (long long)t * G, B.g[(T - 1 - t) & 1],
                                        B.g[(T - t) & 1], B.gvc);
  k_cav_seed_bwd<<<nb, 256, 0, s>>>(Ls, B.faces, B.g[T & 1], B.gvc);
  if (B.mcr > 0) {
    k_cav_pool_bwd<<<nb, 256, 0, s>>>(Ls, B.mcr, B.periodic, B.argD, B.gvc, B.er, 0);   // er now holds g_er
    k_cav_pool_bwd<<<nb, 256, 0, s>>>(Ls, B.mcr, B.periodic, B.argE, B.er, B.gvoid, 1);
  } else {
    k_cav_add<<<(unsigned)((G + 255) / 256), 256, 0, s>>>(G, B.gvc, B.gvoid);
  }
  k_cav_grad_uc<<<Lu.nblocks, 256, 0, s>>>(Lu, Ls.cells, rho, B.beta, B.thr, B.gvoid, grad);
  return 0;
}
