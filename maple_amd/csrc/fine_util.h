// Device helpers shared by the level-0 operator kernels (fine*.hip): the lane
// map of the element-row kernels, cross-lane moves, per-launch work counting and
// the node of a thread in cell-aligned 256-node blocks.
#pragma once

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

namespace {

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

}  // namespace

// Cross-lane moves of a wavefront that holds one periodic z-pencil of 64 nodes:
// the value of the lane at z-1 / z+1 (DPP wave_ror:1 / wave_rol:1, periodic like the pencil).
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
