// Shared definitions of the MI355X (gfx950) cell-solver library.
//
// Grid layout (the reference's, main_IHD_v3.py:175-197): node id and element
// id are both (x*ny + y)*nz + z with z contiguous; element (x,y,z) owns the
// nodes (x,y,z)+HEX8[k] with a periodic wrap.  Device vectors of one level are
// stored plane by plane: value (case j, component c, node n of cell i) sits at
//   v[(j*ND + c)*Ntot + off_i + n]
// where Ntot is the total node count of that level over every cell of the
// batch and off_i the first node of cell i.  A single cell is a batch of one.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#define MAPLE_MAX_CASES 6

// One cell at one multigrid level.
struct LevelCell {
  int nx, ny, nz;
  int id;          // index of the cell in the batch (per-cell arrays); a launch may list a subset of cells
  long long off;   // first node (= first element) of this cell in the level arrays
};

// Hex8 corner offsets in the reference's order (_utils.py:62-64).
__host__ __device__ constexpr int hex_dx(int k) { return (k == 1 || k == 2 || k == 5 || k == 6) ? 1 : 0; }
__host__ __device__ constexpr int hex_dy(int k) { return (k == 2 || k == 3 || k == 6 || k == 7) ? 1 : 0; }
__host__ __device__ constexpr int hex_dz(int k) { return k >= 4 ? 1 : 0; }
// Corner index of offset (a,b,c) in {0,1}^3.
__host__ __device__ constexpr int hex_corner(int a, int b, int c) {
  return c * 4 + (b ? (a ? 2 : 3) : (a ? 1 : 0));
}

__device__ __forceinline__ int wrap(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// XCD-aware block order.  MI355X dispatches consecutive workgroups round-robin
// over its 8 XCDs, each with its own L2.  Logical block b is given to the
// workgroup that XCD (b / chunk) runs, so every XCD works on one contiguous
// run of logical blocks (a slab of neighbouring pencils that share halos in
// its L2) instead of every eighth block of the whole domain.  Bijective on
// [0, nb).
#define MAPLE_NUM_XCD 8
__host__ __device__ inline int xcd_swizzle(int bid, int nb) {
  const int q = nb / MAPLE_NUM_XCD, r = nb % MAPLE_NUM_XCD;
  const int xcd = bid % MAPLE_NUM_XCD, k = bid / MAPLE_NUM_XCD;
  return xcd * q + (xcd < r ? xcd : r) + k;
}
__device__ __forceinline__ int lbid() { return xcd_swizzle((int)blockIdx.x, (int)gridDim.x); }

// Block -> cell lookup: blk_start has ncells+1 entries, the i-th listed cell owns
// blocks [blk_start[i], blk_start[i+1]).  A launch may list only the cells still
// iterating; per-cell arrays are indexed by LevelCell::id, never by the position.
__device__ __forceinline__ int find_cell(const int* __restrict__ blk_start, int ncells, int b) {
  int lo = 0, hi = ncells - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (blk_start[mid] <= b) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Block-wide sum of NV values per thread into out[0..NV) (valid in thread 0).
// blockDim.x must be a multiple of 64 and at most 1024.
template <int NV, typename T>
__device__ __forceinline__ void block_sum(T (&v)[NV], T* __restrict__ out) {
  __shared__ T red[16][NV];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    T s = wave_sum(v[i]);
    if (lane == 0) red[w][i] = s;
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    T s = 0;
    for (int k = 0; k < nw; ++k) s += red[k][threadIdx.x];
    out[threadIdx.x] = s;
  }
  __syncthreads();
}
