// The OC update without a context and without any allocation.
//
// maple_oc_update runs on a context whose creation allocated its scratch (the
// per-cell bisection state, the per-block partial sums, the a = x sqrt(dJ) array
// and the launch tables).  Callers that cannot allocate device memory -- code under
// HIP graph capture, tight serving loops, harnesses that forbid hipMalloc -- call
// maple_oc_update_noalloc instead: the same kernels (the a array is formed on the
// fly from dJ, to the same bits), their scratch in one file-scope device array.
//
// Each device holds its own copy of `scratch`; its address on a device comes from
// hipGetSymbolAddress, taken once per device with that device current, and only
// that address is handed to kernels.  Every call on a device shares the array and
// finds it as the last call left it: the launch tables and bisection state are
// rewritten at the start of every call and each partial sum is written before it
// is read, so nothing has to start at zero.  Calls that overlap on one device (two
// streams, two host threads) race on the array.
#include "../../include/maple_hip.h"

#include <cstddef>
#include <mutex>
#include <string>

#include "kernels.h"

namespace {

constexpr int kMaxCells = MAPLE_NOALLOC_MAX_CELLS;
constexpr long long kMaxElems = MAPLE_NOALLOC_MAX_ELEMENTS;
// 1024-element OC blocks (each cell rounds up once) and 256-element node blocks
constexpr long long kMaxOcBlocks = kMaxElems / 1024 + kMaxCells;
constexpr int kMaxDevices = 64;

struct Scratch {
  OCState st[kMaxCells];
  LevelCell cells[kMaxCells];
  int blk_node[kMaxCells + 1];
  int blk_oc[kMaxCells + 1];
  double partial[kMaxOcBlocks * 64];
};

// the per-device scratch; host code never uses this name, only its device address
static __device__ __align__(256) unsigned char scratch[sizeof(Scratch)];

// Everything a call needs about its cells, passed by value (no host-to-device copy).
struct Table {
  int ncells;
  int dims[kMaxCells][3];
  double volfrac[kMaxCells], move[kMaxCells], xmin[kMaxCells];
};

// One block: the cell and launch tables and the per-cell OC parameters.
__global__ __launch_bounds__(64) void k_noalloc_setup(Table t, Scratch* __restrict__ S) {
  if (threadIdx.x == 0) {
    long long off = 0;
    int bn = 0, bo = 0;
    for (int i = 0; i < t.ncells; ++i) {
      LevelCell c;
      c.nx = t.dims[i][0];
      c.ny = t.dims[i][1];
      c.nz = t.dims[i][2];
      c.id = i;
      c.off = off;
      S->cells[i] = c;
      S->blk_node[i] = bn;
      S->blk_oc[i] = bo;
      const long long n = (long long)c.nx * c.ny * c.nz;
      off += n;
      bn += (int)((n + 255) / 256);
      bo += (int)((n + 1023) / 1024);
    }
    S->blk_node[t.ncells] = bn;
    S->blk_oc[t.ncells] = bo;
  }
  for (int i = threadIdx.x; i < t.ncells; i += blockDim.x) {
    S->st[i].volfrac = t.volfrac[i];
    S->st[i].move = t.move[i];
    S->st[i].xmin = t.xmin[i];
  }
}

std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

// Device address of `scratch` on `device` (that device current), taken once.
int scratch_ptr(int device, Scratch** out) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      return maple_record_error(MAPLE_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    }
    g_ptr[device] = p;
  }
  *out = (Scratch*)g_ptr[device];
  return 0;
}

}  // namespace

extern "C" int maple_oc_update_noalloc(int device, int ncells, const int* dims_host, const double* x_dev,
                                       const double* dJdx_dev, const double* volfrac_host, const double* move_host,
                                       const double* xmin_host, int nbisect, double* xnew_dev, void* stream) {
  if (!dims_host || !x_dev || !dJdx_dev || !volfrac_host || !move_host || !xmin_host || !xnew_dev)
    return maple_record_error(MAPLE_EINVAL, "null argument");
  if (ncells < 1 || ncells > kMaxCells)
    return maple_record_error(MAPLE_EINVAL, "ncells must lie in [1, " + std::to_string(kMaxCells) +
                                                "] (the allocation-free OC update's scratch limit)");
  if (nbisect < 1 || nbisect > 200) return maple_record_error(MAPLE_EINVAL, "nbisect must lie in [1, 200]");
  Table t;
  t.ncells = ncells;
  long long total = 0;
  int nb_node = 0, nb_oc = 0;
  for (int i = 0; i < ncells; ++i) {
    for (int d = 0; d < 3; ++d) {
      const int v = dims_host[3 * i + d];
      if (v <= 0 || v > 4096) return maple_record_error(MAPLE_EINVAL, "cell dimensions must lie in [1, 4096]");
      t.dims[i][d] = v;
    }
    const long long n = (long long)t.dims[i][0] * t.dims[i][1] * t.dims[i][2];
    total += n;
    if (total > kMaxElems)
      return maple_record_error(MAPLE_EINVAL, "more than " + std::to_string(kMaxElems) +
                                                  " elements (the allocation-free OC update's scratch limit)");
    nb_node += (int)((n + 255) / 256);
    nb_oc += (int)((n + 1023) / 1024);
    t.volfrac[i] = volfrac_host[i];
    t.move[i] = move_host[i];
    t.xmin[i] = xmin_host[i];
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    (void)hipGetLastError();
    return maple_record_error(MAPLE_EDEVICE, "no HIP device available");
  }
  if (device < 0 || device >= ndev || device >= kMaxDevices) return maple_record_error(MAPLE_EINVAL, "device index out of range");
  if (hipSetDevice(device) != hipSuccess) {
    (void)hipGetLastError();
    return maple_record_error(MAPLE_EDEVICE, "hipSetDevice failed");
  }
  Scratch* S = nullptr;
  if (int rc = scratch_ptr(device, &S)) return rc;
  // device addresses of the scratch members (S is a device address: never dereferenced here)
  char* base = (char*)S;
  OCState* st = (OCState*)(base + offsetof(Scratch, st));
  LevelCell* cells = (LevelCell*)(base + offsetof(Scratch, cells));
  int* blk_node = (int*)(base + offsetof(Scratch, blk_node));
  int* blk_oc = (int*)(base + offsetof(Scratch, blk_oc));
  double* partial = (double*)(base + offsetof(Scratch, partial));
  hipStream_t s = (hipStream_t)stream;
  k_noalloc_setup<<<1, 64, 0, s>>>(t, S);
  Launch Ln, Lo;
  Ln.cells = Lo.cells = cells;
  Ln.ncells = Lo.ncells = ncells;
  Ln.Ntot = Lo.Ntot = total;
  Ln.blk_start = blk_node;
  Ln.nblocks = nb_node;
  Lo.blk_start = blk_oc;
  Lo.nblocks = nb_oc;
  oc_start(ncells, st, s);
  for (int done = 0; done < nbisect; done += 6) {
    const int depth = nbisect - done < 6 ? nbisect - done : 6;
    oc_pass_dj(Lo, st, x_dev, dJdx_dev, partial, s);
    oc_walk(blk_oc, cells, ncells, depth, partial, st, s);
  }
  oc_final_dj(Ln, st, x_dev, dJdx_dev, xnew_dev, s);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return maple_record_error(MAPLE_EDEVICE, std::string("OC launch: ") + hipGetErrorString(e));
  return 0;
}
