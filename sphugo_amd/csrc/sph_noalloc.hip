// Neighbour readback for callers that cannot allocate device memory.
//
// sph_read_neighbours stages the rows in device buffers it allocates on its
// first call on a handle.  sph_read_neighbours_noalloc runs the same kernel
// (k_read_nbrs) into `scratch`, a file-scope device array of fixed size, and
// calls no allocating routine: for code that runs where hipMalloc is not
// allowed (under stream capture, in a serving loop that must not allocate,
// in harnesses that forbid it).
//
// `scratch` holds kNoAllocEntries neighbour entries (u32 ID + FP64 squared
// distance, 12 MiB); every GPU the library runs on keeps its own copy until
// the process ends.  In host code the array's name is only a stand-in for
// the symbol, so its device address is looked up once per device
// (hipGetSymbolAddress, with that device current) and that address is what
// the kernel and the copies use.  The kernel writes every entry of the
// np * kout rows (entries past a simulation's k get a sentinel), so nothing
// has to be cleared between calls.  Calls that overlap on one GPU share the
// array and race; one call at a time per GPU.
#include <mutex>

#include "sph_sim.hpp"

using namespace sph;

namespace {

constexpr int64_t kNoAllocEntries = 1 << 20;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char
    scratch[kNoAllocEntries * (sizeof(uint32_t) + sizeof(double))];

std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on `device` (looked up on the first call per
// device, with that device current).
int scratch_address(int device, void** out) {
  if (device < 0 || device >= kMaxDevices) return fail(SPH_EINVAL, "device %d out of range", device);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    int cur = 0;
    HIPCHK(hipGetDevice(&cur));
    if (cur != device) HIPCHK(hipSetDevice(device));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) HIPCHK(hipSetDevice(cur));
    if (e != hipSuccess || !p)
      return fail(SPH_EDEVICE, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch_ptr[device] = p;
  }
  *out = g_scratch_ptr[device];
  return 0;
}

}  // namespace

extern "C" int sph_read_neighbours_noalloc(sph_sim* s, uint32_t* ids, double* d2) {
  if (!s || !ids) return fail(SPH_EINVAL, "null argument");
  const int64_t total = s->np * s->pb.kout;
  if (total > kNoAllocEntries)
    return fail(SPH_EINVAL,
                "%lld neighbour entries (particles x k) exceed the %lld of the fixed readback "
                "buffer; use sph_read_neighbours",
                (long long)total, (long long)kNoAllocEntries);
  void* base = nullptr;
  int rc = scratch_address(s->device, &base);
  if (rc) return rc;
  uint32_t* d_ids = (uint32_t*)base;
  double* d_d2 = (double*)((unsigned char*)base + kNoAllocEntries * sizeof(uint32_t));
  s->bind();
  rc = enqueue_read_neighbours(s->pb, d_ids, d_d2, nullptr, nullptr, s->st);
  if (rc) return fail(SPH_EDEVICE, "neighbour readback launch failed");
  HIPCHK(hipMemcpyAsync(ids, d_ids, total * sizeof(uint32_t), hipMemcpyDeviceToHost, s->st));
  if (d2) HIPCHK(hipMemcpyAsync(d2, d_d2, total * sizeof(double), hipMemcpyDeviceToHost, s->st));
  HIPCHK(hipStreamSynchronize(s->st));
  return 0;
}
