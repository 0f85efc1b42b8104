// The inclusive scan for callers that cannot allocate device memory (code
// under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc): tcs_scan_noalloc runs the same kernels as tcs_scan, with the
// look-back workspace in a file-scope device array instead of a caller's
// buffer.
//
// The array lives once per GPU (each device holds its own copy at its own
// address) for the life of the process.  Host code never uses its name as
// an address: the first call on a device, with that device current, asks
// the runtime for that device's address (hipGetSymbolAddress) and keeps it;
// kernels and the fill kernel only ever get that pointer.  The same first
// call zeroes the array on the caller's stream, which the workspace contract
// asks of a workspace before its first use, and every
// TCS_WORKSPACE_MAX_CALLS calls on a device zero it again so that launch
// tags never repeat.
//
// Every call on one GPU shares that GPU's array: calls that overlap on one
// GPU (two streams, or two host threads without ordering) race on it.
// Each call leaves the array ready for the next one; nothing of one call's
// input or output stays in it.
#include <mutex>

#include "scan_kernel.hpp"

namespace tcs {
namespace {

__device__ __align__(256) uint64_t scratch[TCS_NOALLOC_SCRATCH_BYTES / 8];

struct DeviceScratch {
  void* ptr = nullptr;   // this device's address of `scratch`
  uint64_t calls = 0;    // calls since the last zeroing
};

std::mutex g_scratch_mu;
DeviceScratch g_scratch[kMaxCachedDevices];

// This device's scratch, zeroed on `st` first when it is new or due for a
// re-zeroing; nullptr (and *rc set) on failure.
void* device_scratch(hipStream_t st, int* rc) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxCachedDevices) {
    *rc = TCS_ERR_INVALID_ARG;
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  DeviceScratch& d = g_scratch[dev];
  bool zero = false;
  if (!d.ptr) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
      *rc = TCS_ERR_LAUNCH;
      return nullptr;
    }
    d.ptr = p;
    zero = true;
  } else if (d.calls >= TCS_WORKSPACE_MAX_CALLS) {
    zero = true;
  }
  if (zero) {
    const int z = fill_words(d.ptr, TCS_NOALLOC_SCRATCH_BYTES, 0, st);
    if (z != TCS_OK) {
      *rc = z;
      return nullptr;
    }
    d.calls = 0;
  }
  ++d.calls;
  *rc = TCS_OK;
  return d.ptr;
}

}  // namespace
}  // namespace tcs

extern "C" {

int tcs_scan_noalloc(const void* x, int x_dtype, void* y, int64_t n, int S, int tiles_per_wg, double start,
                     void* stream) {
  using namespace tcs;
  if (n < 0 || S <= 0) return TCS_ERR_INVALID_ARG;
  if (n == 0) return TCS_OK;
  if (tcs_scan_workspace_bytes(n, S, tiles_per_wg) > (size_t)TCS_NOALLOC_SCRATCH_BYTES)
    return TCS_ERR_SCRATCH_LIMIT;
  hipStream_t st = static_cast<hipStream_t>(stream);
  int rc = TCS_OK;
  void* ws = device_scratch(st, &rc);
  if (!ws) return rc;
  return scan_dispatch<kSegNone>(x, x_dtype, y, nullptr, n, 0, S, tiles_per_wg, start, ws,
                                 TCS_NOALLOC_SCRATCH_BYTES, stream);
}

}  // extern "C"
