// Lower-bound search of int32 needles in a sorted int32 array
// (torch.searchsorted(sorted, values, side='left', out_int32=True)), as the
// reference's run_searchsorted (src/torch/torch_searchsorted.h:33-82,
// tests/test_searchsorted.py).
//
// Each workgroup first stages every B-th haystack element in LDS
// (B = ceil(len / kSamples), so at most kSamples of them; the whole
// haystack when it fits).  A needle's lower bound c among the samples
// brackets its answer: sorted[(c-1)B] < key <= sorted[cB], so the answer
// lies in [(c-1)B + 1, min(cB, len)] (0 when c = 0), and a binary search
// over that window of fewer than B elements in global memory finishes it.
// The LDS stage replaces the upper levels of the search, whose probes all
// needles share; the lower levels are random probes into the haystack and
// bound the kernel.
#include "tcs_common.hpp"

namespace tcs {

constexpr int kSsThreads = 1024;
constexpr int kSsSamples = 16384;  // 64 KiB of LDS per workgroup

__global__ __launch_bounds__(kSsThreads) void searchsorted_kernel(const int32_t* __restrict__ sorted, int32_t len,
                                                                  const int32_t* __restrict__ values,
                                                                  int32_t* __restrict__ out, int64_t nv) {
  __shared__ int32_t s_samp[kSsSamples];
  const int32_t B = (int32_t)(((int64_t)len + kSsSamples - 1) / kSsSamples);
  const int32_t ns = len > 0 ? (int32_t)(((int64_t)len + B - 1) / B) : 0;  // <= kSsSamples
  for (int i = threadIdx.x; i < ns; i += kSsThreads) s_samp[i] = sorted[(int64_t)i * B];
  __syncthreads();
  const int64_t stride = (int64_t)gridDim.x * kSsThreads;
  for (int64_t i = (int64_t)blockIdx.x * kSsThreads + threadIdx.x; i < nv; i += stride) {
    const int32_t key = values[i];
    // Invariant of both searches: the answer lies in [lo, lo + count].
    int32_t lo = 0, count = ns;
    while (count > 0) {
      const int32_t half = count >> 1;
      const bool right = s_samp[lo + half] < key;
      lo = right ? lo + half + 1 : lo;
      count = right ? count - half - 1 : half;
    }
    const int64_t hi = (int64_t)lo * B < len ? (int64_t)lo * B : len;
    const int32_t glo = lo == 0 ? 0 : (lo - 1) * B + 1;
    count = lo == 0 ? 0 : (int32_t)hi - glo;
    lo = glo;
    while (count > 0) {
      const int32_t half = count >> 1;
      const bool right = sorted[lo + half] < key;
      lo = right ? lo + half + 1 : lo;
      count = right ? count - half - 1 : half;
    }
    out[i] = lo;
  }
}

// Four workgroups per CU of the current device (two are resident at a
// time with 64 KiB of LDS each; the second pair evens out the tail).
static uint32_t searchsorted_grid() {
  static std::atomic<int> cached[kMaxCachedDevices];
  return (uint32_t)device_cached(cached, [](int dev) {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) cus = 256;
    return 4 * cus;
  });
}

}  // namespace tcs

extern "C" {

int tcs_searchsorted(const int32_t* sorted, int64_t sorted_len, const int32_t* values, int32_t* out,
                     int64_t num_values, void* stream) {
  using namespace tcs;
  if (sorted_len < 0 || num_values < 0 || sorted_len >= ((int64_t)1 << 31)) return TCS_ERR_INVALID_ARG;
  if (num_values == 0) return TCS_OK;
  if (!values || !out || (sorted_len > 0 && !sorted)) return TCS_ERR_INVALID_ARG;
  const uint64_t need = (uint64_t)((num_values + kSsThreads - 1) / kSsThreads);
  const uint64_t cap = searchsorted_grid();
  const uint32_t blocks = (uint32_t)(need < cap ? need : cap);
  hipLaunchKernelGGL(searchsorted_kernel, dim3(blocks), dim3(kSsThreads), 0, static_cast<hipStream_t>(stream),
                     sorted, (int32_t)sorted_len, values, out, num_values);
  TCS_CHECK_LAUNCH();
  return TCS_OK;
}

}  // extern "C"
