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

// The samples of `sorted` in LDS: every B-th element, at most kSsSamples.
struct SsSamples {
  int32_t B, ns;
};

__device__ __forceinline__ SsSamples stage_samples(int32_t* s_samp, const int32_t* __restrict__ sorted,
                                                   int32_t len) {
  const int32_t B = (int32_t)(((int64_t)len + kSsSamples - 1) / kSsSamples);
  const int32_t ns = len > 0 ? (int32_t)(((int64_t)len + B - 1) / B) : 0;  // <= kSsSamples
  for (int i = threadIdx.x; i < ns; i += kSsThreads) s_samp[i] = sorted[(int64_t)i * B];
  __syncthreads();
  return SsSamples{B, ns};
}

// Lower bound of key in sorted[lo, lo + count] (the answer is known to lie
// there): a binary search in global memory.
__device__ __forceinline__ int32_t lower_bound_in(const int32_t* __restrict__ sorted, int32_t lo, int32_t count,
                                                  int32_t key) {
  while (count > 0) {
    const int32_t half = count >> 1;
    const bool right = sorted[lo + half] < key;
    lo = right ? lo + half + 1 : lo;
    count = right ? count - half - 1 : half;
  }
  return lo;
}

// Lower bound of key in sorted[0, len): the samples bracket it, then a
// search over the window of fewer than B elements between two samples.
__device__ __forceinline__ int32_t lower_bound_staged(const int32_t* s_samp, SsSamples sm,
                                                      const int32_t* __restrict__ sorted, int32_t len,
                                                      int32_t key) {
  // Invariant of both searches: the answer lies in [lo, lo + count].
  int32_t lo = 0, count = sm.ns;
  while (count > 0) {
    const int32_t half = count >> 1;
    const bool right = s_samp[lo + half] < key;
    lo = right ? lo + half + 1 : lo;
    count = right ? count - half - 1 : half;
  }
  const int64_t hi = (int64_t)lo * sm.B < len ? (int64_t)lo * sm.B : len;
  const int32_t glo = lo == 0 ? 0 : (lo - 1) * sm.B + 1;
  return lower_bound_in(sorted, glo, lo == 0 ? 0 : (int32_t)hi - glo, key);
}

__global__ __launch_bounds__(kSsThreads) void searchsorted_kernel(const int32_t* __restrict__ sorted, int32_t len,
                                                                  const int32_t* __restrict__ values,
                                                                  int32_t* __restrict__ out, int64_t nv) {
  __shared__ int32_t s_samp[kSsSamples];
  const SsSamples sm = stage_samples(s_samp, sorted, len);
  const int64_t stride = (int64_t)gridDim.x * kSsThreads;
  for (int64_t i = (int64_t)blockIdx.x * kSsThreads + threadIdx.x; i < nv; i += stride)
    out[i] = lower_bound_staged(s_samp, sm, sorted, len, values[i]);
}

// ---------------------------------------------------------------------------
// Bucket index (tcs_searchsorted_indexed).  Keys map to order keys
// u = bits ^ 0x80000000 (unsigned order = signed order); the haystack's
// order keys span [umin, umax] (its two ends).  Bucket of u:
// (u - umin) >> shift, with shift the smallest that leaves at most 2^bits
// buckets.  start[k] = first index whose bucket is >= k (k <= 2^bits;
// len for the buckets past the last), so a needle in bucket k has its
// lower bound in [start[k], start[k + 1]]: elements before start[k] are
// in lower buckets, hence smaller, and elements from start[k + 1] on are in
// higher buckets, hence larger.  With keys spread over their range a
// bucket holds about 16 elements and the search is one table read and one
// cache line; a wide window (clustered keys) takes the sampled search.
constexpr int32_t kSsWindowMax = 2048;

struct BucketMap {
  uint32_t umin, umax;
  int shift;
  uint32_t used;  // buckets that hold order keys of [umin, umax]
};

__device__ __forceinline__ uint32_t order_key(int32_t v) { return (uint32_t)v ^ 0x80000000u; }

__device__ __forceinline__ BucketMap bucket_map(const int32_t* __restrict__ sorted, int32_t len, int bits) {
  BucketMap m;
  m.umin = order_key(sorted[0]);
  m.umax = order_key(sorted[len - 1]);
  const uint32_t range = m.umax - m.umin;
  const int width = range == 0 ? 0 : 32 - __builtin_clz(range);
  m.shift = width > bits ? width - bits : 0;
  m.used = (range >> m.shift) + 1u;  // <= 2^bits
  return m;
}

// start[k] for k in [0, 2^bits], one thread per entry.
__global__ __launch_bounds__(kSsThreads) void bucket_start_kernel(const int32_t* __restrict__ sorted, int32_t len,
                                                                  int bits, int32_t* __restrict__ start) {
  __shared__ int32_t s_samp[kSsSamples];
  const SsSamples sm = stage_samples(s_samp, sorted, len);
  const BucketMap m = bucket_map(sorted, len, bits);
  const int64_t entries = ((int64_t)1 << bits) + 1;
  const int64_t stride = (int64_t)gridDim.x * kSsThreads;
  for (int64_t k = (int64_t)blockIdx.x * kSsThreads + threadIdx.x; k < entries; k += stride) {
    int32_t r = len;
    if (k < (int64_t)m.used) {
      // First order key of bucket k: below umax + 1 <= 2^32, so it fits.
      const uint32_t bu = m.umin + ((uint32_t)k << m.shift);
      r = lower_bound_staged(s_samp, sm, sorted, len, (int32_t)(bu ^ 0x80000000u));
    }
    start[k] = r;
  }
}

__global__ __launch_bounds__(kSsThreads) void searchsorted_indexed_kernel(const int32_t* __restrict__ sorted,
                                                                          int32_t len, int bits,
                                                                          const int32_t* __restrict__ start,
                                                                          const int32_t* __restrict__ values,
                                                                          int32_t* __restrict__ out, int64_t nv) {
  __shared__ int32_t s_samp[kSsSamples];
  const SsSamples sm = stage_samples(s_samp, sorted, len);
  const BucketMap m = bucket_map(sorted, len, bits);
  const int64_t stride = (int64_t)gridDim.x * kSsThreads;
  for (int64_t i = (int64_t)blockIdx.x * kSsThreads + threadIdx.x; i < nv; i += stride) {
    const int32_t key = values[i];
    const uint32_t u = order_key(key);
    int32_t r;
    if (u < m.umin) {
      r = 0;
    } else if (u > m.umax) {
      r = len;
    } else {
      const uint32_t k = (u - m.umin) >> m.shift;
      const int32_t lo = start[k], hi = start[k + 1];
      r = hi - lo <= kSsWindowMax ? lower_bound_in(sorted, lo, hi - lo, key)
                                  : lower_bound_staged(s_samp, sm, sorted, len, key);
    }
    out[i] = r;
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

// Table bits for a haystack of len elements: about 16 elements per bucket,
// between 2^10 and 2^20 buckets (a table of at most 4 MiB).
static int bucket_bits(int64_t len) {
  int b = 0;
  while (b < 62 && ((int64_t)1 << (b + 1)) <= len) ++b;  // floor(log2(len))
  b -= 4;
  return b < 10 ? 10 : (b > 20 ? 20 : b);
}

size_t tcs_searchsorted_workspace_bytes(int64_t sorted_len) {
  if (sorted_len <= 0) return 0;
  return (((size_t)1 << bucket_bits(sorted_len)) + 1) * sizeof(int32_t);
}

int tcs_searchsorted_indexed(const int32_t* sorted, int64_t sorted_len, const int32_t* values, int32_t* out,
                             int64_t num_values, void* workspace, size_t workspace_bytes, void* stream) {
  using namespace tcs;
  if (sorted_len < 0 || num_values < 0 || sorted_len >= ((int64_t)1 << 31)) return TCS_ERR_INVALID_ARG;
  if (num_values == 0) return TCS_OK;
  if (!values || !out || (sorted_len > 0 && !sorted)) return TCS_ERR_INVALID_ARG;
  if (sorted_len == 0) return tcs_searchsorted(sorted, sorted_len, values, out, num_values, stream);
  if (!workspace || workspace_bytes < tcs_searchsorted_workspace_bytes(sorted_len)) return TCS_ERR_WORKSPACE;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int bits = bucket_bits(sorted_len);
  auto* start = static_cast<int32_t*>(workspace);
  const uint64_t cap = searchsorted_grid();
  const uint64_t entries = ((uint64_t)1 << bits) + 1;
  const uint64_t tb = (entries + kSsThreads - 1) / kSsThreads;
  hipLaunchKernelGGL(bucket_start_kernel, dim3((uint32_t)(tb < cap ? tb : cap)), dim3(kSsThreads), 0, st, sorted,
                     (int32_t)sorted_len, bits, start);
  TCS_CHECK_LAUNCH();
  const uint64_t need = (uint64_t)((num_values + kSsThreads - 1) / kSsThreads);
  hipLaunchKernelGGL(searchsorted_indexed_kernel, dim3((uint32_t)(need < cap ? need : cap)), dim3(kSsThreads), 0, st,
                     sorted, (int32_t)sorted_len, bits, start, values, out, num_values);
  TCS_CHECK_LAUNCH();
  return TCS_OK;
}

}  // extern "C"
