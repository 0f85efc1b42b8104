// Lower-bound search of int32 needles in a sorted int32 array
// (torch.searchsorted(sorted, values, side='left', out_int32=True)), as the
// reference's run_searchsorted (src/torch/torch_searchsorted.h:33-82,
// tests/test_searchsorted.py).  One lane per needle: a branch-free binary
// search over the haystack in global memory (its upper levels stay in the
// caches), so a wave's 64 searches run side by side.
#include "tcs_common.hpp"

namespace tcs {

__global__ __launch_bounds__(256) void searchsorted_kernel(const int32_t* __restrict__ sorted, int64_t len,
                                                           const int32_t* __restrict__ values,
                                                           int32_t* __restrict__ out, int64_t nv) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= nv) return;
  const int32_t key = values[i];
  // Invariant: sorted[lo - 1] < key <= sorted[lo + count] (virtual bounds).
  int64_t lo = 0, count = len;
  while (count > 0) {
    const int64_t half = count >> 1;
    const bool right = sorted[lo + half] < key;
    lo = right ? lo + half + 1 : lo;
    count = right ? count - half - 1 : half;
  }
  out[i] = (int32_t)lo;
}

}  // namespace tcs

extern "C" {

int tcs_searchsorted(const int32_t* sorted, int64_t sorted_len, const int32_t* values, int32_t* out,
                     int64_t num_values, void* stream) {
  using namespace tcs;
  if (sorted_len < 0 || num_values < 0 || sorted_len >= ((int64_t)1 << 31)) return TCS_ERR_INVALID_ARG;
  if (num_values == 0) return TCS_OK;
  if (!values || !out || (sorted_len > 0 && !sorted)) return TCS_ERR_INVALID_ARG;
  const uint64_t blocks = (uint64_t)((num_values + 255) / 256);
  if (blocks > 0x7fffffffull) return TCS_ERR_INVALID_ARG;
  hipLaunchKernelGGL(searchsorted_kernel, dim3((uint32_t)blocks), dim3(256), 0, static_cast<hipStream_t>(stream),
                     sorted, sorted_len, values, out, num_values);
  TCS_CHECK_LAUNCH();
  return TCS_OK;
}

}  // extern "C"
