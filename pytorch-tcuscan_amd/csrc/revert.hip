// Segmented scan recovered from plain scans (the reference's
// run_seg_scan_mc_revert, src/torch/torch_seg_ops.h:108-153,
// src/seg_scan_mc_revert.cpp; tests/test_seg_scan_mc_revert.py): given the
// inclusive scan sx of x, the inclusive scan sf of the segment-start flags
// and d[j] = sx at the last element of segment j, the segmented scan is
// z[i] = sx[i] - d[sf[i] - 1] (z[i] = sx[i] in segment 0), one float32
// subtraction per element.
#include "tcs_common.hpp"

namespace tcs {

constexpr int kRevG = 4;

__global__ __launch_bounds__(256) void seg_revert_kernel(const float* __restrict__ sx, const int32_t* __restrict__ sf,
                                                         const float* __restrict__ d, int64_t d_len,
                                                         float* __restrict__ z, int64_t n) {
  const int64_t b = ((int64_t)blockIdx.x * 256 + threadIdx.x) * kRevG;
  if (b >= n) return;
  float v[kRevG], o[kRevG];
  int32_t s[kRevG];
  const bool full = b + kRevG <= n;
  if (full) {
    load_items<float, kRevG, true>(sx, b, n, v);
    load_items<int32_t, kRevG, true>(sf, b, n, s);
  } else {
    load_items<float, kRevG, false>(sx, b, n, v);
    load_items<int32_t, kRevG, false>(sf, b, n, s);
  }
#pragma unroll
  for (int i = 0; i < kRevG; ++i) {
    const int64_t j = (int64_t)s[i] - 1;  // segment ids past d (bad input) subtract nothing
    o[i] = (j >= 0 && j < d_len) ? v[i] - d[j] : v[i];
  }
  if (full) {
    store_items<float, kRevG, true>(z, b, n, o);
  } else {
    store_items<float, kRevG, false>(z, b, n, o);
  }
}

}  // namespace tcs

extern "C" {

int tcs_seg_scan_revert(const float* scan_x, const int32_t* scan_f, const float* seg_ends, int64_t num_seg_ends,
                        float* z, int64_t n, void* stream) {
  using namespace tcs;
  if (n < 0 || num_seg_ends < 0) return TCS_ERR_INVALID_ARG;
  if (n == 0) return TCS_OK;
  if (!scan_x || !scan_f || !z || (num_seg_ends > 0 && !seg_ends)) return TCS_ERR_INVALID_ARG;
  if (!aligned16(scan_x) || !aligned16(scan_f) || !aligned16(z)) return TCS_ERR_INVALID_ARG;
  const int64_t blocks = (n + 256 * kRevG - 1) / (256 * kRevG);
  if (blocks > 0x7fffffffll) return TCS_ERR_INVALID_ARG;
  hipLaunchKernelGGL(seg_revert_kernel, dim3((uint32_t)blocks), dim3(256), 0, static_cast<hipStream_t>(stream),
                     scan_x, scan_f, seg_ends, num_seg_ends, z, n);
  TCS_CHECK_LAUNCH();
  return TCS_OK;
}

}  // extern "C"
