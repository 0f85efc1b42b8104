// Kept weights of top-p (nucleus) filtering over probabilities already
// sorted largest first, and their inclusive prefix mass: the mask step of
// the reference's top-p recipe (scripts/profiling/profile_tcuscan_ops.py:
// 335-358 and 376-398: `mask = probs_sum - probs_sort > p;
// probs_sort[mask] = 0`), fused into one pass:
//   kept[i] = (mass[i] - x[i] > p) ? 0 : x[i]   (float32 arithmetic)
// with x float16 and mass float32, as torch evaluates that expression.
#include "tcs_common.hpp"

namespace tcs {

constexpr int kTopPG = 8;

__global__ __launch_bounds__(256) void top_p_keep_kernel(const uint16_t* __restrict__ x,
                                                         const float* __restrict__ mass, float p,
                                                         float* __restrict__ kept, int64_t n) {
  const int64_t b = ((int64_t)blockIdx.x * 256 + threadIdx.x) * kTopPG;
  if (b >= n) return;
  uint16_t h[kTopPG];
  float m[kTopPG], o[kTopPG];
  const bool full = b + kTopPG <= n;
  if (full) {
    load_items<uint16_t, kTopPG, true>(x, b, n, h);
    load_items<float, kTopPG, true>(mass, b, n, m);
  } else {
    load_items<uint16_t, kTopPG, false>(x, b, n, h);
    load_items<float, kTopPG, false>(mass, b, n, m);
  }
#pragma unroll
  for (int i = 0; i < kTopPG; ++i) {
    const float v = (float)__builtin_bit_cast(_Float16, h[i]);
    o[i] = m[i] - v > p ? 0.0f : v;
  }
  if (full) {
    store_items<float, kTopPG, true>(kept, b, n, o);
  } else {
    store_items<float, kTopPG, false>(kept, b, n, o);
  }
}

}  // namespace tcs

extern "C" {

int tcs_top_p_keep(const void* sorted_probs, const float* mass, float p, float* kept, int64_t n, void* stream) {
  using namespace tcs;
  if (n < 0) return TCS_ERR_INVALID_ARG;
  if (n == 0) return TCS_OK;
  if (!sorted_probs || !mass || !kept) return TCS_ERR_INVALID_ARG;
  if (!aligned16(sorted_probs) || !aligned16(mass) || !aligned16(kept)) return TCS_ERR_INVALID_ARG;
  const int64_t per = 256 * kTopPG;
  const uint64_t blocks = (uint64_t)((n + per - 1) / per);
  if (blocks > 0x7fffffffull) return TCS_ERR_INVALID_ARG;
  hipLaunchKernelGGL(top_p_keep_kernel, dim3((uint32_t)blocks), dim3(256), 0, static_cast<hipStream_t>(stream),
                     static_cast<const uint16_t*>(sorted_probs), mass, p, kept, n);
  TCS_CHECK_LAUNCH();
  return TCS_OK;
}

}  // extern "C"
