// Measurement kernels: the random-gather ceiling of CSR SpMV and the
// streaming (read, write, copy) ceilings of HBM on this GPU.
//
// tcs_calib_gather_f16 reads n int32 column indices (nontemporal, 16 bytes
// per lane, as the SpMV kernel reads them) and gathers x[c] for each, with
// no products, no row bookkeeping and no look-back: the least work any
// kernel that forms y = A x from a CSR matrix must do.  With parts = 2 the
// even workgroups gather only columns in the lower half of x and the odd
// ones the upper half, as the split SpMV does (MI355X places workgroup b
// on XCD b mod 8, so each XCD's L2 then holds half of x).  Each thread
// writes one float (the sum of what it gathered) so that nothing is
// optimised away.  bench.py reports the SpMV kernel's time against this.
#include "tcs_common.hpp"
#include "../../include/tcuscan_calib.h"

namespace tcs {

constexpr int kCalibBlock = 512, kCalibK = 2, kCalibItems = 8 * kCalibK;  // columns per thread

template <int PARTS>
__global__ __launch_bounds__(kCalibBlock) void calib_gather_kernel(const int32_t* __restrict__ cols,
                                                                   int64_t n, const uint16_t* __restrict__ x,
                                                                   uint32_t x_len, float* __restrict__ out) {
  const uint32_t b = blockIdx.x;
  const int part = PARTS > 1 ? (int)(b % PARTS) : 0;
  const int64_t blk = (int64_t)(b / PARTS);
  const int lane = threadIdx.x & 63;
  const int64_t wave = blk * (kCalibBlock / 64) + (threadIdx.x >> 6);
  const int64_t wb = wave * (64 * kCalibItems);
  const uint32_t w = (x_len + PARTS - 1) / PARTS;
  const uint32_t lo = w * (uint32_t)part < x_len ? w * (uint32_t)part : x_len;
  const uint32_t cnt = (lo + w < x_len ? lo + w : x_len) - lo;
  float s = 0.f;
#pragma unroll
  for (int k = 0; k < kCalibK; ++k) {
    const int64_t gb = wb + k * 512 + lane * 8;
    int32_t c[8];
    if (gb + 8 <= n) {
      const u32x4* p = reinterpret_cast<const u32x4*>(cols + gb);
      const u32x4 a = __builtin_nontemporal_load(p), e = __builtin_nontemporal_load(p + 1);
      c[0] = (int32_t)a[0]; c[1] = (int32_t)a[1]; c[2] = (int32_t)a[2]; c[3] = (int32_t)a[3];
      c[4] = (int32_t)e[0]; c[5] = (int32_t)e[1]; c[6] = (int32_t)e[2]; c[7] = (int32_t)e[3];
    } else {
#pragma unroll
      for (int i = 0; i < 8; ++i) c[i] = gb + i < n ? cols[gb + i] : -1;
    }
    uint16_t g[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) g[i] = ((uint32_t)c[i] - lo < cnt) ? x[c[i]] : (uint16_t)0;
#pragma unroll
    for (int i = 0; i < 8; ++i) s += (float)__builtin_bit_cast(_Float16, g[i]);
  }
  out[(int64_t)b * kCalibBlock + threadIdx.x] = s;
}

// Streaming ceilings: every thread moves one 16-byte vector (nontemporal,
// as the scan's whole-line stores are).  mode 0 reads (a per-thread xor
// that is stored only if it matches a value it never takes, so nothing is
// written), 1 writes, 2 copies.
template <int MODE>
__global__ __launch_bounds__(kCalibBlock) void calib_stream_kernel(const u32x4* __restrict__ src,
                                                                   u32x4* __restrict__ dst, int64_t nvec) {
  const int64_t i = (int64_t)blockIdx.x * kCalibBlock + threadIdx.x;
  if (i >= nvec) return;
  if constexpr (MODE == 0) {
    const u32x4 v = __builtin_nontemporal_load(src + i);
    if ((v[0] ^ v[1] ^ v[2] ^ v[3]) == 0x9e3779b9u && v[0] == 0x7f4a7c15u) dst[i] = v;
  } else if constexpr (MODE == 1) {
    const u32x4 v = {(uint32_t)i, 0u, 0u, 0u};
    __builtin_nontemporal_store(v, dst + i);
  } else {
    __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
  }
}

}  // namespace tcs

extern "C" {

int tcs_calib_stream(const void* src, void* dst, size_t bytes, int mode, void* stream) {
  using namespace tcs;
  if (mode < 0 || mode > 2 || bytes % 16 != 0) return TCS_ERR_INVALID_ARG;
  if (bytes == 0) return TCS_OK;
  if (!dst || (mode != 1 && !src) || !aligned16(dst) || (src && !aligned16(src))) return TCS_ERR_INVALID_ARG;
  const int64_t nvec = (int64_t)(bytes / 16);
  const int64_t blocks = (nvec + kCalibBlock - 1) / kCalibBlock;
  if (blocks >= ((int64_t)1 << 31)) return TCS_ERR_INVALID_ARG;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const auto* s = static_cast<const u32x4*>(src);
  auto* d = static_cast<u32x4*>(dst);
  if (mode == 0) {
    hipLaunchKernelGGL((calib_stream_kernel<0>), dim3((uint32_t)blocks), dim3(kCalibBlock), 0, st, s, d, nvec);
  } else if (mode == 1) {
    hipLaunchKernelGGL((calib_stream_kernel<1>), dim3((uint32_t)blocks), dim3(kCalibBlock), 0, st, s, d, nvec);
  } else {
    hipLaunchKernelGGL((calib_stream_kernel<2>), dim3((uint32_t)blocks), dim3(kCalibBlock), 0, st, s, d, nvec);
  }
  TCS_CHECK_LAUNCH();
  return TCS_OK;
}

size_t tcs_calib_gather_out_elems(int64_t n, int parts) {
  if (n <= 0 || (parts != 1 && parts != 2)) return 0;
  const int64_t per_block = (int64_t)tcs::kCalibBlock * tcs::kCalibItems;
  return (size_t)((n + per_block - 1) / per_block) * (size_t)parts * tcs::kCalibBlock;
}

int tcs_calib_gather_f16(const int32_t* cols, int64_t n, const void* x, int64_t x_len, int parts, float* out,
                         void* stream) {
  using namespace tcs;
  if (n < 0 || x_len < 0 || x_len >= ((int64_t)1 << 31) || (parts != 1 && parts != 2)) return TCS_ERR_INVALID_ARG;
  if (n == 0) return TCS_OK;
  if (!cols || !x || !out || !aligned16(cols)) return TCS_ERR_INVALID_ARG;
  const int64_t per_block = (int64_t)kCalibBlock * kCalibItems;
  const int64_t blocks = (n + per_block - 1) / per_block * parts;
  if (blocks >= ((int64_t)1 << 31)) return TCS_ERR_INVALID_ARG;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const auto* xv = static_cast<const uint16_t*>(x);
  if (parts == 2) {
    hipLaunchKernelGGL((calib_gather_kernel<2>), dim3((uint32_t)blocks), dim3(kCalibBlock), 0, st, cols, n, xv,
                       (uint32_t)x_len, out);
  } else {
    hipLaunchKernelGGL((calib_gather_kernel<1>), dim3((uint32_t)blocks), dim3(kCalibBlock), 0, st, cols, n, xv,
                       (uint32_t)x_len, out);
  }
  TCS_CHECK_LAUNCH();
  return TCS_OK;
}

}  // extern "C"
