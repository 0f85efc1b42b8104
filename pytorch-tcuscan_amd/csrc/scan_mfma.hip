// The plain scan with the matrix-unit local scan (tcs_mfma.hpp) for fp16,
// bf16 and int8 inputs, reached from scan_dispatch through launch_scan_mfma.
#include "scan_kernel.hpp"

namespace tcs {

int launch_scan_mfma(int x_dtype, int64_t tile, const void* x, void* y, int64_t n, double start,
                     void* ws, uint32_t tiles, int window, hipStream_t st) {
  if (x_dtype == TCS_F16)
    return launch_scan_tile<CodecF16, kSegNone, kLocMfma>(tile, x, y, nullptr, n, 0, (double)(float)start,
                                                          ws, tiles, window, st);
  if (x_dtype == TCS_BF16)
    return launch_scan_tile<CodecBF16, kSegNone, kLocMfma>(tile, x, y, nullptr, n, 0, (double)(float)start,
                                                           ws, tiles, window, st);
  if (x_dtype == TCS_I8)
    return launch_scan_tile<CodecI8U, kSegNone, kLocMfma>(tile, x, y, nullptr, n, 0, (int)start, ws,
                                                         tiles, window, st);
  return TCS_ERR_UNSUPPORTED_DTYPE;
}

}  // namespace tcs
