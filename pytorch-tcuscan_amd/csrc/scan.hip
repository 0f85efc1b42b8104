// C entry points of the plain inclusive scan (DPP local scan; the
// matrix-unit local scan lives in scan_mfma.hip).  The kernel itself is in
// scan_kernel.hpp.
#include "scan_kernel.hpp"

extern "C" {

size_t tcs_scan_workspace_bytes(int64_t n, int S, int tiles_per_wg) {
  if (n <= 0 || S <= 0) return tcs::lookback_ws_bytes(0);
  // Sized for the smallest tile any dtype uses with these parameters (the
  // default tile of short inputs is 8K).
  int64_t tile = tcs::tile_elems_for(S, tiles_per_wg);
  if (tile > 16384) tile = 16384;
  if (tiles_per_wg <= 0 && tile > 8192) tile = 8192;
  return tcs::lookback_ws_bytes((uint64_t)((n + tile - 1) / tile));
}

int tcs_scan(const void* x, int x_dtype, void* y, int64_t n, int S, int tiles_per_wg, double start,
             void* workspace, size_t workspace_bytes, void* stream) {
  return tcs::scan_dispatch<tcs::kSegNone>(x, x_dtype, y, nullptr, n, 0, S, tiles_per_wg, start,
                                           workspace, workspace_bytes, stream);
}

int tcs_scan_method(const void* x, int x_dtype, void* y, int64_t n, int S, int tiles_per_wg,
                    double start, int method, int lookback_window, void* workspace,
                    size_t workspace_bytes, void* stream) {
  return tcs::scan_dispatch<tcs::kSegNone>(x, x_dtype, y, nullptr, n, 0, S, tiles_per_wg, start,
                                           workspace, workspace_bytes, stream, method, lookback_window);
}

}  // extern "C"
