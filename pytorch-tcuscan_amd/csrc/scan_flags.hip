// C entry point of the flag-segmented scan (run_seg_scan,
// run_seg_scan_vec).  The kernel itself is in scan_kernel.hpp.
#include "scan_kernel.hpp"

namespace tcs {

// The flag-segmented scan with an explicit look-back window (for the batch
// plan's tuner; 0 = default).
int seg_scan_flags_cfg(const void* x, int x_dtype, const int8_t* flags, void* y, int64_t n, int S,
                       int tiles_per_wg, int window, void* workspace, size_t workspace_bytes,
                       void* stream) {
  return scan_dispatch<kSegFlags>(x, x_dtype, y, flags, n, 0, S, tiles_per_wg, 0.0, workspace,
                                  workspace_bytes, stream, TCS_METHOD_AUTO, window);
}

}  // namespace tcs

extern "C" {

int tcs_seg_scan_flags(const void* x, int x_dtype, const int8_t* flags, void* y, int64_t n, int S,
                       int tiles_per_wg, void* workspace, size_t workspace_bytes, void* stream) {
  return tcs::scan_dispatch<tcs::kSegFlags>(x, x_dtype, y, flags, n, 0, S, tiles_per_wg, 0.0,
                                            workspace, workspace_bytes, stream);
}

}  // extern "C"
