// C entry point of the scans of fixed-length rows (run_scan_batch,
// run_row_scan, run_block_scan).  The kernel itself is in scan_kernel.hpp.
#include "scan_kernel.hpp"

extern "C" {

int tcs_scan_rows(const void* x, int x_dtype, void* y, int64_t n, int64_t row_len, int S,
                  int tiles_per_wg, void* workspace, size_t workspace_bytes, void* stream) {
  return tcs::scan_dispatch<tcs::kSegFixed>(x, x_dtype, y, nullptr, n, row_len, S, tiles_per_wg,
                                            0.0, workspace, workspace_bytes, stream);
}

}  // extern "C"
