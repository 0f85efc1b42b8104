/*
 * tcuscan for AMD Instinct MI355X (gfx950) -- C ABI of the HIP kernel library
 * libtcuscan_hip.so.
 *
 * Every entry point takes plain device pointers, element counts, a dtype
 * code and a HIP stream passed as `void*` (a hipStream_t; NULL = the legacy
 * default stream).  Calls are stream-ordered and asynchronous: they validate
 * their arguments, enqueue kernels and return a tcs_status code without
 * synchronising.  Invalid arguments are rejected before anything is
 * launched.  No entry point allocates device memory; kernels that need
 * scratch space take a caller-provided workspace whose size is given by the
 * matching *_workspace_bytes function.
 *
 * Workspace contract (look-back kernels).  A workspace must be all-zero
 * before its first use.  Each call leaves it ready for the next call, so a
 * workspace is allocated and zeroed once and then reused, by one stream at
 * a time (calls that overlap on one workspace race).
 *
 * The Ascend reference (huawei-csl/pytorch-tcuscan) launches AscendC
 * kernels `kernel(GM_ADDR in..., GM_ADDR out..., GM_ADDR workspace,
 * GM_ADDR tiling)` from the C++ wrappers under src/torch/; each declaration
 * below names the reference kernel entry points and wrapper it replaces.
 */
#ifndef TCUSCAN_H_
#define TCUSCAN_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
enum tcs_status {
  TCS_OK = 0,
  TCS_ERR_INVALID_ARG = 1,       /* bad size, null pointer, misalignment */
  TCS_ERR_UNSUPPORTED_DTYPE = 2, /* dtype outside the op's accepted set */
  TCS_ERR_WORKSPACE = 3,         /* workspace missing or too small */
  TCS_ERR_LAUNCH = 4             /* the HIP runtime rejected a launch */
};

/* Element type codes (match the tensor dtypes of the reference API). */
enum tcs_dtype {
  TCS_F16 = 0,
  TCS_BF16 = 1,
  TCS_F32 = 2,
  TCS_I8 = 3,
  TCS_I16 = 4,
  TCS_I32 = 5,
  TCS_U32 = 6,
  TCS_I64 = 7
};

/* Human-readable text of a tcs_status code. */
const char* tcs_status_string(int status);

/* ABI version of this library (major * 100 + minor). */
int tcs_abi_version(void);

/* ------------------------------------------------------------------------
 * Scan family
 * --------------------------------------------------------------------- */

/* Workspace bytes for tcs_scan / tcs_scan_rows / tcs_seg_scan_flags over n
 * elements with tile side S and tiles_per_wg S*S tiles per workgroup
 * (tiles_per_wg <= 0 selects the default). */
size_t tcs_scan_workspace_bytes(int64_t n, int S, int tiles_per_wg);

/* Inclusive prefix sum y[i] = start + sum(x[0..i]).
 * x dtype: F16, BF16, F32 -> y float32 (accumulated in double, so the fp16
 * result is the correctly rounded prefix whenever the exact prefix fits in
 * 53 bits); I8, I32 -> y int32 (exact, two's-complement wrap).
 * `start` is rounded to float (int32 for integer inputs) before use.
 * Single pass over HBM with decoupled look-back between tiles.
 * Replaces: scan_multi_core_{fp16,int8}[_no_l2] (src/scan_multi_core.cpp:101,
 * 117,134,150) behind tcuscan::run_scan_multi_core / _no_l2
 * (src/torch/torch_scan.h:184,251); scan_single_core_{int8,fp16,fp32}
 * (src/scan_single_core.cpp:22,48,75) behind run_scan_single_core
 * (src/torch/torch_scan.h:47); scan_multi_cube_fp16 (src/scan_multi_cube.cpp:79)
 * behind run_scan_multi_cube (src/torch/torch_scan.h:416). */
int tcs_scan(const void* x, int x_dtype, void* y, int64_t n, int S, int tiles_per_wg,
             double start, void* workspace, size_t workspace_bytes, void* stream);

/* Independent inclusive scans of consecutive rows of row_len elements
 * (the last row may be shorter).  Same dtypes as tcs_scan.
 * Replaces: scan_batch_{fp16,fp32} (src/scan_batch.cpp:21,59) behind
 * run_scan_batch (src/torch/torch_scan.h:119); row_scan_fp16
 * (src/row_scan.cpp:52) behind run_row_scan (src/torch/torch_scan.h:312);
 * block_scan_fp16 (src/block_scan.cpp:58) behind run_block_scan
 * (src/torch/torch_scan.h:362). */
int tcs_scan_rows(const void* x, int x_dtype, void* y, int64_t n, int64_t row_len, int S,
                  int tiles_per_wg, void* workspace, size_t workspace_bytes, void* stream);

/* Flag-segmented inclusive scan: flags[i] != 0 starts a new segment at i.
 * Same dtypes as tcs_scan.
 * Replaces: seg_scan_single_core / seg_scan_vec_single_core
 * (src/seg_scan_single_core.cpp, src/seg_scan_vec_single_core.cpp) behind
 * run_seg_scan / run_seg_scan_vec (src/torch/torch_seg_ops.h:46,164). */
int tcs_seg_scan_flags(const void* x, int x_dtype, const int8_t* flags, void* y, int64_t n,
                       int S, int tiles_per_wg, void* workspace, size_t workspace_bytes,
                       void* stream);

#ifdef __cplusplus
}
#endif

#endif /* TCUSCAN_H_ */
