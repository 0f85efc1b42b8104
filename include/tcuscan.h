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
 * before its first use (tcs_workspace_reset zeroes one on a stream).  Each
 * call leaves it ready for the next call, so a workspace is allocated and
 * zeroed once and then reused, by one stream at a time (calls that overlap
 * on one workspace race), and re-zeroed at least once every
 * TCS_WORKSPACE_MAX_CALLS calls: look-back launches tag their tile
 * descriptors with a 30-bit launch tag kept in the workspace, and tags
 * repeat after 2^30 - 1 launches (one call launches at most 8 look-back
 * kernels).  The workspace must be 8-byte aligned.
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

/* Calls between re-zeroings of one workspace (see the workspace contract). */
#define TCS_WORKSPACE_MAX_CALLS (1u << 24)

/* Status codes. */
enum tcs_status {
  TCS_OK = 0,
  TCS_ERR_INVALID_ARG = 1,       /* bad size, null pointer, misalignment */
  TCS_ERR_UNSUPPORTED_DTYPE = 2, /* dtype outside the op's accepted set */
  TCS_ERR_WORKSPACE = 3,         /* workspace missing or too small */
  TCS_ERR_LAUNCH = 4,            /* the HIP runtime rejected a launch */
  TCS_ERR_SCRATCH_LIMIT = 5      /* input needs more scratch than the fixed
                                    array of tcs_scan_noalloc holds */
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

/* Zeroes a workspace on `stream` (a fill kernel, so it may be captured in a
 * HIP graph): before its first use, and again at least once every
 * TCS_WORKSPACE_MAX_CALLS calls.  This also restores the default poll
 * limit.  No counterpart in the reference, which allocates a fresh
 * workspace per call (src/torch/commons.h:25-55). */
int tcs_workspace_reset(void* workspace, size_t workspace_bytes, void* stream);

/* Look-back poll limit of launches on this workspace: a tile whose
 * predecessor has not published after `polls` polls computes that
 * predecessor's aggregate from the input itself (the forward-progress
 * fallback).  0 falls back at the first poll that finds a predecessor
 * missing, which exercises the fallback on every walk; a negative value
 * restores the default (512).  Results do not depend on it.  Stream-ordered
 * (a one-word write by a kernel).  No counterpart in the reference, whose
 * multi-core kernels wait at SyncAll barriers instead
 * (src/kernels/kernel_scan_multi_core.h:92,111,143). */
int tcs_workspace_set_poll_limit(void* workspace, size_t workspace_bytes, int polls, void* stream);

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
 * Single pass over HBM with decoupled look-back between tiles; the local
 * scan method is TCS_METHOD_AUTO (see tcs_scan_method).
 * Replaces: scan_multi_core_{fp16,int8}[_no_l2] (src/scan_multi_core.cpp:101,
 * 117,134,150) behind tcuscan::run_scan_multi_core / _no_l2
 * (src/torch/torch_scan.h:184,251); scan_single_core_{int8,fp16,fp32}
 * (src/scan_single_core.cpp:22,48,75) behind run_scan_single_core
 * (src/torch/torch_scan.h:47); scan_multi_cube_fp16 (src/scan_multi_cube.cpp:79)
 * behind run_scan_multi_cube (src/torch/torch_scan.h:416). */
int tcs_scan(const void* x, int x_dtype, void* y, int64_t n, int S, int tiles_per_wg,
             double start, void* workspace, size_t workspace_bytes, void* stream);

/* Local scan method of tcs_scan_method. */
enum tcs_scan_method {
  TCS_METHOD_AUTO = 0, /* the shipped method, the faster one measured on MI355X
                          (DPP for every dtype) */
  TCS_METHOD_DPP = 1,  /* lane-serial group sums joined by DPP wavefront scans */
  TCS_METHOD_MFMA = 2  /* 32x32 units scanned on the matrix cores (F16, BF16, I8 only) */
};

/* tcs_scan with an explicit local scan method and look-back window
 * (predecessor tiles polled per step, 1..64; 0 = the default: 6 for DPP, 32
 * for MFMA); results are
 * identical for every choice (bit-exact).  TCS_METHOD_MFMA multiplies each 1024-element
 * unit, as 32 rows of 32, by a lower-triangular ones matrix on
 * v_mfma_f32_32x32x16_f16 (fp16, split into two exactly-summed parts),
 * v_mfma_f32_32x32x16_bf16 (bf16, split into three exactly-summed bands
 * of each row's exponent range) or v_mfma_i32_32x32x32_i8 (int8), and is
 * accepted for F16, BF16 and I8 inputs.
 * Replaces: the Cube-unit row scan of scan_multi_core_{fp16,int8}
 * (src/kernels/kernel_row_scan.h:97-131 via
 * src/kernels/kernel_scan_multi_core.h:70-152) behind run_scan_multi_core
 * (src/torch/torch_scan.h:184). */
int tcs_scan_method(const void* x, int x_dtype, void* y, int64_t n, int S, int tiles_per_wg,
                    double start, int method, int lookback_window, void* workspace,
                    size_t workspace_bytes, void* stream);

/* Bytes of the fixed per-GPU scratch array of tcs_scan_noalloc. */
#define TCS_NOALLOC_SCRATCH_BYTES (1u << 20)

/* tcs_scan without a workspace argument, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference,
 * whose wrappers allocate a workspace per call (src/torch/commons.h:25-55).
 * The look-back workspace is a file-scope device array of
 * TCS_NOALLOC_SCRATCH_BYTES that every GPU holds until the process ends;
 * the call allocates nothing (no hipMalloc, hipFree or allocating library
 * routine).  Inputs whose workspace (tcs_scan_workspace_bytes(n, S,
 * tiles_per_wg)) exceeds that size are rejected with
 * TCS_ERR_SCRATCH_LIMIT before anything is launched: with the default
 * tiles that is n above about 5.3e8 elements.  All calls on one GPU share
 * its array, so calls that overlap on one GPU (other streams, other host
 * threads) race; stream-ordered calls do not.  Same arguments, dtypes and
 * results as tcs_scan otherwise. */
int tcs_scan_noalloc(const void* x, int x_dtype, void* y, int64_t n, int S, int tiles_per_wg,
                     double start, void* stream);

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


/* ------------------------------------------------------------------------
 * Reduction and down-sweep building blocks
 *
 * These keep the reference's block geometry.  reduce_tiles / complete_rows
 * split x into blocks [b*BL, min((b+1)*BL, n)) with BL = T * ceil(ceil(n/T)
 * / num_blocks) for a tile of T elements (scalar::GetWorkDistribution,
 * src/kernels/tcuscan_utils.h:1389-1408).  Every workspace starts with a
 * 4.5 KiB header of launch state; *_workspace_bytes include it.
 * --------------------------------------------------------------------- */

size_t tcs_reduce_tiles_workspace_bytes(int64_t n, int64_t tile_len, int64_t num_blocks);

/* out[b] = sum of block b (T = tile_len).  x: F16, BF16, F32 -> float32;
 * I8, I32 -> int32 (exact).  Deterministic (fixed-order partials).
 * Replaces: reduce_tiles_{fp16,int8} (src/reduce_tiles.cpp:48,70) behind
 * tcuscan::run_reduce_tiles (src/torch/torch_reduce.h:42). */
int tcs_reduce_tiles(const void* x, int x_dtype, void* out, int64_t n, int64_t tile_len,
                     int64_t num_blocks, void* workspace, size_t workspace_bytes, void* stream);

size_t tcs_complete_blocks_workspace_bytes(int64_t num_blocks);

/* y[i] = x[i] + sum(sums[:b]) with b = i / (n / num_blocks); n must be a
 * multiple of num_blocks.  dtype: F32 or I32 for x, sums and y.
 * Replaces: complete_blocks_{fp32,int32} (src/complete_blocks.cpp:56,75)
 * behind run_complete_blocks (src/torch/torch_reduce.h:197). */
int tcs_complete_blocks(const void* x, int dtype, const void* sums, int64_t num_blocks, void* y,
                        int64_t n, int64_t tile_len, void* workspace, size_t workspace_bytes,
                        void* stream);

size_t tcs_complete_rows_workspace_bytes(int64_t n, int64_t num_sums, int64_t tile_width,
                                         int64_t tile_height);

/* Down-sweep of the multi-core scan: blocks as above with T = tile_width *
 * min(tile_height, ceil(8192 / tile_width)); inside block b, rows of
 * tile_width elements get a carry starting at sum(sums[:b]), and each row's
 * last output becomes the next row's carry.  dtype: F32 or I32.
 * Replaces: complete_rows_{fp32,int32} (src/complete_rows.cpp:53,72) behind
 * run_complete_rows (src/torch/torch_reduce.h:143). */
int tcs_complete_rows(const void* x, int dtype, const void* sums, int64_t num_sums, void* y,
                      int64_t n, int64_t tile_width, int64_t tile_height, void* workspace,
                      size_t workspace_bytes, void* stream);

/* ------------------------------------------------------------------------
 * Segmented sums and SpMV
 * --------------------------------------------------------------------- */

size_t tcs_seg_sum_workspace_bytes(int64_t nnz);

/* y[r] = sum(x[end(r-1) : end(r)]) for r < num_segments, with end(-1) = 0,
 * end(r) = ends[r] for r < num_ends and nnz otherwise (num_ends <=
 * num_segments).  For a full CSR indptr pass ends = indptr + 1 and
 * num_ends = num_segments; for the start-only convention of
 * run_seg_sum_single_cube pass ends = starts + 1 and num_ends =
 * num_segments - 1.  Empty segments give 0.  ends dtype: I32, U32, I64
 * (non-decreasing).  x: F16, BF16, F32 -> float32 (double accumulation);
 * I8 -> int32 (exact).  Rows spanning tiles are completed by look-back, so
 * the result does not depend on how rows are distributed.
 * Replaces: seg_sum_single_core_{fp16,int8}, seg_sum_multi_core_{fp16,int8},
 * seg_sum_single_cube_fp16, seg_sum_multi_cube_fp16
 * (src/seg_sum_single_core.cpp:30,60, src/seg_sum_multi_core.cpp:118,150,
 * src/seg_sum_single_cube.cpp:79, src/seg_sum_multi_cube.cpp:115) behind
 * run_seg_sum_single_core / _multi_core / _single_cube / _multi_cube
 * (src/torch/torch_seg_ops.h:211,333,278,443). */
int tcs_seg_sum_indptr(const void* x, int x_dtype, const void* ends, int ends_dtype,
                       int64_t num_segments, int64_t num_ends, void* y, int64_t nnz,
                       void* workspace, size_t workspace_bytes, void* stream);

/* Workspace bytes for tcs_spmv_csr with this shape: room for the split
 * gather (vectors above 2 MiB: two look-back chains and two double partial
 * sums per row).  Any workspace of at least tcs_seg_sum_workspace_bytes(nnz)
 * is accepted; one too small for the split runs the single-chain gather,
 * with identical results. */
size_t tcs_spmv_workspace_bytes(int64_t nnz, int64_t num_rows, int64_t x_len, int vals_dtype);

/* CSR SpMV y = A x: y[r] = sum over row r of vals[p] * x[cols[p]], rows
 * delimited as in tcs_seg_sum_indptr with ends = indptr + 1 (num_rows
 * entries).  vals/x: F16 or F32 (same dtype); products are formed exactly
 * and summed in double, y is float32 (the correctly rounded row sum while
 * the exact sum fits in double).  Columns outside [0, x_len) contribute
 * exactly 0.  Workspace: tcs_spmv_workspace_bytes (or at least
 * tcs_seg_sum_workspace_bytes(nnz)).
 * Replaces: spmv_v2_{fp16,fp32} (src/spmv_v2.cpp:126,162) behind
 * run_spmv_v2 (src/torch/torch_spmv.h:112). */
int tcs_spmv_csr(const void* vals, int vals_dtype, const void* indptr_ends, int ends_dtype,
                 const int32_t* cols, const void* x, int64_t x_len, float* y, int64_t num_rows,
                 int64_t nnz, void* workspace, size_t workspace_bytes, void* stream);

/* z[i] = vals[i] * x[cols[i]] rounded to the input dtype (F16, F32: exact
 * product rounded to nearest even; I16: two's-complement wrap).  Columns
 * outside [0, x_len) give 0.
 * Replaces: csr_gather_{fp16,int16,fp32} (src/csr_gather.cpp:5,16,27)
 * behind run_csr_gather (src/torch/torch_gather.h:92). */
int tcs_csr_gather(const void* vals, int dtype, const int32_t* cols, const void* x, int64_t x_len,
                   void* z, int64_t nnz, void* stream);

/* ------------------------------------------------------------------------
 * Compaction
 * --------------------------------------------------------------------- */

size_t tcs_compress_workspace_bytes(int64_t n);

/* *count_out (device int64) = number of i with mask[i] == 1, n < 2^31.
 * Also leaves in the workspace the count of every 2048-element chunk and
 * the offset of every 64K-element block, which tcs_compress_counted reads.
 * Replaces the sizing step of run_compress / run_compress_ind /
 * _no_arange (run_reduce_tiles + .item(), src/torch/torch_compress.h:62-64). */
int tcs_count_mask(const int8_t* mask, int64_t n, int64_t* count_out, void* workspace,
                   size_t workspace_bytes, void* stream);

/* tcs_compress for a mask that tcs_count_mask(mask, n) counted last on the
 * same workspace (stream-ordered, nothing else on that workspace between
 * the two calls): output offsets come from the counts it left, so every
 * 2048-element chunk is compacted independently, without the look-back.
 * Same arguments and results as tcs_compress.
 * Replaces: compress_fp16/_fp32, compress_ind_*, compress_ind_no_arange_*
 * (src/compress.cpp:26-242) behind run_compress, run_compress_ind and
 * run_compress_ind_no_arange (src/torch/torch_compress.h:41,172,249),
 * which use the counts of their own sizing pass the same way
 * (num_ones_per_block, src/torch/torch_compress.h:196-197). */
int tcs_compress_counted(const void* x, int x_elem_bytes, const void* idx_in, int idx_mode, int idx_elem_bytes,
                         const int8_t* mask, void* out_x, void* out_idx, int64_t n, int64_t out_capacity,
                         void* workspace, size_t workspace_bytes, void* stream);

/* Stable compaction of the elements with mask[i] == 1 into out_x (element
 * size x_elem_bytes: 1, 2, 4 or 8; bits are copied).  idx_mode 0: values
 * only; 1: also out_idx[k] = idx_in[i] (idx_elem_bytes 4 or 8); 2: also
 * out_idx[k] = i (int32); 3: also out_idx[k] = i + 1 (int32).  At most
 * out_capacity elements are written.  n < 2^31.
 * Replaces: compress_{fp16,fp32}, compress_with_sums_*, compress_ind_*,
 * compress_ind_no_arange_* (src/compress.cpp:26-242) behind run_compress,
 * run_compress_pos, run_compress_ind, run_compress_ind_no_arange
 * (src/torch/torch_compress.h:41,109,172,249). */
int tcs_compress(const void* x, int x_elem_bytes, const void* idx_in, int idx_mode, int idx_elem_bytes,
                 const int8_t* mask, void* out_x, void* out_idx, int64_t n, int64_t out_capacity,
                 void* workspace, size_t workspace_bytes, void* stream);

/* ------------------------------------------------------------------------
 * Filters, where, searchsorted
 * --------------------------------------------------------------------- */

/* Comparison of x[i] with a pivot; the values are those of AscendC's CMPMODE
 * that run_count_if takes (tests/test_count_if.py:44-50). */
enum tcs_cmp {
  TCS_CMP_LT = 0,
  TCS_CMP_GT = 1,
  TCS_CMP_EQ = 2,
  TCS_CMP_LE = 3,
  TCS_CMP_GE = 4,
  TCS_CMP_NE = 5
};

/* *count_out (device int64) = number of i with (x[i] cmp pivot), compared
 * in x's type after rounding pivot to float and then to that type; a NaN
 * counts only for TCS_CMP_NE.  x: F16, BF16, F32.  Uses
 * tcs_compress_workspace_bytes(n), where it also leaves the chunk counts
 * that tcs_filter_counted reads.  Replaces count_if_fp16 behind
 * run_count_if (src/torch/torch_compare.h:33-67, src/compare.cpp:25-39) and
 * the mask + count steps of run_filter_greater_eq / run_filter_less_eq /
 * run_where (src/torch/torch_compress.h:319-398). */
int tcs_count_filter(const void* x, int x_dtype, double pivot, int cmp, int64_t n, int64_t* count_out,
                     void* workspace, size_t workspace_bytes, void* stream);

/* Stable compaction of the x[i] with x[i] >= pivot (TCS_CMP_GE) or
 * x[i] <= pivot (TCS_CMP_LE), compared as in tcs_count_filter and fused
 * into one pass (no mask in memory): values into out_vals (x's type) and/or
 * their positions into out_pos (int32); at least one must be given; at most
 * out_capacity are written; n < 2^31.  Replaces run_filter_greater_eq,
 * run_filter_less_eq (values) and run_where (positions of x >= pivot;
 * where_fp16, src/kernels/kernel_where.h) (src/torch/torch_compress.h:319-398). */
int tcs_filter(const void* x, int x_dtype, double pivot, int cmp, void* out_vals, int32_t* out_pos,
               int64_t n, int64_t out_capacity, void* workspace, size_t workspace_bytes, void* stream);

/* tcs_filter after tcs_count_filter: identical results, computed from the
 * kept count of every 2048-element chunk that tcs_count_filter (same x,
 * pivot, cmp, n) left in `workspace` just before on the same stream; each
 * chunk is compacted on its own (no look-back).  cmp: TCS_CMP_GE or
 * TCS_CMP_LE.  Replaces the compaction step of run_filter_greater_eq,
 * run_filter_less_eq and run_where (src/torch/torch_compress.h:319-398),
 * whose wrappers count first and size their outputs from the count. */
int tcs_filter_counted(const void* x, int x_dtype, double pivot, int cmp, void* out_vals, int32_t* out_pos,
                       int64_t n, int64_t out_capacity, void* workspace, size_t workspace_bytes,
                       void* stream);

/* Segmented scan from plain scans: z[i] = scan_x[i] - seg_ends[scan_f[i] - 1]
 * (z[i] = scan_x[i] where scan_f[i] == 0), one float32 subtraction each;
 * scan_x: inclusive scan of x, scan_f: inclusive scan of the segment-start
 * flags, seg_ends[j]: scan_x at the last element of segment j.  Segment ids
 * past num_seg_ends subtract nothing.  Replaces seg_scan_mc_revert behind
 * run_seg_scan_mc_revert (src/torch/torch_seg_ops.h:108-153). */
int tcs_seg_scan_revert(const float* scan_x, const int32_t* scan_f, const float* seg_ends, int64_t num_seg_ends,
                        float* z, int64_t n, void* stream);

/* Top-p kept weights over float16 probabilities sorted largest first and
 * their float32 inclusive prefix mass: kept[i] = (mass[i] - x[i] > p) ? 0 :
 * x[i] in float32 arithmetic.  All three buffers 16-byte aligned.  Replaces
 * the mask step `probs_sum - probs_sort > p; probs_sort[mask] = 0` of the
 * reference's top-p recipe (scripts/profiling/profile_tcuscan_ops.py:
 * 352-353 and 392-393), which is torch code there. */
int tcs_top_p_keep(const void* sorted_probs, const float* mass, float p, float* kept, int64_t n, void* stream);

/* ------------------------------------------------------------------------
 * Split (stable partition)
 * --------------------------------------------------------------------- */

size_t tcs_split_workspace_bytes(int64_t n);

/* out_x = [x[i] for mask[i] == 1] ++ [x[i] for mask[i] != 1], both parts in
 * input order; with idx_in (int32, may be NULL) out_idx receives idx_in
 * permuted the same way.  x: elements of 1, 2, 4 or 8 bytes; mask: int8;
 * n < 2^31; all 16-byte aligned.  Stream-ordered, no host sync.  Replaces
 * split_uint16 / split_ind_uint16 behind run_split / run_split_ind
 * (src/torch/torch_split.h:25-141). */
int tcs_split(const void* x, int x_elem_bytes, const int8_t* mask, const int32_t* idx_in, void* out_x,
              int32_t* out_idx, int64_t n, void* workspace, size_t workspace_bytes, void* stream);

/* ------------------------------------------------------------------------
 * Radix sort
 * --------------------------------------------------------------------- */

/* Workspace of tcs_sort (look-back descriptors and digit counts: zeroed
 * once, reusable like every workspace) and its scratch (one temporary copy
 * of keys and positions, any contents); 0 for an unsupported dtype. */
size_t tcs_sort_workspace_bytes(int64_t n, int dtype);
size_t tcs_sort_scratch_bytes(int64_t n, int dtype);

/* Stable sort of x (dtype F16, BF16, F32, I8, I16, I32 or U32) into out_x,
 * with out_idx[k] = the input position of out_x[k] (or idx_in[that
 * position] when idx_in is given).  Order as torch.sort(stable=True):
 * ascending, or descending when `descending` != 0; -0.0 and +0.0 compare
 * equal and NaN sorts after +inf.  Values keep their exact bits.  n < 2^31;
 * x and idx_in 16-byte aligned.  Replaces radix_sort_fp16 / radix_sort_int16
 * behind run_radix_sort (src/torch/torch_sort.h:21-81, src/radix_sort.cpp). */
int tcs_sort(const void* x, int dtype, const int32_t* idx_in, void* out_x, int32_t* out_idx, int64_t n,
             int descending, void* workspace, size_t workspace_bytes, void* scratch, size_t scratch_bytes,
             void* stream);

/* ------------------------------------------------------------------------
 * Top-k
 * --------------------------------------------------------------------- */

/* Workspace (zeroed once, reusable) and scratch (any contents) of tcs_topk;
 * 0 for an unsupported dtype. */
size_t tcs_topk_workspace_bytes(int64_t n, int64_t k, int dtype);
size_t tcs_topk_scratch_bytes(int64_t n, int64_t k, int dtype);

/* The k largest (largest != 0) or smallest elements of x, by radix select,
 * into out_vals with their input positions in out_idx, ordered largest
 * first (smallest first), ties by input position; among equal elements the
 * lowest positions are taken.  Order as tcs_sort (NaN above +inf).  dtype
 * F16, BF16, F32, I8, I16, I32 or U32; 0 <= k <= n < 2^31.  No host sync.
 * Replaces topk_fp16 / topk_int16 behind run_topk_fp16 / run_topk_int16
 * (src/torch/torch_topk.h:69-207), whose x_min / x_max bounds it does not
 * need. */
int tcs_topk(const void* x, int dtype, int64_t n, int64_t k, int largest, void* out_vals, int32_t* out_idx,
             void* workspace, size_t workspace_bytes, void* scratch, size_t scratch_bytes, void* stream);

/* Number of estimates tcs_topk_pivot writes: min(64, ceil(n / (32 *
 * num_samples))). */
int64_t tcs_topk_pivot_blocks(int64_t n, int num_samples);

/* Per block, an fp16 estimate of the k-th largest value of x (fp16): over
 * the block's first eight tiles of 32 * num_samples elements, the largest
 * tile value of "4th largest of the 8th largests of each 32" (at least 32
 * elements of that tile are >= it).  4 <= num_samples <= 64.  Replaces
 * topk_pivot_fp16 behind run_topk_pivot_fp16 (src/torch/torch_topk.h:23-60,
 * src/kernels/kernel_topk_pivot.h). */
int tcs_topk_pivot(const void* x, int64_t n, int num_samples, void* out, int64_t out_len, void* stream);

/* out[i] = first j with sorted[j] >= values[i] (sorted ascending, int32;
 * sorted_len < 2^31).  Replaces searchsorted_int32 behind run_searchsorted
 * (src/torch/torch_searchsorted.h:33-82). */
int tcs_searchsorted(const int32_t* sorted, int64_t sorted_len, const int32_t* values, int32_t* out,
                     int64_t num_values, void* stream);

/* Workspace bytes of tcs_searchsorted_indexed for a haystack of sorted_len
 * elements: a bucket table of 2^b + 1 int32, b = clamp(floor(log2(len)) - 4,
 * 10, 20) (at most 4 MiB + 4 bytes). */
size_t tcs_searchsorted_workspace_bytes(int64_t sorted_len);

/* tcs_searchsorted with a bucket index built in `workspace` first (no host
 * synchronisation): the haystack's order keys between its first and last
 * element are cut into 2^b buckets of equal width, the start of every
 * bucket is found once, and each needle then searches only its bucket
 * (about 16 elements when the keys spread over their range; clustered keys
 * fall back to the sampled search of tcs_searchsorted).  Results identical
 * to tcs_searchsorted.  `workspace` is scratch that this call overwrites
 * from its first byte (any contents before, nothing kept after): it must
 * not be a look-back workspace that other calls use, unless that is zeroed
 * again (tcs_workspace_reset) before its next look-back use.  Replaces
 * searchsorted_int32 behind run_searchsorted
 * (src/torch/torch_searchsorted.h:33-82) for large haystacks. */
int tcs_searchsorted_indexed(const int32_t* sorted, int64_t sorted_len, const int32_t* values, int32_t* out,
                             int64_t num_values, void* workspace, size_t workspace_bytes, void* stream);

/* ------------------------------------------------------------------------
 * Batched work plans
 *
 * No counterpart in the reference, whose wrappers launch and synchronise
 * per call (e.g. src/torch/torch_scan.h:107) and fix their tiling by host
 * rules (the src/tiling headers).  A plan records ops over fixed device buffers,
 * tunes them and replays them all as one HIP graph launch.  Buffers must
 * keep their addresses while the plan is used.  Each op owns a region of
 * the plan's workspace (zeroed once, like every workspace), so ops run on
 * parallel graph branches without sharing look-back state.  A plan is used
 * by one thread and one stream at a time.
 * --------------------------------------------------------------------- */

typedef struct tcs_batch tcs_batch;

int tcs_batch_create(tcs_batch** out);
void tcs_batch_destroy(tcs_batch* b);
int tcs_batch_num_ops(const tcs_batch* b);

/* Each add returns the op index (>= 0) or a negated tcs_status.  Semantics
 * are those of tcs_scan (start 0), tcs_seg_scan_flags, tcs_seg_sum_indptr and
 * tcs_spmv_csr.  Adding an op drops a captured graph. */
int tcs_batch_add_scan(tcs_batch* b, const void* x, int x_dtype, void* y, int64_t n);
int tcs_batch_add_seg_scan(tcs_batch* b, const void* x, int x_dtype, const int8_t* flags, void* y,
                           int64_t n);
int tcs_batch_add_seg_sum(tcs_batch* b, const void* x, int x_dtype, const void* ends, int ends_dtype,
                          int64_t num_segments, int64_t num_ends, void* y, int64_t nnz);
int tcs_batch_add_spmv(tcs_batch* b, const void* vals, int vals_dtype, const void* indptr_ends,
                       int ends_dtype, const int32_t* cols, const void* x, int64_t x_len, float* y,
                       int64_t num_rows, int64_t nnz);

/* Workspace bytes of the whole plan (every op's region). */
size_t tcs_batch_workspace_bytes(const tcs_batch* b);

/* Autotuner.  For every scan op, times each candidate tiling -- tiles of
 * S*S*tiles_per_wg elements for S in {32, 64, 128} and tiles_per_wg in
 * {1, ..., 32} (distinct tiles only), the local scan method (DPP, and MFMA
 * for plain F16/I8 scans) and look-back windows {6, 8, 16, 32} -- on the op's
 * own buffers with HIP events (one warm-up, then `reps` timed runs), and
 * keeps the fastest.  Results are cached per (op kind, dtype, ceil(log2 n))
 * for the process; `force` re-tunes cached keys.  Synchronises `stream`. */
int tcs_batch_tune(tcs_batch* b, void* workspace, size_t workspace_bytes, int reps, int force,
                   void* stream);

/* The tuner's cache of this process, to persist it across processes:
 * entry i (0 <= i < tcs_batch_tuned_count(), in key order) is the best
 * configuration found for scans of kind `kind` (0 = plain scan, 1 =
 * flag-segmented scan), dtype `dtype` and ceil(log2 n) = log2n.
 * tcs_batch_tuned_set adds or replaces one (validated as a candidate); ops
 * added to a plan afterwards, and tcs_batch_tune without `force`, use it.
 * No counterpart in the reference, whose tiling comes from fixed host
 * heuristics (src/tiling/heuristics/). */
int tcs_batch_tuned_count(void);
int tcs_batch_tuned_entry(int i, int* kind, int* dtype, int* log2n, int* S, int* tiles_per_wg, int* method,
                          int* window, float* ms);
int tcs_batch_tuned_set(int kind, int dtype, int log2n, int S, int tiles_per_wg, int method, int window,
                        float ms);

/* Configuration of op `op` (S, tiles_per_wg, method, window; ms = the
 * tuner's best time, -1 if untuned). */
int tcs_batch_op_config(const tcs_batch* b, int op, int* S, int* tiles_per_wg, int* method,
                        int* window, float* ms);

/* Runs every op in order on `stream` without a graph. */
int tcs_batch_run_eager(tcs_batch* b, void* workspace, size_t workspace_bytes, void* stream);

/* Captures every op into one HIP graph (on a stream of the plan's own;
 * `stream` is synchronised first): ops are dealt round-robin to up to
 * max_branches parallel branches forked from and joined to the origin. */
int tcs_batch_instantiate(tcs_batch* b, int max_branches, void* workspace, size_t workspace_bytes,
                          void* stream);

/* One launch of the captured graph on `stream`. */
int tcs_batch_launch(tcs_batch* b, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* TCUSCAN_H_ */
