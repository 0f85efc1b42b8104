/*
 * Measurement kernels of libtcuscan_hip.so.  They have no counterpart in the
 * reference and compute nothing a caller needs: bench.py runs them to state
 * the ceilings the operators are measured against.
 */
#ifndef TCUSCAN_CALIB_H_
#define TCUSCAN_CALIB_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Elements of `out` that tcs_calib_gather_f16 writes for n columns. */
size_t tcs_calib_gather_out_elems(int64_t n, int parts);

/* The random-gather ceiling of CSR SpMV: reads n int32 column indices
 * (16-byte aligned) and gathers the fp16 x[c] of each c in [0, x_len),
 * with parts = 1 (every workgroup gathers every column) or 2 (even
 * workgroups the lower half of x, odd ones the upper half, as the split
 * SpMV does).  Writes per-thread sums to out
 * (tcs_calib_gather_out_elems(n, parts) floats). */
int tcs_calib_gather_f16(const int32_t* cols, int64_t n, const void* x, int64_t x_len, int parts, float* out,
                         void* stream);

/* The streaming ceilings of HBM: moves `bytes` (a multiple of 16, both
 * buffers 16-byte aligned, at least `bytes` long) as one 16-byte vector per
 * thread with nontemporal accesses: mode 0 reads src, 1 writes dst, 2
 * copies src to dst.  Mode 0 writes nothing (dst must still be valid). */
int tcs_calib_stream(const void* src, void* dst, size_t bytes, int mode, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* TCUSCAN_CALIB_H_ */
