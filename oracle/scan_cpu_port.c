/*
 * CPU port of the reference's run_scan_cpu -- TEST INFRASTRUCTURE ONLY.
 *
 * Restates tcuscan::parallel_inclusive_scan_blocked and run_scan_cpu
 * (src/torch/torch_scan_cpu.h:23-107 of huawei-csl/pytorch-tcuscan):
 * the input is widened to the output type (fp16 -> float, int8 -> int32),
 * then scanned in three OpenMP phases: a sequential inclusive scan of each of
 * min(max(1, threads), n) equal blocks, a serial exclusive scan of the block
 * totals, and a parallel add of each block's offset.  Used by the tests as
 * a checker of the numpy oracle and by bench.py as the timed CPU baseline;
 * the product never links it.
 */
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static float half_to_float(uint16_t h) {
  const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
  uint32_t exp = (h >> 10) & 0x1fu;
  uint32_t man = h & 0x3ffu;
  uint32_t bits;
  if (exp == 0) {
    if (man == 0) {
      bits = sign;
    } else { /* subnormal: renormalise */
      exp = 127 - 15 + 1;
      while ((man & 0x400u) == 0) { man <<= 1; exp--; }
      man &= 0x3ffu;
      bits = sign | (exp << 23) | (man << 13);
    }
  } else if (exp == 31) {
    bits = sign | 0x7f800000u | (man << 13);
  } else {
    bits = sign | ((exp + 127 - 15) << 23) | (man << 13);
  }
  float f;
  memcpy(&f, &bits, sizeof f);
  return f;
}

#define DEFINE_BLOCKED_SCAN(NAME, T)                                         \
  static void NAME(T* a, int64_t n, int threads) {                           \
    if (n <= 0) return;                                                      \
    int64_t blocks = threads < 1 ? 1 : threads;                              \
    if (blocks > n) blocks = n;                                              \
    const int64_t bs = (n + blocks - 1) / blocks;                            \
    T* totals = (T*)calloc((size_t)blocks, sizeof(T));                       \
    T* prefix = (T*)calloc((size_t)blocks, sizeof(T));                       \
    _Pragma("omp parallel for schedule(static) num_threads(threads)")        \
    for (int64_t b = 0; b < blocks; ++b) {                                   \
      const int64_t s = b * bs;                                              \
      const int64_t e = s + bs < n ? s + bs : n;                             \
      if (s >= e) { totals[b] = (T)0; continue; }                            \
      T sum = a[s];                                                          \
      for (int64_t i = s + 1; i < e; ++i) { sum += a[i]; a[i] = sum; }       \
      totals[b] = sum;                                                       \
    }                                                                        \
    T running = (T)0;                                                        \
    for (int64_t b = 0; b < blocks; ++b) { prefix[b] = running; running += totals[b]; } \
    _Pragma("omp parallel for schedule(static) num_threads(threads)")        \
    for (int64_t b = 0; b < blocks; ++b) {                                   \
      const T off = prefix[b];                                               \
      if (off == (T)0) continue;                                             \
      const int64_t s = b * bs;                                              \
      const int64_t e = s + bs < n ? s + bs : n;                             \
      for (int64_t i = s; i < e; ++i) a[i] += off;                           \
    }                                                                        \
    free(totals);                                                            \
    free(prefix);                                                            \
  }

DEFINE_BLOCKED_SCAN(blocked_scan_f32, float)
DEFINE_BLOCKED_SCAN(blocked_scan_i32, int32_t)

/* fp16 bits -> float32 inclusive scan (run_scan_cpu for torch.float16). */
void port_scan_cpu_f16(const uint16_t* x, float* y, int64_t n, int threads) {
  _Pragma("omp parallel for schedule(static) num_threads(threads)")
  for (int64_t i = 0; i < n; ++i) y[i] = half_to_float(x[i]);
  blocked_scan_f32(y, n, threads);
}

/* float32 -> float32 inclusive scan (run_scan_cpu for torch.float32). */
void port_scan_cpu_f32(const float* x, float* y, int64_t n, int threads) {
  _Pragma("omp parallel for schedule(static) num_threads(threads)")
  for (int64_t i = 0; i < n; ++i) y[i] = x[i];
  blocked_scan_f32(y, n, threads);
}

/* int8 -> int32 inclusive scan (run_scan_cpu for torch.int8). */
void port_scan_cpu_i8(const int8_t* x, int32_t* y, int64_t n, int threads) {
  _Pragma("omp parallel for schedule(static) num_threads(threads)")
  for (int64_t i = 0; i < n; ++i) y[i] = (int32_t)x[i];
  blocked_scan_i32(y, n, threads);
}
