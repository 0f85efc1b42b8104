// Host/device helpers shared by the tcuscan kernel sources: element codecs,
// vectorised tile loads/stores, launch configuration and status plumbing.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#include "tcs_device.hpp"
#include "../../include/tcuscan.h"

namespace tcs {

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// Codecs: storage type, accumulator type and output type of each input dtype.
// fp16/bf16/fp32 accumulate in double and emit float; int8/int32 accumulate
// and emit int32 (two's-complement wrap, as torch.cumsum(dtype=int32)).
// ---------------------------------------------------------------------------
struct CodecF16 {
  using In = uint16_t; using Acc = double; using Out = float;
  __device__ static Acc conv(In v) { return to_acc_f(v, F16{}); }
  __device__ static Out out(Acc a) { return (float)a; }
};
struct CodecBF16 {
  using In = uint16_t; using Acc = double; using Out = float;
  __device__ static Acc conv(In v) { return to_acc_f(v, BF16{}); }
  __device__ static Out out(Acc a) { return (float)a; }
};
struct CodecF32 {
  using In = float; using Acc = double; using Out = float;
  __device__ static Acc conv(In v) { return (double)v; }
  __device__ static Out out(Acc a) { return (float)a; }
};
struct CodecI8 {
  using In = int8_t; using Acc = int; using Out = int;
  __device__ static Acc conv(In v) { return (int)v; }
  __device__ static Out out(Acc a) { return a; }
};
struct CodecI32 {
  using In = int; using Acc = int; using Out = int;
  __device__ static Acc conv(In v) { return v; }
  __device__ static Out out(Acc a) { return a; }
};

// ---------------------------------------------------------------------------
// Per-thread contiguous loads/stores of ITEMS elements.  Full tiles use
// 16-byte vector accesses; the ragged last tile falls back to guarded scalar
// accesses.  Pointers handed to the kernels are 16-byte aligned (checked on
// the host).
// ---------------------------------------------------------------------------
template <class T, int ITEMS>
__device__ __forceinline__ void load_items(const T* __restrict__ p, int64_t base, int64_t n, T (&r)[ITEMS]) {
  constexpr int kBytes = ITEMS * (int)sizeof(T);
  if (base + ITEMS <= n) {
    if constexpr (kBytes % 16 == 0) {
      const u32x4* vp = reinterpret_cast<const u32x4*>(p + base);
      u32x4 tmp[kBytes / 16];
#pragma unroll
      for (int k = 0; k < kBytes / 16; ++k) tmp[k] = __builtin_nontemporal_load(vp + k);
      __builtin_memcpy(r, tmp, kBytes);
    } else {
#pragma unroll
      for (int i = 0; i < ITEMS; ++i) r[i] = p[base + i];
    }
  } else {
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) r[i] = (base + i < n) ? p[base + i] : T(0);
  }
}

template <class T, int ITEMS>
__device__ __forceinline__ void store_items(T* __restrict__ p, int64_t base, int64_t n, const T (&r)[ITEMS]) {
  constexpr int kBytes = ITEMS * (int)sizeof(T);
  if (base + ITEMS <= n) {
    if constexpr (kBytes % 16 == 0) {
      u32x4 tmp[kBytes / 16];
      __builtin_memcpy(tmp, r, kBytes);
      u32x4* vp = reinterpret_cast<u32x4*>(p + base);
#pragma unroll
      for (int k = 0; k < kBytes / 16; ++k) __builtin_nontemporal_store(tmp[k], vp + k);
    } else {
#pragma unroll
      for (int i = 0; i < ITEMS; ++i) p[base + i] = r[i];
    }
  } else {
#pragma unroll
    for (int i = 0; i < ITEMS; ++i)
      if (base + i < n) p[base + i] = r[i];
  }
}

// ---------------------------------------------------------------------------
// Workspace layout for look-back kernels: a 256-byte header whose first
// word is the ticket/epoch word, then 16 bytes of descriptor per tile.
// ---------------------------------------------------------------------------
constexpr size_t kHeaderBytes = 256;
constexpr size_t kDescBytes = 16;

inline size_t lookback_ws_bytes(uint64_t num_tiles) {
  const size_t b = kHeaderBytes + (size_t)num_tiles * kDescBytes;
  return (b + 255) & ~(size_t)255;
}

__device__ __forceinline__ uint64_t* ws_header(void* ws) { return reinterpret_cast<uint64_t*>(ws); }
__device__ __forceinline__ uint64_t* ws_desc(void* ws) {
  return reinterpret_cast<uint64_t*>(reinterpret_cast<char*>(ws) + kHeaderBytes);
}

// Tile length (elements) for a look-back kernel: tiles_per_wg matrix tiles
// of S*S elements, clamped to the instantiated range [1024, 16384] and
// rounded down to a power of two.  tiles_per_wg <= 0 selects the default.
inline int64_t tile_elems_for(int S, int tiles_per_wg) {
  int64_t s2 = (int64_t)S * S;
  if (tiles_per_wg <= 0) tiles_per_wg = s2 >= 16384 ? 1 : (int)((16384 + s2 - 1) / s2);
  int64_t t = s2 * tiles_per_wg;
  if (t < 1024) t = 1024;
  if (t > 16384) t = 16384;
  int64_t p = 1024;
  while (p * 2 <= t) p *= 2;
  return p;
}

inline bool aligned16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

#define TCS_CHECK_LAUNCH()                                  \
  do {                                                      \
    hipError_t _e = hipGetLastError();                      \
    if (_e != hipSuccess) return TCS_ERR_LAUNCH;            \
  } while (0)

}  // namespace tcs
