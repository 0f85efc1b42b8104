// Host/device helpers shared by the tcuscan kernel sources: element codecs,
// vectorised tile loads/stores, launch configuration and status plumbing.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#include <atomic>

#include "tcs_device.hpp"
#include "../../include/tcuscan.h"

namespace tcs {

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// Codecs: storage type, accumulator type and output type of each input dtype.
// fp16/bf16/fp32 accumulate in double and emit float; int8/int32 accumulate
// and emit int32 (two's-complement wrap, as torch.cumsum(dtype=int32)).
// ---------------------------------------------------------------------------
// G is the group width: contiguous elements one lane loads per group
// (16 bytes of fp16/bf16/fp32 input, 8 bytes of int8).
struct CodecF16 {
  static constexpr int G = 8;
  using In = uint16_t; using Acc = double; using Out = float;
  __device__ static Acc conv(In v) { return to_acc_f(v, F16{}); }
  __device__ static Out out(Acc a) { return (float)a; }
};
struct CodecBF16 {
  static constexpr int G = 8;
  using In = uint16_t; using Acc = double; using Out = float;
  __device__ static Acc conv(In v) { return to_acc_f(v, BF16{}); }
  __device__ static Out out(Acc a) { return (float)a; }
};
struct CodecF32 {
  static constexpr int G = 4;
  using In = float; using Acc = double; using Out = float;
  __device__ static Acc conv(In v) { return (double)v; }
  __device__ static Out out(Acc a) { return (float)a; }
};
struct CodecI8 {
  static constexpr int G = 8;
  using In = int8_t; using Acc = int; using Out = int;
  __device__ static Acc conv(In v) { return (int)v; }
  __device__ static Out out(Acc a) { return a; }
};
// int8 for the matrix-unit scan: a lane group is a lane's whole share of a
// 1024-element unit (16 bytes), so one wave instruction loads 1 KiB.
struct CodecI8U {
  static constexpr int G = 16;
  using In = int8_t; using Acc = int; using Out = int;
  __device__ static Acc conv(In v) { return (int)v; }
  __device__ static Out out(Acc a) { return a; }
};
struct CodecI32 {
  static constexpr int G = 4;
  using In = int; using Acc = int; using Out = int;
  __device__ static Acc conv(In v) { return v; }
  __device__ static Out out(Acc a) { return a; }
};

// ---------------------------------------------------------------------------
// Per-lane loads/stores of ITEMS contiguous elements.  Full groups use 16-,
// 8- or 4-byte vector accesses; a ragged group falls back to guarded scalar
// accesses.  The kernels stripe these groups across the 64 lanes so that one
// wave instruction covers one contiguous span (a 16-byte group per lane ->
// 1 KiB per instruction).  Plain (temporal) stores are used: partial-line
// nontemporal stores measured 2.5x slower on gfx950 for 32-byte lane groups.
// Pointers handed to the kernels are 16-byte aligned (checked on the host).
// ---------------------------------------------------------------------------
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

// FULL = true asserts that the whole group is in range (the caller checked
// it once per tile), so only the vector path is emitted: mixing the guarded
// scalar path into the same code serialises the loads behind waitcnts.
template <class T, int ITEMS, bool FULL = false>
__device__ __forceinline__ void load_items(const T* __restrict__ p, int64_t base, int64_t n, T (&r)[ITEMS]) {
  constexpr int kBytes = ITEMS * (int)sizeof(T);
  if (FULL || base + ITEMS <= n) {
    if constexpr (kBytes % 16 == 0) {
      const u32x4* vp = reinterpret_cast<const u32x4*>(p + base);
      u32x4 tmp[kBytes / 16];
#pragma unroll
      for (int k = 0; k < kBytes / 16; ++k) tmp[k] = vp[k];
      __builtin_memcpy(r, tmp, kBytes);
    } else if constexpr (kBytes == 8) {
      const u32x2 v = *reinterpret_cast<const u32x2*>(p + base);
      __builtin_memcpy(r, &v, 8);
    } else if constexpr (kBytes == 4) {
      const uint32_t v = *reinterpret_cast<const uint32_t*>(p + base);
      __builtin_memcpy(r, &v, 4);
    } else {
#pragma unroll
      for (int i = 0; i < ITEMS; ++i) r[i] = p[base + i];
    }
  } else {
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) r[i] = (base + i < n) ? p[base + i] : T(0);
  }
}

template <class T, int ITEMS, bool FULL = false>
__device__ __forceinline__ void store_items(T* __restrict__ p, int64_t base, int64_t n, const T (&r)[ITEMS]) {
  constexpr int kBytes = ITEMS * (int)sizeof(T);
  if (FULL || base + ITEMS <= n) {
    if constexpr (kBytes % 16 == 0) {
      u32x4 tmp[kBytes / 16];
      __builtin_memcpy(tmp, r, kBytes);
      u32x4* vp = reinterpret_cast<u32x4*>(p + base);
#pragma unroll
      for (int k = 0; k < kBytes / 16; ++k) vp[k] = tmp[k];
    } else if constexpr (kBytes == 8) {
      u32x2 v;
      __builtin_memcpy(&v, r, 8);
      *reinterpret_cast<u32x2*>(p + base) = v;
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

// N elements of a 1-, 2- or 4-byte type held as packed 32-bit words.  A
// lane's inputs stay packed in VGPRs (an element array of int8 or 16-bit
// values tends to take one VGPR per element) and elements are extracted
// where they are used.
template <class T, int N>
struct Packed {
  static_assert(sizeof(T) == 1 || sizeof(T) == 2 || sizeof(T) == 4 || sizeof(T) == 8, "element size");
  static_assert((N * sizeof(T)) % 4 == 0, "whole words");
  static constexpr int kW = (int)(N * sizeof(T) / 4);
  uint32_t w[kW];
  __device__ __forceinline__ T operator[](int i) const {
    if constexpr (sizeof(T) == 8) {
      return __builtin_bit_cast(T, (uint64_t)w[2 * i] | ((uint64_t)w[2 * i + 1] << 32));
    } else if constexpr (sizeof(T) == 4) {
      return __builtin_bit_cast(T, w[i]);
    } else if constexpr (sizeof(T) == 2) {
      return __builtin_bit_cast(T, (uint16_t)(w[i >> 1] >> (16 * (i & 1))));
    } else {
      return __builtin_bit_cast(T, (uint8_t)(w[i >> 2] >> (8 * (i & 3))));
    }
  }
};

// load_items into packed words.
template <class T, int ITEMS, bool FULL = false>
__device__ __forceinline__ void load_items(const T* __restrict__ p, int64_t base, int64_t n,
                                           Packed<T, ITEMS>& r) {
  constexpr int kBytes = ITEMS * (int)sizeof(T);
  if (FULL || base + ITEMS <= n) {
    if constexpr (kBytes % 16 == 0) {
      const u32x4* vp = reinterpret_cast<const u32x4*>(p + base);
#pragma unroll
      for (int k = 0; k < kBytes / 16; ++k) {
        const u32x4 v = vp[k];
        r.w[4 * k] = v[0];
        r.w[4 * k + 1] = v[1];
        r.w[4 * k + 2] = v[2];
        r.w[4 * k + 3] = v[3];
      }
    } else if constexpr (kBytes == 8) {
      const u32x2 v = *reinterpret_cast<const u32x2*>(p + base);
      r.w[0] = v[0];
      r.w[1] = v[1];
    } else {
      r.w[0] = *reinterpret_cast<const uint32_t*>(p + base);
    }
  } else {
    T e[ITEMS];
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) e[i] = (base + i < n) ? p[base + i] : T(0);
    __builtin_memcpy(r.w, e, kBytes);
  }
}

// Marks the raw words of `r` as rewritten so the compiler recomputes values
// derived from them instead of keeping those (wider) values live in VGPRs.
template <class T, int ITEMS>
__device__ __forceinline__ void reg_barrier(T (&r)[ITEMS]) {
  constexpr int kWords = (ITEMS * (int)sizeof(T) + 3) / 4;
  uint32_t w[kWords] = {};
  __builtin_memcpy(w, r, ITEMS * sizeof(T));
#pragma unroll
  for (int k = 0; k < kWords; ++k) asm volatile("" : "+v"(w[k]));
  __builtin_memcpy(r, w, ITEMS * sizeof(T));
}

template <class T, int N>
__device__ __forceinline__ void reg_barrier(Packed<T, N>& r) {
#pragma unroll
  for (int k = 0; k < Packed<T, N>::kW; ++k) asm volatile("" : "+v"(r.w[k]));
}

// ---------------------------------------------------------------------------
// Workspace layout.  Every workspace starts with a header of 18 slots of
// 256 bytes (one word each, apart so that their atomics do not serialise on
// one address):
//   slots 0-7    ticket counters of the persistent scan;
//   slot 8       the launch epoch (see launch_arrive), and 128 bytes into
//                the slot the look-back poll limit (0: the default);
//   slot 9       groups arrived, slots 10-17: workgroups arrived per group.
// All but the epoch and the poll limit are zero between calls (each launch
// resets them), and no op puts scratch data in the header.  Look-back
// kernels keep 16 bytes of descriptor per tile right after the header;
// other scratch follows the descriptors.  One workspace serves many ops and
// sizes, so bytes that one call used as scratch a later call may read as
// descriptors.  Two rules keep such bytes from ever reading as a current
// descriptor:
//   * scratch an op leaves behind has a zero high half in every 64-bit
//     word (counts, row numbers, key digits); an op whose scratch holds
//     other values (doubles, negative integers) zeroes it before it
//     returns;
//   * a reader accepts only status words with flag 1 or 2 carrying the
//     current launch's tag, and tags repeat only after 2^30 - 1 launches
//     on one workspace, so every owner re-zeroes its workspace well before
//     that (TCS_WORKSPACE_MAX_CALLS calls, tcs_workspace_reset).
// ---------------------------------------------------------------------------
constexpr int kTicketCounters = 8;
constexpr int kSlotEpoch = 8, kSlotGroupsArrived = 9, kSlotArrived0 = 10;
constexpr size_t kSlotBytes = 256;
constexpr size_t kHeaderBytes = (kSlotArrived0 + kTicketCounters) * kSlotBytes;
constexpr size_t kDescBytes = 16;
constexpr size_t kPollWordOffset = kSlotEpoch * kSlotBytes + 128;

inline size_t lookback_ws_bytes(uint64_t num_tiles) {
  const size_t b = kHeaderBytes + (size_t)num_tiles * kDescBytes;
  return (b + 255) & ~(size_t)255;
}

__device__ __forceinline__ uint64_t* ws_slot(void* ws, int k) {
  return reinterpret_cast<uint64_t*>(reinterpret_cast<char*>(ws) + (size_t)k * kSlotBytes);
}
__device__ __forceinline__ uint64_t* ws_desc(void* ws) {
  return reinterpret_cast<uint64_t*>(reinterpret_cast<char*>(ws) + kHeaderBytes);
}

// Launch state of the current launch on workspace `ws`: its tag and the
// look-back poll limit, read by one thread of every workgroup at its start.
// The tag comes from the epoch word in the workspace, so it is right under
// HIP graph replay too (a tag passed as a kernel argument would be frozen
// into the graph).  Each workgroup then arrives on its group's counter
// (group = blockIdx.x % K); the last arrival of a group arrives on the
// groups counter, and the last group advances the epoch for the next
// launch.  Every workgroup's epoch read has returned before its arrival is
// issued (the wait below), so the advance comes after all reads.  The
// atomics are relaxed: acquire/release at agent scope would write back and
// invalidate L2 in every workgroup, which costs streaming kernels with
// thousands of workgroups several times their run time.
//
// The poll limit word holds limit + 1, or 0 for the default
// (kFallbackPolls); the host sets it with tcs_workspace_set_poll_limit.
struct LaunchTag {
  uint32_t tag;
  int polls;
};

__device__ __forceinline__ LaunchTag launch_arrive(void* ws) {
  const uint64_t e = __hip_atomic_load(ws_slot(ws, kSlotEpoch), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const uint64_t pw = __hip_atomic_load(
      reinterpret_cast<uint64_t*>(reinterpret_cast<char*>(ws) + kPollWordOffset), __ATOMIC_RELAXED,
      __HIP_MEMORY_SCOPE_AGENT);
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  const uint32_t K = gridDim.x < (uint32_t)kTicketCounters ? gridDim.x : (uint32_t)kTicketCounters;
  const uint32_t g = blockIdx.x % K;
  const uint64_t members = (gridDim.x - g + K - 1) / K;
  uint64_t* c = ws_slot(ws, kSlotArrived0 + (int)g);
  if (__hip_atomic_fetch_add(c, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == members - 1) {
    __hip_atomic_store(c, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint64_t* d = ws_slot(ws, kSlotGroupsArrived);
    if (__hip_atomic_fetch_add(d, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == K - 1) {
      __hip_atomic_store(d, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(ws_slot(ws, kSlotEpoch), e + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
  const int polls = pw == 0 ? kFallbackPolls : (pw > (uint64_t)kFallbackPolls + 1 ? kFallbackPolls : (int)(pw - 1));
  return LaunchTag{tag_of_epoch((uint32_t)(e % 0x3FFFFFFFull)), polls};
}

// Tile length (elements) for a look-back kernel: tiles_per_wg matrix tiles
// of S*S elements, clamped to the instantiated range [1024, 65536] and
// rounded down to a power of two.  tiles_per_wg <= 0 selects the default
// tile of 32768 elements: large enough that the per-tile ticket atomic
// (about 12 ns, serialised on one address) stays off the critical path.
constexpr int64_t kDefaultTile = 32768;
inline int64_t tile_elems_for(int S, int tiles_per_wg) {
  int64_t s2 = (int64_t)S * S;
  if (tiles_per_wg <= 0) tiles_per_wg = (int)((kDefaultTile + s2 - 1) / s2);
  int64_t t = s2 * tiles_per_wg;
  if (t < 1024) t = 1024;
  if (t > 65536) t = 65536;
  int64_t p = 1024;
  while (p * 2 <= t) p *= 2;
  return p;
}

inline bool aligned16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

// A launch-size figure of the current device, computed once per device by
// compute(device) and cached in `cache` (kMaxCachedDevices entries of a
// static array, zero = not yet computed).  Threads racing on a first call
// compute and store the same value; devices past the array are not cached.
constexpr int kMaxCachedDevices = 64;
template <class F>
inline int device_cached(std::atomic<int>* cache, F&& compute) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0) dev = 0;
  if (dev >= kMaxCachedDevices) return compute(dev);
  int v = cache[dev].load(std::memory_order_relaxed);
  if (v == 0) {
    v = compute(dev);
    cache[dev].store(v, std::memory_order_relaxed);
  }
  return v;
}

// Fills `bytes` (a multiple of 8, 8-byte aligned) at p with 64-bit `value`
// on the stream, by a kernel (status.hip).
int fill_words(void* p, size_t bytes, uint64_t value, hipStream_t st);

// The flag-segmented scan with an explicit look-back window (scan_flags.hip).
int seg_scan_flags_cfg(const void* x, int x_dtype, const int8_t* flags, void* y, int64_t n, int S,
                       int tiles_per_wg, int window, void* workspace, size_t workspace_bytes,
                       void* stream);

#define TCS_CHECK_LAUNCH()                                  \
  do {                                                      \
    hipError_t _e = hipGetLastError();                      \
    if (_e != hipSuccess) return TCS_ERR_LAUNCH;            \
  } while (0)

}  // namespace tcs
