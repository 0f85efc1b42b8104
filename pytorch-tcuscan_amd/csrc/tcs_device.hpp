// Device-side building blocks shared by every tcuscan HIP kernel on gfx950:
// element conversion, 64-wide wavefront scans on DPP, launch tags, and the
// decoupled look-back over tile descriptors.
//
// The look-back follows Merrill & Garland, "Single-pass Parallel Prefix Scan
// with Decoupled Look-back" (NVR-2016-002).  It replaces the three-pass
// row-scan / reduce-tiles / complete-rows pipeline of the reference
// (src/kernels/kernel_scan_multi_core.h:70-152) with one pass over HBM.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tcs {

constexpr int kWave = 64;

// ---------------------------------------------------------------------------
// Element conversion.  Floating inputs accumulate in double: every fp16 value
// is an integer multiple of 2^-24, so any partial sum of fp16 values whose
// magnitude stays below 2^29 is exact in double and the final cast to float
// is the correctly rounded prefix sum.
// ---------------------------------------------------------------------------
struct F16 { uint16_t bits; };
struct BF16 { uint16_t bits; };

__device__ __forceinline__ double to_acc_f(uint16_t h, F16) {
  return (double)(float)__builtin_bit_cast(_Float16, h);
}
__device__ __forceinline__ double to_acc_f(uint16_t h, BF16) {
  return (double)__uint_as_float((uint32_t)h << 16);
}

// ---------------------------------------------------------------------------
// DPP lane movement.  Lanes whose source is out of range, or that the
// row/bank mask disables, read the additive identity 0.
// ---------------------------------------------------------------------------
template <int CTRL, int RM = 0xf, int BM = 0xf>
__device__ __forceinline__ int dpp_i32(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, RM, BM, true);
}
template <int CTRL, int RM = 0xf, int BM = 0xf>
__device__ __forceinline__ float dpp_f32(float v) {
  return __int_as_float(dpp_i32<CTRL, RM, BM>(__float_as_int(v)));
}
template <int CTRL, int RM = 0xf, int BM = 0xf>
__device__ __forceinline__ double dpp_f64(double v) {
  const uint64_t b = __builtin_bit_cast(uint64_t, v);
  const int lo = dpp_i32<CTRL, RM, BM>((int)(uint32_t)b);
  const int hi = dpp_i32<CTRL, RM, BM>((int)(uint32_t)(b >> 32));
  return __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

template <int CTRL, int RM = 0xf, int BM = 0xf>
__device__ __forceinline__ int dpp_mov(int v) { return dpp_i32<CTRL, RM, BM>(v); }
template <int CTRL, int RM = 0xf, int BM = 0xf>
__device__ __forceinline__ uint32_t dpp_mov(uint32_t v) { return (uint32_t)dpp_i32<CTRL, RM, BM>((int)v); }
template <int CTRL, int RM = 0xf, int BM = 0xf>
__device__ __forceinline__ float dpp_mov(float v) { return dpp_f32<CTRL, RM, BM>(v); }
template <int CTRL, int RM = 0xf, int BM = 0xf>
__device__ __forceinline__ double dpp_mov(double v) { return dpp_f64<CTRL, RM, BM>(v); }
template <int CTRL, int RM = 0xf, int BM = 0xf>
__device__ __forceinline__ int64_t dpp_mov(int64_t v) {
  return __builtin_bit_cast(int64_t, dpp_f64<CTRL, RM, BM>(__builtin_bit_cast(double, v)));
}

// DPP controls used below (GFX9 encoding).
constexpr int kRowShr1 = 0x111, kRowShr2 = 0x112, kRowShr4 = 0x114, kRowShr8 = 0x118;
constexpr int kRowBcast15 = 0x142, kRowBcast31 = 0x143, kWaveShr1 = 0x138;

__device__ __forceinline__ int readlane(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ __forceinline__ float readlane(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double readlane(double v, int l);
__device__ __forceinline__ int64_t readlane(int64_t v, int l) {
  return __builtin_bit_cast(int64_t, readlane(__builtin_bit_cast(double, v), l));
}
__device__ __forceinline__ double readlane(double v, int l) {
  const uint64_t b = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, l);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), l);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// Inclusive sum scan across the 64 lanes: four row shifts inside each
// 16-lane DPP row, then two row broadcasts carry rows 0-1 into rows 1-3.
template <class T>
__device__ __forceinline__ T wave_incl_scan(T v) {
  v = v + dpp_mov<kRowShr1>(v);
  v = v + dpp_mov<kRowShr2>(v);
  v = v + dpp_mov<kRowShr4>(v);
  v = v + dpp_mov<kRowShr8>(v);
  v = v + dpp_mov<kRowBcast15, 0xa>(v);
  v = v + dpp_mov<kRowBcast31, 0xc>(v);
  return v;
}

// One step of the segmented operator (f_l, v_l) . (f_r, v_r) =
// (f_l | f_r, f_r ? v_r : v_l + v_r).
template <int CTRL, int RM = 0xf, int BM = 0xf, class T>
__device__ __forceinline__ void seg_step(int& f, T& v) {
  const T vm = dpp_mov<CTRL, RM, BM>(v);
  const int fm = dpp_i32<CTRL, RM, BM>(f);
  v = f ? v : v + vm;
  f = f | fm;
}

// Inclusive segmented scan across the wave; f != 0 marks a lane whose value
// already restarts a segment.
template <class T>
__device__ __forceinline__ void wave_seg_incl_scan(int& f, T& v) {
  seg_step<kRowShr1>(f, v);
  seg_step<kRowShr2>(f, v);
  seg_step<kRowShr4>(f, v);
  seg_step<kRowShr8>(f, v);
  seg_step<kRowBcast15, 0xa>(f, v);
  seg_step<kRowBcast31, 0xc>(f, v);
}

// Inclusive sum scan inside each half wave (lanes 0-31 and 32-63
// separately): the 64-lane scan without the row-1 -> rows-2/3 broadcast.
template <class T>
__device__ __forceinline__ T half_wave_incl_scan(T v) {
  v = v + dpp_mov<kRowShr1>(v);
  v = v + dpp_mov<kRowShr2>(v);
  v = v + dpp_mov<kRowShr4>(v);
  v = v + dpp_mov<kRowShr8>(v);
  v = v + dpp_mov<kRowBcast15, 0xa>(v);
  return v;
}

template <class T>
__device__ __forceinline__ T wave_sum(T v) {
  return readlane(wave_incl_scan(v), kWave - 1);
}

// ---------------------------------------------------------------------------
// Launch tags.
//
// Every launch on a workspace gets a tag in [1, 2^30 - 1] that differs from
// the tags of recent launches on that workspace (an epoch kept in the
// workspace advances once per launch), and descriptors carry it, so
// descriptors left by earlier launches are never mistaken for current ones
// and the workspace needs no clearing between calls.  Zeroed memory carries
// tag 0, which no launch uses.
// ---------------------------------------------------------------------------
__host__ __device__ inline uint32_t tag_of_epoch(uint32_t epoch) { return epoch % 0x3FFFFFFFu + 1u; }

// ---------------------------------------------------------------------------
// Tile descriptors.
//
// A descriptor is a pair of 64-bit words, each {payload32 | status32}; a
// 32-bit payload uses only the first word.  status = tag << 2 | flag with
// flag 1 = aggregate (the tile alone), 2 = inclusive prefix.  Each word is
// stored and loaded with a single agent-scope atomic, so a reader accepts a
// descriptor only when both words carry the same current status: a reader
// that races with the aggregate->inclusive update sees mismatched words and
// simply polls again.  No fences are needed because nothing but the
// descriptor itself is published.
// ---------------------------------------------------------------------------
constexpr uint32_t kFlagAgg = 1, kFlagIncl = 2;

// A status word is current when it carries the launch's tag and one of the
// two flags a publisher writes (a word with flag 0 or 3 is never a
// descriptor, whatever its tag bits): s - (tag << 2 | 1) is then 0 or 1,
// and any other tag or flag gives a larger unsigned difference.  One
// compare: the separate tag and flag tests cost the radix sort's per-digit
// walks 10% (2^28 fp16 keys: 3.04 against 2.76 ms per sort).
__device__ __forceinline__ bool status_current(uint32_t s, uint32_t tag) {
  return s - ((tag << 2) | kFlagAgg) <= kFlagIncl - kFlagAgg;
}

template <class T> struct DescWords { static constexpr int value = sizeof(T) == 8 ? 2 : 1; };

template <class T>
__device__ __forceinline__ void desc_publish(uint64_t* desc, uint32_t tile, T v, uint32_t tag, uint32_t flag) {
  const uint64_t status = (uint64_t)((tag << 2) | flag) << 32;
  if constexpr (sizeof(T) == 8) {
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    uint64_t* d = desc + 2 * (uint64_t)tile;
    __hip_atomic_store(d, status | (uint32_t)b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(d + 1, status | (uint32_t)(b >> 32), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  } else {
    const uint32_t b = __builtin_bit_cast(uint32_t, v);
    __hip_atomic_store(desc + tile, status | b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

template <class T>
__device__ __forceinline__ bool desc_try_read(const uint64_t* desc, int64_t tile, uint32_t tag, T& v, int& flag) {
  if constexpr (sizeof(T) == 8) {
    const uint64_t* d = desc + 2 * tile;
    const uint64_t w0 = __hip_atomic_load(d, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint64_t w1 = __hip_atomic_load(d + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t s0 = (uint32_t)(w0 >> 32), s1 = (uint32_t)(w1 >> 32);
    if (s0 != s1 || !status_current(s0, tag)) return false;
    flag = (int)(s0 & 3u);
    v = __builtin_bit_cast(T, ((w1 & 0xffffffffull) << 32) | (w0 & 0xffffffffull));
    return true;
  } else {
    const uint64_t w0 = __hip_atomic_load(desc + tile, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t s0 = (uint32_t)(w0 >> 32);
    if (!status_current(s0, tag)) return false;
    flag = (int)(s0 & 3u);
    v = __builtin_bit_cast(T, (uint32_t)w0);
    return true;
  }
}

// Exclusive prefix of `tile`: the sum of predecessor descriptors back to and
// including the nearest inclusive one.  Called by one whole wavefront; lane
// l < win (1 <= win <= 64) inspects tile - 1 - l, so every poll covers win
// predecessors (a 16-wide window measured faster than 64 on gfx950: fewer
// device-scope loads on the hot descriptors near the frontier).  Tiles before index 0
// read as an inclusive zero, which bounds the walk.
//
// Forward progress does not rely on the order in which the hardware starts
// workgroups: when a predecessor the walk needs has not published after
// kFallbackPolls polls, the wave computes that tile's aggregate itself from
// the (read-only) input through `fallback(tile) -> {flag, value}` and
// carries on, so a waiting tile can never wait on a tile that has not been
// scheduled.  The limit is kFallbackPolls unless the workspace sets a
// smaller one (tcs_workspace_set_poll_limit; 0 gives up at the first poll
// that finds a needed predecessor missing, which exercises the fallback on
// every walk).
constexpr int kFallbackPolls = 512;
constexpr int kLookbackWindow = 16;

template <class T>
struct FlagVal {
  int flag;
  T value;
};

template <class T, class Fallback>
__device__ T lookback(const uint64_t* desc, uint32_t tile, uint32_t tag, int max_polls, Fallback&& fallback,
                      int win = kLookbackWindow) {
  const int lane = __lane_id();
  T acc = T(0);
  int64_t base = (int64_t)tile - 1;
  for (;;) {
    const int64_t idx = base - lane;
    T val = T(0);
    int flag = (int)kFlagIncl;
    bool ok = true;
    if (lane >= win) {
      flag = (int)kFlagAgg;  // outside the window: contributes nothing
    } else if (idx >= 0) {
      ok = desc_try_read(desc, idx, tag, val, flag);
    }
    for (int polls = 0;; ++polls) {
      const uint64_t valid = __ballot(ok);
      const uint64_t incl = __ballot(ok && flag == (int)kFlagIncl);
      uint64_t need = ~0ull;
      int k = 63;
      if (incl) {
        k = __builtin_ctzll(incl);
        need = (k == 63) ? ~0ull : ((2ull << k) - 1ull);
      }
      if ((valid & need) == need) {
        acc = acc + wave_sum(lane <= k ? val : T(0));
        if (incl) return acc;
        break;
      }
      if (polls >= max_polls) {
        // Nearest predecessor the walk still needs.
        const int m = __builtin_ctzll(need & ~valid);
        const FlagVal<T> fv = fallback(base - m);
        if (lane == m) {
          val = fv.value;
          flag = fv.flag ? (int)kFlagIncl : (int)kFlagAgg;
          ok = true;
        }
        polls = 0;
        continue;
      }
      __builtin_amdgcn_s_sleep(1);
      if (!ok) ok = desc_try_read(desc, idx, tag, val, flag);
    }
    base -= win;
  }
}

}  // namespace tcs
