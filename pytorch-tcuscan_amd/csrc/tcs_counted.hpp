// Counted layouts shared by the compactions, the filters and the split.
//
// A count pass (tcs_count_mask, tcs_count_filter) leaves in the workspace
// the kept count of every 2048-element chunk and the exclusive offset of
// every 64K-element block, so the pass that follows knows where each
// chunk's output starts without a look-back: every wave handles one chunk
// on its own, with no block barrier and no waiting on other tiles.
//
// Workspace after the header: partial[np] (block counts, int64),
// block_off[np] (exclusive block offsets, int64), chunk_cnt[nch] (chunk
// counts as 64-bit words); np = ceil(n / 64K), nch = ceil(n / 2048).  Every
// word is a non-negative count below 2^31, so none reads as a look-back
// descriptor later.
#pragma once

#include <type_traits>

#include "tcs_common.hpp"

namespace tcs {

template <int BYTES> struct Bits;
template <> struct Bits<1> { using T = uint8_t; };
template <> struct Bits<2> { using T = uint16_t; };
template <> struct Bits<4> { using T = uint32_t; };
template <> struct Bits<8> { using T = uint64_t; };

// Count passes: 256-thread workgroups over 64K-element blocks.
constexpr int kCountBlock = 256;
constexpr int64_t kCountChunk = 1 << 16;
constexpr int64_t kChunk = 2048;                              // elements per wave chunk
constexpr int kChunksPerBlock = (int)(kCountChunk / kChunk);  // 32

struct CountedLayout {
  size_t partial, block_off, chunk_cnt, end;
};

inline size_t up256(size_t b) { return (b + 255) & ~(size_t)255; }

inline CountedLayout counted_layout(int64_t n) {
  const size_t np = (size_t)((n + kCountChunk - 1) / kCountChunk);
  const size_t nch = (size_t)((n + kChunk - 1) / kChunk);
  CountedLayout l;
  l.partial = kHeaderBytes;
  l.block_off = up256(l.partial + np * 8);
  l.chunk_cnt = up256(l.block_off + np * 8);
  l.end = up256(l.chunk_cnt + nch * 8);
  return l;
}

// Bit 7 of each byte of the result: that byte of v equals 1 (exact: the
// low seven bits plus 0x7f carry into bit 7 unless zero, and no carry
// leaves a byte).
__device__ __forceinline__ uint32_t ones_mask(uint32_t v) {
  const uint32_t z = v ^ 0x01010101u;  // bytes equal to 1 become 0
  return ~(((z & 0x7f7f7f7fu) + 0x7f7f7f7fu) | z | 0x7f7f7f7fu);
}

// Number of bytes equal to 1 in a 32-bit word.
__device__ __forceinline__ uint32_t ones_in_word(uint32_t v) { return (uint32_t)__builtin_popcount(ones_mask(v)); }

// Bits 0..3: bytes 0..3 of v equal 1 (a multiply gathers bits 7, 15, 23,
// 31 of ones_mask into bits 21..24).
__device__ __forceinline__ uint32_t ones_bits4(uint32_t v) {
  return ((ones_mask(v) >> 7) * 0x00204081u) >> 21 & 0xfu;
}

// Nontemporal load of G contiguous elements (a whole 8- or 16-byte multiple)
// into packed words.
template <class T, int N>
__device__ __forceinline__ void load_nt(const T* __restrict__ p, int64_t base, Packed<T, N>& r) {
  constexpr int kBytes = N * (int)sizeof(T);
  if constexpr (kBytes % 16 == 0) {
    const u32x4* vp = reinterpret_cast<const u32x4*>(p + base);
#pragma unroll
    for (int k = 0; k < kBytes / 16; ++k) {
      const u32x4 v = __builtin_nontemporal_load(vp + k);
      r.w[4 * k] = v[0];
      r.w[4 * k + 1] = v[1];
      r.w[4 * k + 2] = v[2];
      r.w[4 * k + 3] = v[3];
    }
  } else {
    static_assert(kBytes == 8, "8- or 16-byte multiples");
    const u32x2 v = __builtin_nontemporal_load(reinterpret_cast<const u32x2*>(p + base));
    r.w[0] = v[0];
    r.w[1] = v[1];
  }
}

// Orders one wave's LDS writes before its reads (the wave's lanes run in
// lockstep; the fences keep the compiler from moving accesses across).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One wave writes the run stage[0, wn) to out[base, base + wn): a head up to
// the first 128-byte line boundary of the output address (outputs need only
// be element-aligned), whole lines as 16-byte nontemporal vectors (one per
// lane per store), and a tail, head and tail with nontemporal element
// stores.  fp32 values + int32 indices at 2^28, p = 0.3: 0.529 ms with the
// vectors starting at the first 16-byte boundary (so the first and last line
// of every run's vector body were partial), 0.520 with line-aligned vectors
// and cached element stores, 0.502 with line-aligned vectors and
// nontemporal element stores.
template <class OT, class ST>
__device__ __forceinline__ void flush_run(const ST* stage, OT* __restrict__ out, int64_t base, int wn, int lane) {
  constexpr int V = 16 / (int)sizeof(OT);
  constexpr int L = 128 / (int)sizeof(OT);  // elements per line
  const uintptr_t addr = reinterpret_cast<uintptr_t>(out + base);
  const int h0 = (int)(((128u - (uint32_t)(addr & 127u)) & 127u) / sizeof(OT));
  const int h = h0 < wn ? h0 : wn;
  const int body = (wn - h) / L * L;
  for (int k = lane; k < h; k += kWave) __builtin_nontemporal_store((OT)stage[k], out + base + k);
  for (int k = lane; k * V < body; k += kWave) {
    OT v[V];
#pragma unroll
    for (int q = 0; q < V; ++q) v[q] = (OT)stage[h + k * V + q];
    u32x4 w;
    __builtin_memcpy(&w, v, 16);
    __builtin_nontemporal_store(w, reinterpret_cast<u32x4*>(out + base + h + (int64_t)k * V));
  }
  for (int t = h + body + lane; t < wn; t += kWave) __builtin_nontemporal_store((OT)stage[t], out + base + t);
}

// Lane l's share of the chunk's output offset among the kept elements: the
// kept count of chunk l of the block for the chunks before this one, and the
// block's offset on the last lane; the wave's sum (wave_sum) is the offset.
// Issued before a chunk's data loads, so its words arrive first (loads
// return in order).
__device__ __forceinline__ int64_t chunk_base_part(int64_t ch, const int64_t* __restrict__ block_off,
                                                   const uint64_t* __restrict__ chunk_cnt, int lane) {
  const int64_t blk = ch / kChunksPerBlock;
  const int before = (int)(ch - blk * kChunksPerBlock);
  int64_t part = lane < before ? (int64_t)chunk_cnt[blk * kChunksPerBlock + lane] : 0;
  if (lane == kWave - 1) part = block_off[blk];
  return part;
}

}  // namespace tcs
