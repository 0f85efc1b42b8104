// Single-pass inclusive scans on gfx950: plain 1-D, fixed-length rows and
// flag-delimited segments, all on one decoupled look-back kernel.
//
// Tiles of NW * J * 64 * G consecutive elements (the reference's
// tiles_per_wg matrix tiles of S*S elements) are handed out by a ticket
// counter to persistent workgroups of NW data waves plus one look-back wave.
// Inside a data wave the tile is cut into J groups of 64 * G elements; lane
// l holds elements [l*G, l*G + G) of each group, so every load instruction
// covers one contiguous span (1 KiB for fp16).
//
// Each workgroup runs a two-stage software pipeline.  While the look-back
// wave walks back for tile t's carry, the data waves load and reduce the
// next tile and the last of them to finish publishes that tile's aggregate
// at once (found with an LDS counter, no barrier), so no tile's aggregate
// waits on another tile's look-back.  After one barrier the data waves add
// t's carry and store t.  HBM is touched once per input and output byte
// plus 16 bytes of descriptor per tile.
//
// Reference semantics followed: src/torch/torch_scan.h:47-241 (dtype
// routing, starting_sum), src/kernels/kernel_scan_batch.h:319-390 (row-wise
// scans), src/kernels/kernel_row_scan.h:19-46 (scan of each S-row),
// src/kernels/kernel_block_scan.h:16-45 (scan of each S*S block),
// tests/test_seg_scan_single_core.py:40-63 (f=1 starts a segment).
#pragma once

#include <type_traits>

#include "tcs_common.hpp"
#include "tcs_mfma.hpp"

namespace tcs {

enum SegMode { kSegNone = 0, kSegFixed = 1, kSegFlags = 2 };

// How a data wave scans its part of a tile.  kLocDpp: lane-serial sums of
// contiguous groups joined by DPP wave scans.  kLocMfma: 1024-element units
// scanned on the matrix cores (tcs_mfma.hpp), plain scans of fp16 and int8.
enum LocalScan { kLocDpp = 0, kLocMfma = 1 };

template <class C, int NW, int J>
struct ScanGeom {
  static constexpr int G = C::G;
  static constexpr int64_t GSPAN = (int64_t)kWave * G;  // one group across the wave
  static constexpr int64_t WSEG = GSPAN * J;            // elements per data wave
  static constexpr int64_t TILE = WSEG * NW;
};

// Per-lane registers of one tile held by a data wave.
template <class C, int J>
struct TileRegs {
  Packed<typename C::In, C::G> raw[J];
  // Segment heads: bit i of group j is bit (G * j + i) % 32 of word
  // (G * j) / 32 (groups of G <= 16 items; packed so that the heads of a
  // whole tile take few registers).
  uint32_t heads[(J * C::G + 31) / 32];
  __device__ __forceinline__ uint32_t head_bits(int j) const {
    return (heads[(C::G * j) >> 5] >> ((C::G * j) & 31)) & ((1u << C::G) - 1u);
  }
};

template <class C, int SEG, int G>
__device__ __forceinline__ uint32_t group_heads(const int8_t* __restrict__ flags, int64_t gb,
                                                int64_t n, int64_t seg_len, bool full) {
  uint32_t h = 0;
  if constexpr (SEG == kSegFixed) {
    int64_t q = gb % seg_len;
#pragma unroll
    for (int i = 0; i < G; ++i) {
      if (q == 0) h |= 1u << i;
      q = (q + 1 == seg_len) ? 0 : q + 1;
    }
  } else if constexpr (SEG == kSegFlags) {
    static_assert(G % 4 == 0, "flag groups are whole 32-bit words");
    Packed<int8_t, G> f;
    if (full) {
      load_items<int8_t, G, true>(flags, gb, n, f);
    } else {
      load_items<int8_t, G, false>(flags, gb, n, f);
    }
    // Four flags per word: bit 7 of each byte becomes (byte != 0) (the low
    // seven bits plus 0x7f carry into bit 7 unless zero; no carry leaves the
    // byte), then a multiply gathers bits 7, 15, 23, 31 into bits 21..24.
#pragma unroll
    for (int w = 0; w < G / 4; ++w) {
      const uint32_t v = f.w[w];
      const uint32_t nz = ((((v & 0x7f7f7f7fu) + 0x7f7f7f7fu) | v) & 0x80808080u) >> 7;
      h |= ((nz * 0x00204081u) >> 21 & 0xfu) << (4 * w);
    }
  }
  return h;
}

// Segmented aggregate {any head, sum since the last head} of tile t,
// computed by one wavefront straight from the input (the look-back's
// forward-progress fallback).
template <class C, int NW, int J, int SEG>
__device__ FlagVal<typename C::Acc> scan_tile_aggregate(const typename C::In* __restrict__ x,
                                                        const int8_t* __restrict__ flags,
                                                        int64_t n, int64_t seg_len, int64_t t) {
  using In = typename C::In;
  using Acc = typename C::Acc;
  using Gm = ScanGeom<C, NW, J>;
  constexpr int G = Gm::G;
  constexpr bool kSeg = SEG != kSegNone;
  const int lane = __lane_id();
  Acc run = Acc(0);
  int runf = 0;
  for (int gi = 0; gi < NW * J; ++gi) {
    const int64_t gb = t * Gm::TILE + gi * Gm::GSPAN + (int64_t)lane * G;
    In r[G];
    load_items<In, G, false>(x, gb, n, r);
    uint32_t heads = group_heads<C, SEG, G>(flags, gb, n, seg_len, false);
    Acc tv = Acc(0);
    int tf = 0;
    for (int i = 0; i < G; ++i) {
      const Acc a = C::conv(r[i]);
      if (kSeg && (heads >> i & 1u) && gb + i < n) {
        tv = a;
        tf = 1;
      } else {
        tv = tv + a;
      }
    }
    if constexpr (kSeg) {
      wave_seg_incl_scan(tf, tv);
    } else {
      tv = wave_incl_scan(tv);
    }
    const Acc gt = readlane(tv, kWave - 1);
    const int gf = kSeg ? readlane(tf, kWave - 1) : 0;
    run = gf ? gt : run + gt;
    runf |= gf;
  }
  return FlagVal<Acc>{runf, run};
}

// LDS state of one tile in flight.  Two slots alternate between the tile
// being reduced and the tile being completed.
template <class C, int NW>
struct PipeSlot {
  typename C::Acc wval[NW];
  int wflag[NW];
  typename C::Acc wexcl[NW];
  int wexclf[NW];
  typename C::Acc agg;
  typename C::Acc carry;
  int aggf;
  int first_head;
  int count;  // data waves that have reduced the tile
};

struct ScanArgs {
  int64_t n;
  int64_t seg_len;
  uint32_t num_tiles;
  uint32_t tag;     // set in the kernel from the workspace epoch
  int window;       // look-back window (predecessors polled per step)
  int polls;        // polls before the fallback (set in the kernel from the workspace)
};

// Element offset, within a data wave's part of the tile, of the G elements
// lane `lane` holds as group j: striped groups for the DPP scan, the pieces
// of 1024-element units for the matrix-unit scan.
template <class C, int LOC>
__device__ __forceinline__ int64_t group_offset(int j, int lane) {
  if constexpr (LOC == kLocMfma) {
    constexpr int kG = UnitScan<C>::kGroups;
    return (int64_t)kUnit * (j / kG) + UnitScan<C>::piece(j % kG, lane);
  } else {
    return (int64_t)j * kWave * C::G + (int64_t)lane * C::G;
  }
}

template <class C, int NW, int J, int SEG, int LOC, bool FULL>
__device__ __forceinline__ void tile_load(TileRegs<C, J>& r, const typename C::In* __restrict__ x,
                                          const int8_t* __restrict__ flags, const ScanArgs& a,
                                          uint32_t tile, int wid, int lane) {
  using Gm = ScanGeom<C, NW, J>;
  const int64_t wave0 = (int64_t)tile * Gm::TILE + (int64_t)wid * Gm::WSEG;
#pragma unroll
  for (int j = 0; j < J; ++j)
    load_items<typename C::In, Gm::G, FULL>(x, wave0 + group_offset<C, LOC>(j, lane), a.n, r.raw[j]);
  const int64_t wbase = wave0 + (int64_t)lane * Gm::G;
#pragma unroll
  for (int w = 0; w < (J * Gm::G + 31) / 32; ++w) r.heads[w] = 0;
#pragma unroll
  for (int j = 0; j < J; ++j)
    r.heads[(Gm::G * j) >> 5] |=
        group_heads<C, SEG, Gm::G>(flags, wbase + j * Gm::GSPAN, a.n, a.seg_len, FULL) << ((Gm::G * j) & 31);
}

// Data-wave reduction of a tile: lane-serial sums, one wave scan per group,
// the wave total into the slot.  The last data wave to arrive scans the
// wave totals and publishes the tile's descriptor: inclusive for tile 0 and
// for tiles holding a segment head, aggregate otherwise.
template <class C, int NW, int J, int SEG, int LOC, class GexT>
__device__ __forceinline__ void tile_reduce(const TileRegs<C, J>& r, GexT&& gex,
                                            uint32_t& gexf, PipeSlot<C, NW>& sl,
                                            typename C::Acc start, void* ws, const ScanArgs& a,
                                            uint32_t tile, int wid, int lane) {
  using Acc = typename C::Acc;
  constexpr int G = C::G;
  constexpr bool kSeg = SEG != kSegNone;
  static_assert(LOC == kLocDpp || !kSeg, "the matrix-unit scan covers plain scans only");
  gexf = 0u;
  Acc wtot = Acc(0);
  int wtotf = 0;
  if constexpr (LOC == kLocMfma) {
    // Only the wave total is needed here; the units are scanned in the
    // output pass.
    Acc s = Acc(0);
#pragma unroll
    for (int j = 0; j < J; ++j)
#pragma unroll
      for (int i = 0; i < G; ++i) s = s + C::conv(r.raw[j][i]);
    wtot = wave_sum(s);
  }
#pragma unroll
  for (int j = 0; j < (LOC == kLocMfma ? 0 : J); ++j) {
    // Segment heads are usually rare: a group row of the wave with none
    // takes the plain scan (no per-item selects, no flag steps), which
    // performs the same additions in the same order.
    const bool seg_here = kSeg && __ballot(r.head_bits(j) != 0u) != 0ull;
    Acc tv = Acc(0);
    int wf = 0;
    Acc wv;
    if (!seg_here) {
#pragma unroll
      for (int i = 0; i < G; ++i) tv = tv + C::conv(r.raw[j][i]);
      wv = wave_incl_scan(tv);
    } else {
      int tf = 0;
#pragma unroll
      for (int i = 0; i < G; ++i) {
        const Acc v = C::conv(r.raw[j][i]);
        if (r.head_bits(j) >> i & 1u) {
          tv = v;
          tf = 1;
        } else {
          tv = tv + v;
        }
      }
      wf = tf;
      wv = tv;
      wave_seg_incl_scan(wf, wv);
    }
    const Acc ev = dpp_mov<kWaveShr1>(wv);
    const int ef = seg_here ? dpp_i32<kWaveShr1>(wf) : 0;
    gex[j] = ef ? ev : wtot + ev;
    gexf |= (uint32_t)((wtotf | ef) != 0) << j;
    const Acc gt = readlane(wv, kWave - 1);
    const int gf = seg_here ? readlane(wf, kWave - 1) : 0;
    wtot = gf ? gt : wtot + gt;
    wtotf |= gf;
  }
  int last = 0;
  if (lane == 0) {
    sl.wval[wid] = wtot;
    sl.wflag[wid] = wtotf;
    if (wid == 0) sl.first_head = kSeg ? (int)(r.heads[0] & 1u) : 0;
    last = __hip_atomic_fetch_add(&sl.count, 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP) == NW - 1;
  }
  if (readlane(last, 0)) {
    Acc v = lane < NW ? sl.wval[lane] : Acc(0);
    int vf = lane < NW ? sl.wflag[lane] : 0;
    if constexpr (kSeg) {
      wave_seg_incl_scan(vf, v);
    } else {
      v = wave_incl_scan(v);
    }
    const Acc agg = readlane(v, NW - 1);
    const int aggf = kSeg ? readlane(vf, NW - 1) : 0;
    const Acc xv = dpp_mov<kWaveShr1>(v);
    const int xf = kSeg ? dpp_i32<kWaveShr1>(vf) : 0;
    if (lane < NW) {
      sl.wexcl[lane] = xv;
      sl.wexclf[lane] = xf;
    }
    if (lane == 0) {
      sl.agg = agg;
      sl.aggf = aggf;
      sl.count = 0;
      uint64_t* desc = ws_desc(ws);
      if (tile == 0) {
        if (a.num_tiles > 1) desc_publish<Acc>(desc, 0, (kSeg && aggf) ? agg : start + agg, a.tag, kFlagIncl);
      } else {
        desc_publish<Acc>(desc, tile, agg, a.tag, aggf ? kFlagIncl : kFlagAgg);
      }
    }
  }
}

// Look-back wave: the carry entering tile t, and t's inclusive descriptor.
template <class C, int NW, int J, int SEG>
__device__ __forceinline__ void tile_lookback(PipeSlot<C, NW>& sl,
                                              const typename C::In* __restrict__ x,
                                              const int8_t* __restrict__ flags,
                                              typename C::Acc start, void* ws, const ScanArgs& a,
                                              uint32_t tile, int lane) {
  using Acc = typename C::Acc;
  constexpr bool kSeg = SEG != kSegNone;
  auto fb = [&](int64_t t) {
    FlagVal<Acc> fv = scan_tile_aggregate<C, NW, J, SEG>(x, flags, a.n, a.seg_len, t);
    if (t == 0) {
      if (!kSeg || !fv.flag) fv.value = start + fv.value;
      fv.flag = 1;
    }
    return fv;
  };
  uint64_t* desc = ws_desc(ws);
  Acc carry = Acc(0);
  if (tile == 0) {
    if constexpr (!kSeg) carry = start;
  } else if (sl.aggf) {
    if (!sl.first_head) carry = lookback<Acc>(desc, tile, a.tag, a.polls, fb, a.window);
  } else {
    carry = lookback<Acc>(desc, tile, a.tag, a.polls, fb, a.window);
    if (lane == 0) desc_publish<Acc>(desc, tile, carry + sl.agg, a.tag, kFlagIncl);
  }
  if (lane == 0) sl.carry = carry;
}

// Per-wave LDS staging of one group's outputs: with 8-wide lane groups a
// lane's 8 outputs are 32 contiguous bytes, so direct 16-byte stores leave
// every store instruction writing half of each 32-byte span; staged, each
// store instruction writes 1 KiB contiguously.
// The matrix-unit scan stages a whole 1024-element unit: its lanes hold
// runs of 4 outputs 32 elements apart.
template <class C, int LOC>
struct OutStage {
  static constexpr bool kOn = sizeof(typename C::Out) == 4 && (LOC == kLocMfma || C::G == 8);
  static constexpr int kWords = !kOn ? 1 : (LOC == kLocMfma ? kUnit : kWave * C::G);
};

// Second pass over the tile: add the carries and store.
template <class C, int NW, int J, int SEG, bool FULL, class GexT>
__device__ __forceinline__ void tile_output(TileRegs<C, J>& r, const GexT& gex,
                                            const uint32_t gexf, const PipeSlot<C, NW>& sl,
                                            typename C::Out* __restrict__ y, int64_t n,
                                            uint32_t tile, int wid, int lane,
                                            uint32_t* __restrict__ os) {
  using Acc = typename C::Acc;
  using Out = typename C::Out;
  using Gm = ScanGeom<C, NW, J>;
  constexpr int G = Gm::G;
  constexpr bool kSeg = SEG != kSegNone;
  // Re-convert the raw inputs rather than keeping the converted (double)
  // values live across the look-back.
#pragma unroll
  for (int j = 0; j < J; ++j) reg_barrier(r.raw[j]);
  Acc cin = sl.carry;
  {
    const Acc wx = sl.wexcl[wid];
    cin = sl.wexclf[wid] ? wx : cin + wx;
  }
  const int64_t wbase = (int64_t)tile * Gm::TILE + (int64_t)wid * Gm::WSEG + (int64_t)lane * G;
#pragma unroll
  for (int j = 0; j < J; ++j) {
    Acc run = (gexf >> j & 1u) ? gex[j] : cin + gex[j];
    Out o[G];
    if (!(kSeg && __ballot(r.head_bits(j) != 0u) != 0ull)) {  // as in tile_reduce
#pragma unroll
      for (int i = 0; i < G; ++i) {
        run = run + C::conv(r.raw[j][i]);
        o[i] = C::out(run);
      }
    } else {
#pragma unroll
      for (int i = 0; i < G; ++i) {
        const Acc v = C::conv(r.raw[j][i]);
        if (r.head_bits(j) >> i & 1u) {
          run = v;
        } else {
          run = run + v;
        }
        o[i] = C::out(run);
      }
    }
    if constexpr (FULL && OutStage<C, kLocDpp>::kOn) {
      u32x4 v0, v1;
      __builtin_memcpy(&v0, &o[0], 16);
      __builtin_memcpy(&v1, &o[4], 16);
      *reinterpret_cast<u32x4*>(os + lane * 8) = v0;
      *reinterpret_cast<u32x4*>(os + lane * 8 + 4) = v1;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      const u32x4 w0 = *reinterpret_cast<const u32x4*>(os + lane * 4);
      const u32x4 w1 = *reinterpret_cast<const u32x4*>(os + kWave * 4 + lane * 4);
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      // Whole lines, written once and not read again by this kernel:
      // nontemporal stores (5.60 -> 5.88 TB/s fp16, 5.44 -> 6.56 int8 at 2^28).
      Out* gy = y + (wbase - (int64_t)lane * G) + j * Gm::GSPAN;
      __builtin_nontemporal_store(w0, reinterpret_cast<u32x4*>(gy + lane * 4));
      __builtin_nontemporal_store(w1, reinterpret_cast<u32x4*>(gy + kWave * 4 + lane * 4));
    } else {
      store_items<Out, G, FULL>(y, wbase + j * Gm::GSPAN, n, o);
    }
  }
}

// Second pass of the matrix-unit scan: each pair of groups is one
// 1024-element unit.  The unit's 32 row scans come from the matrix cores;
// a half-wave scan of the 32 row totals gives each row's offset, and the
// running wave carry enters every output in double (int32 for int8).
template <class C, int NW, int J, bool FULL>
__device__ __forceinline__ void tile_output_mfma(TileRegs<C, J>& r, const PipeSlot<C, NW>& sl,
                                                 typename C::Out* __restrict__ y, int64_t n,
                                                 uint32_t tile, int wid, int lane,
                                                 uint16_t* __restrict__ stage,
                                                 uint32_t* __restrict__ os) {
  using Acc = typename C::Acc;
  using Out = typename C::Out;
  using Gm = ScanGeom<C, NW, J>;
  constexpr int kG = UnitScan<C>::kGroups;
  static_assert(J % kG == 0, "a unit is kG lane groups");
#pragma unroll
  for (int j = 0; j < J; ++j) reg_barrier(r.raw[j]);
  Acc run = sl.carry + sl.wexcl[wid];
  const int64_t wave0 = (int64_t)tile * Gm::TILE + (int64_t)wid * Gm::WSEG;
  const int row = lane & 31;
#pragma unroll
  for (int u = 0; u < J / kG; ++u) {
    Acc v[16];
    if constexpr (kG == 2) {
      UnitScan<C>::scan(r.raw[2 * u], r.raw[2 * u + 1], v, stage, lane);
    } else {
      UnitScan<C>::scan(r.raw[u], v, stage, lane);
    }
    // Row total of row `row`: accumulator 15 of lane row + 32 (m = 31).
    const Acc tot = __shfl(v[15], row + 32);
    const Acc incl = half_wave_incl_scan(tot);
    // Exclusive row offset by a one-lane shift (not incl - tot, which turns
    // an infinite row total into NaN).
    const Acc sh = dpp_mov<kWaveShr1>(incl);
    const Acc base = run + (row == 0 ? Acc(0) : sh);
    const int64_t ub = wave0 + (int64_t)kUnit * u;
    if constexpr (FULL) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        Out o[4];
#pragma unroll
        for (int s = 0; s < 4; ++s) o[s] = C::out(base + v[4 * q + s]);
        u32x4 w;
        __builtin_memcpy(&w, o, 16);
        *reinterpret_cast<u32x4*>(os + unit_out_pos(q, lane)) = w;
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      u32x4 w[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) w[k] = *reinterpret_cast<const u32x4*>(os + k * kWave * 4 + lane * 4);
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
      for (int k = 0; k < 4; ++k)
        __builtin_nontemporal_store(w[k], reinterpret_cast<u32x4*>(y + ub + k * kWave * 4 + lane * 4));
    } else {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        Out o[4];
#pragma unroll
        for (int s = 0; s < 4; ++s) o[s] = C::out(base + v[4 * q + s]);
        store_items<Out, 4, FULL>(y, ub + unit_out_pos(q, lane), n, o);
      }
    }
    run = run + readlane(incl, 31);
    // Keep each unit's product, offsets and stores together: hoisting the
    // next units' MFMAs above them would hold their 16 accumulators live.
    __builtin_amdgcn_sched_barrier(0);
  }
}

template <class C, int NW, int J, int SEG, int LOC>
__device__ __forceinline__ void load_any(TileRegs<C, J>& r, const typename C::In* __restrict__ x,
                                         const int8_t* __restrict__ flags, const ScanArgs& a,
                                         uint32_t tile, int wid, int lane) {
  if ((int64_t)(tile + 1) * ScanGeom<C, NW, J>::TILE <= a.n) {
    tile_load<C, NW, J, SEG, LOC, true>(r, x, flags, a, tile, wid, lane);
  } else {
    tile_load<C, NW, J, SEG, LOC, false>(r, x, flags, a, tile, wid, lane);
  }
}

template <class C, int NW, int J, int SEG, int LOC, class GexT>
__device__ __forceinline__ void output_any(TileRegs<C, J>& r, const GexT& gex,
                                           const uint32_t gexf, const PipeSlot<C, NW>& sl,
                                           typename C::Out* __restrict__ y, const ScanArgs& a,
                                           uint32_t tile, int wid, int lane,
                                           uint16_t* __restrict__ stage, uint32_t* __restrict__ os) {
  const bool full = (int64_t)(tile + 1) * ScanGeom<C, NW, J>::TILE <= a.n;
  if constexpr (LOC == kLocMfma) {
    if (full) {
      tile_output_mfma<C, NW, J, true>(r, sl, y, a.n, tile, wid, lane, stage, os);
    } else {
      tile_output_mfma<C, NW, J, false>(r, sl, y, a.n, tile, wid, lane, stage, os);
    }
  } else if (full) {
    tile_output<C, NW, J, SEG, true>(r, gex, gexf, sl, y, a.n, tile, wid, lane, os);
  } else {
    tile_output<C, NW, J, SEG, false>(r, gex, gexf, sl, y, a.n, tile, wid, lane, os);
  }
}

// One wave's group exclusives of one pipeline stage, kept in LDS: element
// j of lane l at base[j * kWave] (base = that lane's first), so a wave
// writes and reads 512 contiguous bytes per group.
template <class Acc>
struct LdsGex {
  Acc* base;
  __device__ __forceinline__ Acc& operator[](int j) const { return base[j * kWave]; }
};

// Persistent pipelined scan.  Workgroups draw tiles from ticket counters in
// the workspace header, one tile ahead.  A single counter serialises every
// draw on one address (about 12 ns each on MI355X, which caps small tiles),
// so there are up to kTicketCounters of them: workgroup b belongs to group
// k = b % K and counter k hands out the tiles k, k + K, k + 2K, ...
// (consecutive workgroups land on different XCDs, so each XCD mostly draws
// from its own counter).  A workgroup only holds tiles while it runs and
// publishes each held tile's aggregate as soon as the tile is reduced;
// every tile below the smallest undrawn one is held by a running workgroup
// or done, so the smallest unfinished tile can always finish and every walk
// terminates (the look-back's fallback covers workgroups that are not
// resident).  A workgroup stops drawing after its first draw past its
// group's last tile, so counter k sees exactly tiles_k + wgs_k draws; the
// last one resets it for the next launch on the workspace.
template <class C, int NW, int J, int SEG, int LOC>
__global__ __launch_bounds__((NW + 1) * kWave) void scan_kernel(
    const typename C::In* __restrict__ x, typename C::Out* __restrict__ y,
    const int8_t* __restrict__ flags, typename C::Acc start, void* ws, ScanArgs a) {
  using Acc = typename C::Acc;
  constexpr int kStage = LOC == kLocMfma ? UnitScan<C>::kStageHalfs : 1;
  __shared__ PipeSlot<C, NW> slot[2];
  __shared__ uint32_t s_draw[2];
  __shared__ LaunchTag s_tag;
  __shared__ __attribute__((aligned(16))) uint16_t stage_all[NW][kStage];
  __shared__ __attribute__((aligned(16))) uint32_t ostage_all[NW][OutStage<C, LOC>::kWords];
  const int lane = __lane_id();
  const int wid = threadIdx.x / kWave;  // wave NW is the look-back wave
  const bool lb_wave = wid == NW;
  // Ticket group of this workgroup; K <= num_tiles (the grid never exceeds
  // the tile count), so every group has at least one tile and one workgroup.
  const uint32_t K = gridDim.x < (uint32_t)kTicketCounters ? gridDim.x : (uint32_t)kTicketCounters;
  const uint32_t grp = blockIdx.x % K;
  uint64_t* ctr = ws_slot(ws, (int)grp);
  const uint64_t grp_tiles = ((uint64_t)a.num_tiles - grp + K - 1) / K;
  const uint64_t grp_wgs = (gridDim.x - grp + K - 1) / K;
  const uint64_t last_draw = grp_tiles + grp_wgs - 1;
  auto draw = [&]() -> uint32_t {
    const uint64_t v = __hip_atomic_fetch_add(ctr, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (v == last_draw) __hip_atomic_store(ctr, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return v >= grp_tiles ? a.num_tiles : (uint32_t)(v * K + grp);
  };

  if (threadIdx.x < 2) {
    slot[threadIdx.x].count = 0;
  }
  if (lb_wave && lane == 0) {
    s_tag = launch_arrive(ws);
    const uint32_t d0 = draw();
    s_draw[0] = d0;
    s_draw[1] = d0 < a.num_tiles ? draw() : a.num_tiles;
  }
  __syncthreads();
  a.tag = s_tag.tag;
  a.polls = s_tag.polls;
  uint32_t t = s_draw[0];
  uint32_t tn = s_draw[1];
  if (t >= a.num_tiles) return;

  TileRegs<C, J> reg[2];
  // Each group's exclusive value from the reduce pass, kept until the
  // output pass.  Segmented scans keep them in LDS: held in registers for
  // both pipeline stages they are 2 * J doubles per lane, which is what made
  // 32K-element fp16 segmented tiles spill.
  constexpr bool kGexLds = SEG != kSegNone && LOC == kLocDpp;
  __shared__ Acc gex_sh[kGexLds ? 2 * NW * J * kWave : 1];
  Acc gex[2][kGexLds ? 1 : J];
  uint32_t gexf[2];  // bit j: group j's exclusive value starts a segment
  uint16_t* stage = stage_all[lb_wave ? 0 : wid];
  uint32_t* ostage = ostage_all[lb_wave ? 0 : wid];
  // The parity is a compile-time constant (std::integral_constant), so
  // every index into the two-stage arrays stays static: a runtime index
  // would put them in scratch memory.
  // The parity is a compile-time constant (std::integral_constant), so
  // every index into the two-stage arrays stays static: a runtime index
  // would put them in scratch memory.
  auto gex_of = [&](auto parity) {
    constexpr int par = decltype(parity)::value;
    if constexpr (kGexLds) {
      return LdsGex<Acc>{&gex_sh[((par * NW + (lb_wave ? 0 : wid)) * J) * kWave + lane]};
    } else {
      return gex[par];
    }
  };
  auto reduce_into = [&](auto parity, uint32_t tile) {
    constexpr int par = decltype(parity)::value;
    if constexpr (kGexLds) {
      tile_reduce<C, NW, J, SEG, LOC>(reg[par], gex_of(parity), gexf[par], slot[par], start, ws, a, tile, wid, lane);
    } else {
      tile_reduce<C, NW, J, SEG, LOC>(reg[par], gex[par], gexf[par], slot[par], start, ws, a, tile, wid, lane);
    }
  };
  auto output_from = [&](auto parity, uint32_t tile) {
    constexpr int par = decltype(parity)::value;
    if constexpr (kGexLds) {
      output_any<C, NW, J, SEG, LOC>(reg[par], gex_of(parity), gexf[par], slot[par], y, a, tile, wid, lane, stage, ostage);
    } else {
      output_any<C, NW, J, SEG, LOC>(reg[par], gex[par], gexf[par], slot[par], y, a, tile, wid, lane, stage, ostage);
    }
  };
  if (!lb_wave) {
    load_any<C, NW, J, SEG, LOC>(reg[0], x, flags, a, t, wid, lane);
    reduce_into(std::integral_constant<int, 0>{}, t);
  }
  __syncthreads();

  // One pipeline step with tile t in parity P and the next tile tn going
  // into parity 1 - P.  Returns false when the workgroup has no next tile.
  auto step = [&](auto parity) -> bool {
    constexpr int P = decltype(parity)::value;
    constexpr int Q = 1 - P;
    if (lb_wave) {
      tile_lookback<C, NW, J, SEG>(slot[P], x, flags, start, ws, a, t, lane);
      if (lane == 0) s_draw[P] = tn < a.num_tiles ? draw() : a.num_tiles;
    } else if (tn < a.num_tiles) {
      load_any<C, NW, J, SEG, LOC>(reg[Q], x, flags, a, tn, wid, lane);
      reduce_into(std::integral_constant<int, Q>{}, tn);
    }
    __syncthreads();
    if (!lb_wave) output_from(std::integral_constant<int, P>{}, t);
    t = tn;
    tn = s_draw[P];
    return t < a.num_tiles;
  };
  for (;;) {
    if (!step(std::integral_constant<int, 0>{})) break;
    if (!step(std::integral_constant<int, 1>{})) break;
  }
}

// Resident workgroups of one scan_kernel instantiation on the current
// device: the persistent grid size.
template <class C, int NW, int J, int SEG, int LOC>
static uint32_t resident_grid() {
  static std::atomic<int> cached[kMaxCachedDevices];
  return (uint32_t)device_cached(cached, [](int dev) {
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, scan_kernel<C, NW, J, SEG, LOC>,
                                                     (NW + 1) * kWave, 0) != hipSuccess ||
        per_cu < 1)
      per_cu = 1;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus < 1)
      cus = 256;
    return per_cu * cus;
  });
}

// (NW, J) for a tile of `tile` elements and group width G.  The matrix-unit
// scan needs J even (a unit is two groups).
template <class C, int SEG, int LOC>
static int launch_scan_tile(int64_t tile, const void* x, void* y, const int8_t* flags, int64_t n,
                            int64_t seg_len, typename C::Acc start, void* ws, uint32_t tiles,
                            int window, hipStream_t st) {
  const auto* xi = static_cast<const typename C::In*>(x);
  auto* yo = static_cast<typename C::Out*>(y);
  constexpr int G = C::G;
  const int64_t groups = tile / (kWave * G);  // NW * J
#define TCS_LAUNCH(NW, J)                                                                        \
  do {                                                                                           \
    uint32_t grid = resident_grid<C, NW, J, SEG, LOC>();                                         \
    if (grid > tiles) grid = tiles;                                                              \
    hipLaunchKernelGGL((scan_kernel<C, NW, J, SEG, LOC>), dim3(grid), dim3((NW + 1) * kWave), 0, \
                       st, xi, yo, flags, start, ws, ScanArgs{n, seg_len, tiles, 0, window, 0});  \
  } while (0)
  if constexpr (LOC == kLocMfma && UnitScan<C>::kGroups == 1) {
    switch (groups) {
      case 1: TCS_LAUNCH(1, 1); break;
      case 2: TCS_LAUNCH(2, 1); break;
      case 4: TCS_LAUNCH(4, 1); break;
      case 8: TCS_LAUNCH(4, 2); break;
      case 16: TCS_LAUNCH(4, 4); break;
      case 32: TCS_LAUNCH(8, 4); break;
      case 64: TCS_LAUNCH(8, 8); break;
      default: return TCS_ERR_INVALID_ARG;
    }
  } else if constexpr (LOC == kLocMfma) {
    switch (groups) {
      case 2: TCS_LAUNCH(1, 2); break;
      case 4: TCS_LAUNCH(2, 2); break;
      case 8: TCS_LAUNCH(4, 2); break;
      case 16: TCS_LAUNCH(4, 4); break;
      case 32: TCS_LAUNCH(8, 4); break;
      case 64: TCS_LAUNCH(8, 8); break;
      default: return TCS_ERR_INVALID_ARG;
    }
  } else {
    switch (groups) {
      case 2: TCS_LAUNCH(2, 1); break;
      case 4: TCS_LAUNCH(4, 1); break;
      case 8: TCS_LAUNCH(4, 2); break;
      case 16: TCS_LAUNCH(4, 4); break;
      case 32: TCS_LAUNCH(8, 4); break;
      case 64: TCS_LAUNCH(8, 8); break;
      case 128:
        // 64K int8 tiles (plain scans): packed int8 leaves room for 16 groups.
        if constexpr (sizeof(typename C::In) == 1 && SEG == kSegNone) {
          TCS_LAUNCH(8, 16);
          break;
        } else {
          return TCS_ERR_INVALID_ARG;
        }
      default: return TCS_ERR_INVALID_ARG;
    }
  }
#undef TCS_LAUNCH
  TCS_CHECK_LAUNCH();
  return TCS_OK;
}

// Argument checks, tile choice and launch of one scan flavour.  The
// matrix-unit flavour (plain scans of F16 and I8 only) is compiled in its
// own translation unit and reached through launch_scan_mfma.
int launch_scan_mfma(int x_dtype, int64_t tile, const void* x, void* y, int64_t n, double start,
                     void* ws, uint32_t tiles, int window, hipStream_t st);

template <int SEG>
static int scan_dispatch(const void* x, int x_dtype, void* y, const int8_t* flags, int64_t n,
                         int64_t seg_len, int S, int tiles_per_wg, double start, void* ws,
                         size_t ws_bytes, void* stream, int method = TCS_METHOD_AUTO,
                         int window = 0) {
  if (n < 0 || S <= 0) return TCS_ERR_INVALID_ARG;
  if (window < 0 || window > kWave) return TCS_ERR_INVALID_ARG;
  if (method != TCS_METHOD_AUTO && method != TCS_METHOD_DPP && method != TCS_METHOD_MFMA)
    return TCS_ERR_INVALID_ARG;
  if (method == TCS_METHOD_MFMA &&
      (SEG != kSegNone || (x_dtype != TCS_F16 && x_dtype != TCS_BF16 && x_dtype != TCS_I8)))
    return TCS_ERR_UNSUPPORTED_DTYPE;
  // The shipped local scan is the faster one measured on MI355X at
  // N = 2^28 with the default tiles: DPP for every dtype (fp16 5.6 vs 3.7
  // TB/s on the matrix units; int8 5.1 vs 4.4).
  if (method == TCS_METHOD_AUTO) method = TCS_METHOD_DPP;
  // Default look-back window, measured the same way: 6 for DPP (fp16 265 us
  // at 6 vs 272 at 8, 284 at 16, 420 at 1; bf16 263 vs 270; int8 level from
  // 5 to 8), 32 for the matrix-unit scan.
  if (window == 0) window = method == TCS_METHOD_MFMA ? 32 : 6;
  // An integer scan's start is an int32: reject what does not convert
  // (NaN, infinities, values outside the int32 range) rather than cast it.
  if ((x_dtype == TCS_I8 || x_dtype == TCS_I32) && !(start >= -2147483648.0 && start <= 2147483647.0))
    return TCS_ERR_INVALID_ARG;
  if (n == 0) return TCS_OK;
  if (!x || !y || (SEG == kSegFlags && !flags)) return TCS_ERR_INVALID_ARG;
  if (SEG == kSegFixed && seg_len <= 0) return TCS_ERR_INVALID_ARG;
  if (!aligned16(x) || !aligned16(y) || (flags && !aligned16(flags))) return TCS_ERR_INVALID_ARG;
  // Default tile (tiles_per_wg <= 0), measured best on MI355X: 64K elements
  // for plain int8 scans (5.41 vs 5.09 TB/s at 32K), 32K for other 16-bit
  // and 8-bit inputs, 16K for 32-bit inputs (the 64-group maximum of 4-wide
  // groups); the matrix-unit flavour uses 16K for fp16 (at 32K its
  // double-buffered tiles spill), 8K for bf16 (its three band pieces spill
  // at 16K) and 32K for int8.
  const bool i8_plain_dpp = SEG == kSegNone && method == TCS_METHOD_DPP && x_dtype == TCS_I8;
  int64_t tile = tiles_per_wg > 0 ? tile_elems_for(S, tiles_per_wg)
                 : (method == TCS_METHOD_MFMA && x_dtype == TCS_F16) ? 16384
                 : (method == TCS_METHOD_MFMA && x_dtype == TCS_BF16) ? 8192
                 : (x_dtype == TCS_F32 || x_dtype == TCS_I32) ? 16384
                 : i8_plain_dpp ? 65536 : 32768;
  // Short inputs take smaller DPP tiles, so that more workgroups share
  // them (fp16 back to back, look-back window 6: 2^18 8.4 us at 8K, 8.6 at
  // 16K; 2^20 10.4 at 8K, 10.0 at 16K, 13.9 at 32K; 2^21 12.1 at 8K, 11.2 at
  // 16K; at 2^22 16K and 32K tie).
  if (tiles_per_wg <= 0 && method == TCS_METHOD_DPP) {
    if (n <= ((int64_t)1 << 18) && tile > 8192) tile = 8192;
    else if (n <= ((int64_t)1 << 23) && tile > 16384) tile = 16384;
  }
  // Lane group width: 4 for 32-bit inputs, 8 for 16-bit and int8, 16 for
  // int8 on the matrix-unit scan (one group per 1024-element unit).
  const int g = (x_dtype == TCS_F32 || x_dtype == TCS_I32) ? 4
                : (method == TCS_METHOD_MFMA && x_dtype == TCS_I8) ? 16 : 8;
  // At most 64 lane groups per tile; 128 for plain int8 DPP scans.
  const int max_groups = i8_plain_dpp ? 128 : 64;
  if (tile > max_groups * kWave * g) tile = max_groups * kWave * g;
  const uint64_t tiles = (uint64_t)((n + tile - 1) / tile);
  if (tiles >= 0xFFFFFFFFull) return TCS_ERR_INVALID_ARG;
  if (!ws || ws_bytes < lookback_ws_bytes(tiles)) return TCS_ERR_WORKSPACE;
  if (x_dtype != TCS_F16 && x_dtype != TCS_BF16 && x_dtype != TCS_F32 && x_dtype != TCS_I8 &&
      x_dtype != TCS_I32)
    return TCS_ERR_UNSUPPORTED_DTYPE;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const uint32_t nt = (uint32_t)tiles;
  if (method == TCS_METHOD_MFMA) return launch_scan_mfma(x_dtype, tile, x, y, n, start, ws, nt, window, st);
  switch (x_dtype) {
    case TCS_F16:
      return launch_scan_tile<CodecF16, SEG, kLocDpp>(tile, x, y, flags, n, seg_len, (double)(float)start, ws, nt, window, st);
    case TCS_BF16:
      return launch_scan_tile<CodecBF16, SEG, kLocDpp>(tile, x, y, flags, n, seg_len, (double)(float)start, ws, nt, window, st);
    case TCS_F32:
      return launch_scan_tile<CodecF32, SEG, kLocDpp>(tile, x, y, flags, n, seg_len, (double)(float)start, ws, nt, window, st);
    case TCS_I8:
      return launch_scan_tile<CodecI8, SEG, kLocDpp>(tile, x, y, flags, n, seg_len, (int)start, ws, nt, window, st);
    default:
      return launch_scan_tile<CodecI32, SEG, kLocDpp>(tile, x, y, flags, n, seg_len, (int)start, ws, nt, window, st);
  }
}

}  // namespace tcs
