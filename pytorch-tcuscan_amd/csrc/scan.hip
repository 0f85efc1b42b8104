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
#include <type_traits>

#include "tcs_common.hpp"

namespace tcs {

enum SegMode { kSegNone = 0, kSegFixed = 1, kSegFlags = 2 };

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
  typename C::In raw[J][C::G];
  uint32_t heads[J];
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
    int8_t f[G];
    if (full) {
      load_items<int8_t, G, true>(flags, gb, n, f);
    } else {
      load_items<int8_t, G, false>(flags, gb, n, f);
    }
#pragma unroll
    for (int i = 0; i < G; ++i)
      if (f[i] != 0) h |= 1u << i;
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
  uint32_t tag;
};

template <class C, int NW, int J, int SEG, bool FULL>
__device__ __forceinline__ void tile_load(TileRegs<C, J>& r, const typename C::In* __restrict__ x,
                                          const int8_t* __restrict__ flags, const ScanArgs& a,
                                          uint32_t tile, int wid, int lane) {
  using Gm = ScanGeom<C, NW, J>;
  const int64_t wbase = (int64_t)tile * Gm::TILE + (int64_t)wid * Gm::WSEG + (int64_t)lane * Gm::G;
#pragma unroll
  for (int j = 0; j < J; ++j)
    load_items<typename C::In, Gm::G, FULL>(x, wbase + j * Gm::GSPAN, a.n, r.raw[j]);
#pragma unroll
  for (int j = 0; j < J; ++j)
    r.heads[j] = group_heads<C, SEG, Gm::G>(flags, wbase + j * Gm::GSPAN, a.n, a.seg_len, FULL);
}

// Data-wave reduction of a tile: lane-serial sums, one wave scan per group,
// the wave total into the slot.  The last data wave to arrive scans the
// wave totals and publishes the tile's descriptor: inclusive for tile 0 and
// for tiles holding a segment head, aggregate otherwise.
template <class C, int NW, int J, int SEG>
__device__ __forceinline__ void tile_reduce(const TileRegs<C, J>& r, typename C::Acc (&gex)[J],
                                            int (&gexf)[J], PipeSlot<C, NW>& sl,
                                            typename C::Acc start, void* ws, const ScanArgs& a,
                                            uint32_t tile, int wid, int lane) {
  using Acc = typename C::Acc;
  constexpr int G = C::G;
  constexpr bool kSeg = SEG != kSegNone;
  Acc wtot = Acc(0);
  int wtotf = 0;
#pragma unroll
  for (int j = 0; j < J; ++j) {
    Acc tv = Acc(0);
    int tf = 0;
#pragma unroll
    for (int i = 0; i < G; ++i) {
      const Acc v = C::conv(r.raw[j][i]);
      if (kSeg && (r.heads[j] >> i & 1u)) {
        tv = v;
        tf = 1;
      } else {
        tv = tv + v;
      }
    }
    int wf = tf;
    Acc wv = tv;
    if constexpr (kSeg) {
      wave_seg_incl_scan(wf, wv);
    } else {
      wv = wave_incl_scan(wv);
    }
    const Acc ev = dpp_mov<kWaveShr1>(wv);
    const int ef = kSeg ? dpp_i32<kWaveShr1>(wf) : 0;
    gex[j] = ef ? ev : wtot + ev;
    gexf[j] = wtotf | ef;
    const Acc gt = readlane(wv, kWave - 1);
    const int gf = kSeg ? readlane(wf, kWave - 1) : 0;
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
    if (!sl.first_head) carry = lookback<Acc>(desc, tile, a.tag, fb);
  } else {
    carry = lookback<Acc>(desc, tile, a.tag, fb);
    if (lane == 0) desc_publish<Acc>(desc, tile, carry + sl.agg, a.tag, kFlagIncl);
  }
  if (lane == 0) sl.carry = carry;
}

// Second pass over the tile: add the carries and store.
template <class C, int NW, int J, int SEG, bool FULL>
__device__ __forceinline__ void tile_output(TileRegs<C, J>& r, const typename C::Acc (&gex)[J],
                                            const int (&gexf)[J], const PipeSlot<C, NW>& sl,
                                            typename C::Out* __restrict__ y, int64_t n,
                                            uint32_t tile, int wid, int lane) {
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
    Acc run = gexf[j] ? gex[j] : cin + gex[j];
    Out o[G];
#pragma unroll
    for (int i = 0; i < G; ++i) {
      const Acc v = C::conv(r.raw[j][i]);
      if (kSeg && (r.heads[j] >> i & 1u)) {
        run = v;
      } else {
        run = run + v;
      }
      o[i] = C::out(run);
    }
    store_items<Out, G, FULL>(y, wbase + j * Gm::GSPAN, n, o);
  }
}

template <class C, int NW, int J, int SEG>
__device__ __forceinline__ void load_any(TileRegs<C, J>& r, const typename C::In* __restrict__ x,
                                         const int8_t* __restrict__ flags, const ScanArgs& a,
                                         uint32_t tile, int wid, int lane) {
  if ((int64_t)(tile + 1) * ScanGeom<C, NW, J>::TILE <= a.n) {
    tile_load<C, NW, J, SEG, true>(r, x, flags, a, tile, wid, lane);
  } else {
    tile_load<C, NW, J, SEG, false>(r, x, flags, a, tile, wid, lane);
  }
}

template <class C, int NW, int J, int SEG>
__device__ __forceinline__ void output_any(TileRegs<C, J>& r, const typename C::Acc (&gex)[J],
                                           const int (&gexf)[J], const PipeSlot<C, NW>& sl,
                                           typename C::Out* __restrict__ y, const ScanArgs& a,
                                           uint32_t tile, int wid, int lane) {
  if ((int64_t)(tile + 1) * ScanGeom<C, NW, J>::TILE <= a.n) {
    tile_output<C, NW, J, SEG, true>(r, gex, gexf, sl, y, a.n, tile, wid, lane);
  } else {
    tile_output<C, NW, J, SEG, false>(r, gex, gexf, sl, y, a.n, tile, wid, lane);
  }
}

// Persistent pipelined scan.  Workgroups draw tiles from a ticket counter
// in the workspace header, one tile ahead: a workgroup only holds tiles
// while it runs and publishes each held tile's aggregate as soon as the
// tile is reduced, so the smallest unfinished tile is always some running
// workgroup's current tile and every walk terminates.  A workgroup stops
// drawing after its first draw past the last tile, so there are exactly
// num_tiles + gridDim.x draws; the last one resets the counter for the next
// launch on the workspace.
template <class C, int NW, int J, int SEG>
__global__ __launch_bounds__((NW + 1) * kWave) void scan_kernel(
    const typename C::In* __restrict__ x, typename C::Out* __restrict__ y,
    const int8_t* __restrict__ flags, typename C::Acc start, void* ws, ScanArgs a) {
  using Acc = typename C::Acc;
  __shared__ PipeSlot<C, NW> slot[2];
  __shared__ uint32_t s_draw[2];
  const int lane = __lane_id();
  const int wid = threadIdx.x / kWave;  // wave NW is the look-back wave
  const bool lb_wave = wid == NW;
  uint64_t* hdr = ws_header(ws);
  const uint64_t last_draw = (uint64_t)a.num_tiles + gridDim.x - 1;
  auto draw = [&]() -> uint32_t {
    const uint64_t v = __hip_atomic_fetch_add(hdr, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (v == last_draw) __hip_atomic_store(hdr, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return v >= a.num_tiles ? a.num_tiles : (uint32_t)v;
  };

  if (threadIdx.x < 2) {
    slot[threadIdx.x].count = 0;
  }
  if (lb_wave && lane == 0) {
    const uint32_t d0 = draw();
    s_draw[0] = d0;
    s_draw[1] = d0 < a.num_tiles ? draw() : a.num_tiles;
  }
  __syncthreads();
  uint32_t t = s_draw[0];
  uint32_t tn = s_draw[1];
  if (t >= a.num_tiles) return;

  TileRegs<C, J> reg[2];
  Acc gex[2][J];
  int gexf[2][J];
  if (!lb_wave) {
    load_any<C, NW, J, SEG>(reg[0], x, flags, a, t, wid, lane);
    tile_reduce<C, NW, J, SEG>(reg[0], gex[0], gexf[0], slot[0], start, ws, a, t, wid, lane);
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
      load_any<C, NW, J, SEG>(reg[Q], x, flags, a, tn, wid, lane);
      tile_reduce<C, NW, J, SEG>(reg[Q], gex[Q], gexf[Q], slot[Q], start, ws, a, tn, wid, lane);
    }
    __syncthreads();
    if (!lb_wave) output_any<C, NW, J, SEG>(reg[P], gex[P], gexf[P], slot[P], y, a, t, wid, lane);
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
template <class C, int NW, int J, int SEG>
static uint32_t resident_grid() {
  static int cached[16] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 16) dev = 0;
  if (cached[dev] == 0) {
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, scan_kernel<C, NW, J, SEG>,
                                                     (NW + 1) * kWave, 0) != hipSuccess ||
        per_cu < 1)
      per_cu = 1;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus < 1)
      cus = 256;
    cached[dev] = per_cu * cus;
  }
  return (uint32_t)cached[dev];
}

// (NW, J) for a tile of `tile` elements and group width G.
template <class C, int SEG>
static int launch_scan_tile(int64_t tile, const void* x, void* y, const int8_t* flags, int64_t n,
                            int64_t seg_len, typename C::Acc start, void* ws, uint32_t tiles,
                            uint32_t tag, hipStream_t st) {
  const auto* xi = static_cast<const typename C::In*>(x);
  auto* yo = static_cast<typename C::Out*>(y);
  constexpr int G = C::G;
  const int64_t groups = tile / (kWave * G);  // NW * J
#define TCS_LAUNCH(NW, J)                                                                   \
  do {                                                                                      \
    uint32_t grid = resident_grid<C, NW, J, SEG>();                                         \
    if (grid > tiles) grid = tiles;                                                         \
    hipLaunchKernelGGL((scan_kernel<C, NW, J, SEG>), dim3(grid), dim3((NW + 1) * kWave), 0, \
                       st, xi, yo, flags, start, ws, ScanArgs{n, seg_len, tiles, tag});    \
  } while (0)
  switch (groups) {
    case 2: TCS_LAUNCH(2, 1); break;
    case 4: TCS_LAUNCH(4, 1); break;
    case 8: TCS_LAUNCH(4, 2); break;
    case 16: TCS_LAUNCH(4, 4); break;
    case 32: TCS_LAUNCH(8, 4); break;
    case 64: TCS_LAUNCH(8, 8); break;
    default: return TCS_ERR_INVALID_ARG;
  }
#undef TCS_LAUNCH
  TCS_CHECK_LAUNCH();
  return TCS_OK;
}

template <int SEG>
static int scan_dispatch(const void* x, int x_dtype, void* y, const int8_t* flags, int64_t n,
                         int64_t seg_len, int S, int tiles_per_wg, double start, void* ws,
                         size_t ws_bytes, void* stream) {
  if (n < 0 || S <= 0) return TCS_ERR_INVALID_ARG;
  if (n == 0) return TCS_OK;
  if (!x || !y || (SEG == kSegFlags && !flags)) return TCS_ERR_INVALID_ARG;
  if (SEG == kSegFixed && seg_len <= 0) return TCS_ERR_INVALID_ARG;
  if (!aligned16(x) || !aligned16(y) || (flags && !aligned16(flags))) return TCS_ERR_INVALID_ARG;
  // Default tile (tiles_per_wg <= 0): 32K elements for 16-bit inputs, 16K
  // otherwise (measured best on MI355X; int8 at 32K spills registers).  At
  // most 64 lane groups per tile: 32K elements for 8-wide groups, 16K for
  // 4-wide (fp32/int32) groups.
  int64_t tile = tiles_per_wg > 0 ? tile_elems_for(S, tiles_per_wg)
                                  : ((x_dtype == TCS_F16 || x_dtype == TCS_BF16) ? 32768 : 16384);
  const int g = (x_dtype == TCS_F32 || x_dtype == TCS_I32) ? 4 : 8;
  if (tile > 64 * kWave * g) tile = 64 * kWave * g;
  const uint64_t tiles = (uint64_t)((n + tile - 1) / tile);
  if (tiles >= 0xFFFFFFFFull) return TCS_ERR_INVALID_ARG;
  if (!ws || ws_bytes < lookback_ws_bytes(tiles)) return TCS_ERR_WORKSPACE;
  if (x_dtype != TCS_F16 && x_dtype != TCS_BF16 && x_dtype != TCS_F32 && x_dtype != TCS_I8 &&
      x_dtype != TCS_I32)
    return TCS_ERR_UNSUPPORTED_DTYPE;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const uint32_t nt = (uint32_t)tiles;
  const uint32_t tag = next_launch_tag(ws);
  switch (x_dtype) {
    case TCS_F16:
      return launch_scan_tile<CodecF16, SEG>(tile, x, y, flags, n, seg_len, (double)(float)start, ws, nt, tag, st);
    case TCS_BF16:
      return launch_scan_tile<CodecBF16, SEG>(tile, x, y, flags, n, seg_len, (double)(float)start, ws, nt, tag, st);
    case TCS_F32:
      return launch_scan_tile<CodecF32, SEG>(tile, x, y, flags, n, seg_len, (double)(float)start, ws, nt, tag, st);
    case TCS_I8:
      return launch_scan_tile<CodecI8, SEG>(tile, x, y, flags, n, seg_len, (int)start, ws, nt, tag, st);
    default:
      return launch_scan_tile<CodecI32, SEG>(tile, x, y, flags, n, seg_len, (int)start, ws, nt, tag, st);
  }
}

}  // namespace tcs

extern "C" {

size_t tcs_scan_workspace_bytes(int64_t n, int S, int tiles_per_wg) {
  if (n <= 0 || S <= 0) return tcs::lookback_ws_bytes(0);
  // Sized for the smallest tile any dtype uses with these parameters.
  int64_t tile = tcs::tile_elems_for(S, tiles_per_wg);
  if (tile > 16384) tile = 16384;
  return tcs::lookback_ws_bytes((uint64_t)((n + tile - 1) / tile));
}

int tcs_scan(const void* x, int x_dtype, void* y, int64_t n, int S, int tiles_per_wg, double start,
             void* workspace, size_t workspace_bytes, void* stream) {
  return tcs::scan_dispatch<tcs::kSegNone>(x, x_dtype, y, nullptr, n, 0, S, tiles_per_wg, start,
                                           workspace, workspace_bytes, stream);
}

int tcs_scan_rows(const void* x, int x_dtype, void* y, int64_t n, int64_t row_len, int S,
                  int tiles_per_wg, void* workspace, size_t workspace_bytes, void* stream) {
  return tcs::scan_dispatch<tcs::kSegFixed>(x, x_dtype, y, nullptr, n, row_len, S, tiles_per_wg,
                                            0.0, workspace, workspace_bytes, stream);
}

int tcs_seg_scan_flags(const void* x, int x_dtype, const int8_t* flags, void* y, int64_t n, int S,
                       int tiles_per_wg, void* workspace, size_t workspace_bytes, void* stream) {
  return tcs::scan_dispatch<tcs::kSegFlags>(x, x_dtype, y, flags, n, 0, S, tiles_per_wg, 0.0,
                                            workspace, workspace_bytes, stream);
}

}  // extern "C"
