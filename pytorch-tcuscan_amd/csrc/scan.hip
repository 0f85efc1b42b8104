// Single-pass inclusive scans on gfx950: plain 1-D, fixed-length rows and
// flag-delimited segments, all on one decoupled look-back kernel.
//
// Each workgroup owns one tile of BLOCK*ITEMS consecutive elements (the
// reference's tiles_per_wg matrix tiles of S*S elements).  A thread holds
// ITEMS contiguous elements loaded with 16-byte nontemporal loads, reduces
// them serially, the 64-wide wavefront scans the thread totals with DPP,
// wave 0 scans the wave totals and performs the look-back, and every thread
// then re-walks its elements adding the carry and stores them with 16-byte
// nontemporal stores.  HBM is touched exactly once per input and output
// byte plus 16 bytes of descriptor traffic per tile.
//
// Reference semantics followed: src/torch/torch_scan.h:47-241 (dtype
// routing, starting_sum), src/kernels/kernel_scan_batch.h:319-390 (row-wise
// scans), src/kernels/kernel_row_scan.h:19-46 (scan of each S-row),
// src/kernels/kernel_block_scan.h:16-45 (scan of each S*S block),
// tests/test_seg_scan_single_core.py:40-63 (f=1 starts a segment).
#include "tcs_common.hpp"

namespace tcs {

enum SegMode { kSegNone = 0, kSegFixed = 1, kSegFlags = 2 };

template <class C, int BLOCK, int ITEMS, int SEG>
__global__ __launch_bounds__(BLOCK) void scan_kernel(const typename C::In* __restrict__ x,
                                                     typename C::Out* __restrict__ y,
                                                     const int8_t* __restrict__ flags, int64_t n,
                                                     int64_t seg_len, typename C::Acc start,
                                                     void* ws, uint32_t num_tiles) {
  using In = typename C::In;
  using Acc = typename C::Acc;
  using Out = typename C::Out;
  constexpr int NW = BLOCK / kWave;
  constexpr int64_t TILE = (int64_t)BLOCK * ITEMS;
  static_assert(ITEMS <= 32, "head mask is 32 bits");

  __shared__ Acc s_wval[NW];
  __shared__ int s_wflag[NW];
  __shared__ Acc s_wexcl[NW];
  __shared__ int s_wexclf[NW];
  __shared__ Acc s_carry;
  __shared__ uint32_t s_tile, s_tag;

  const int tid = threadIdx.x;
  const int lane = __lane_id();
  const int wid = tid / kWave;

  if (tid == 0) {
    const Ticket t = draw_ticket(ws_header(ws), num_tiles);
    s_tile = t.tile;
    s_tag = t.tag;
  }
  __syncthreads();
  const uint32_t tile = s_tile;
  const uint32_t tag = s_tag;
  const int64_t tile_base = (int64_t)tile * TILE;
  const int64_t base = tile_base + (int64_t)tid * ITEMS;

  In raw[ITEMS];
  load_items<In, ITEMS>(x, base, n, raw);

  uint32_t heads = 0;
  if constexpr (SEG == kSegFixed) {
    int64_t r = base % seg_len;
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
      if (r == 0) heads |= 1u << i;
      r = (r + 1 == seg_len) ? 0 : r + 1;
    }
  } else if constexpr (SEG == kSegFlags) {
    int8_t f[ITEMS];
    load_items<int8_t, ITEMS>(flags, base, n, f);
#pragma unroll
    for (int i = 0; i < ITEMS; ++i)
      if (f[i] != 0) heads |= 1u << i;
  }

  // Thread-serial segmented reduction.
  Acc tv = Acc(0);
  int tf = 0;
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    const Acc a = C::conv(raw[i]);
    if (SEG != kSegNone && (heads >> i & 1u)) {
      tv = a;
      tf = 1;
    } else {
      tv = tv + a;
    }
  }

  // Wavefront scan of thread totals; (ef, ev) is the wave-exclusive prefix.
  int wf = tf;
  Acc wv = tv;
  if constexpr (SEG == kSegNone) {
    wv = wave_incl_scan(wv);
  } else {
    wave_seg_incl_scan(wf, wv);
  }
  const Acc ev = dpp_mov<kWaveShr1>(wv);
  const int ef = SEG == kSegNone ? 0 : dpp_i32<kWaveShr1>(wf);
  if (lane == kWave - 1) {
    s_wval[wid] = wv;
    s_wflag[wid] = wf;
  }
  __syncthreads();

  if (wid == 0) {
    Acc a = lane < NW ? s_wval[lane] : Acc(0);
    int af = lane < NW ? s_wflag[lane] : 0;
    if constexpr (SEG == kSegNone) {
      a = wave_incl_scan(a);
    } else {
      wave_seg_incl_scan(af, a);
    }
    const Acc agg = readlane(a, NW - 1);
    const int aggf = SEG == kSegNone ? 0 : readlane(af, NW - 1);
    const Acc xa = dpp_mov<kWaveShr1>(a);
    const int xf = SEG == kSegNone ? 0 : dpp_i32<kWaveShr1>(af);
    if (lane < NW) {
      s_wexcl[lane] = xa;
      s_wexclf[lane] = xf;
    }
    // Does the tile begin with a segment head?  Then no carry enters it.
    const bool first_head = SEG != kSegNone && (readlane((int)(heads & 1u), 0) != 0);
    uint64_t* desc = ws_desc(ws);
    Acc carry = Acc(0);
    if (tile == 0) {
      if constexpr (SEG == kSegNone) carry = start;
      if (lane == 0 && num_tiles > 1)
        desc_publish<Acc>(desc, 0, aggf ? agg : carry + agg, tag, kFlagIncl);
    } else if (aggf) {
      // A head inside the tile fixes the tile's carry-out on its own.
      if (lane == 0) desc_publish<Acc>(desc, tile, agg, tag, kFlagIncl);
      if (!first_head) carry = lookback<Acc>(desc, tile, tag);
    } else {
      if (lane == 0) desc_publish<Acc>(desc, tile, agg, tag, kFlagAgg);
      carry = lookback<Acc>(desc, tile, tag);
      if (lane == 0) desc_publish<Acc>(desc, tile, carry + agg, tag, kFlagIncl);
    }
    if (lane == 0) s_carry = carry;
  }
  __syncthreads();

  // Carry entering this thread: tile carry, then the waves before this one,
  // then the lanes before this one.
  Acc pre = s_carry;
  {
    const Acc wx = s_wexcl[wid];
    const int wxf = s_wexclf[wid];
    pre = wxf ? wx : pre + wx;
  }
  pre = ef ? ev : pre + ev;

  Out o[ITEMS];
  Acc run = pre;
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    const Acc a = C::conv(raw[i]);
    if (SEG != kSegNone && (heads >> i & 1u)) {
      run = a;
    } else {
      run = run + a;
    }
    o[i] = C::out(run);
  }
  store_items<Out, ITEMS>(y, base, n, o);
}

constexpr int kItems = 16;

template <class C, int SEG>
static int launch_scan_block(int block, const void* x, void* y, const int8_t* flags, int64_t n,
                             int64_t seg_len, typename C::Acc start, void* ws, uint32_t tiles,
                             hipStream_t st) {
  const auto* xi = static_cast<const typename C::In*>(x);
  auto* yo = static_cast<typename C::Out*>(y);
  switch (block) {
#define TCS_CASE(B)                                                                              \
  case B:                                                                                        \
    hipLaunchKernelGGL((scan_kernel<C, B, kItems, SEG>), dim3(tiles), dim3(B), 0, st, xi, yo,    \
                       flags, n, seg_len, start, ws, tiles);                                     \
    break;
    TCS_CASE(64)
    TCS_CASE(128)
    TCS_CASE(256)
    TCS_CASE(512)
    TCS_CASE(1024)
#undef TCS_CASE
    default:
      return TCS_ERR_INVALID_ARG;
  }
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
  const int64_t tile = tile_elems_for(S, tiles_per_wg);
  const uint64_t tiles = (uint64_t)((n + tile - 1) / tile);
  if (tiles >= 0xFFFFFFFFull) return TCS_ERR_INVALID_ARG;
  if (!ws || ws_bytes < lookback_ws_bytes(tiles)) return TCS_ERR_WORKSPACE;
  const int block = (int)(tile / kItems);
  hipStream_t st = static_cast<hipStream_t>(stream);
  const uint32_t nt = (uint32_t)tiles;
  switch (x_dtype) {
    case TCS_F16:
      return launch_scan_block<CodecF16, SEG>(block, x, y, flags, n, seg_len, (double)(float)start, ws, nt, st);
    case TCS_BF16:
      return launch_scan_block<CodecBF16, SEG>(block, x, y, flags, n, seg_len, (double)(float)start, ws, nt, st);
    case TCS_F32:
      return launch_scan_block<CodecF32, SEG>(block, x, y, flags, n, seg_len, (double)(float)start, ws, nt, st);
    case TCS_I8:
      return launch_scan_block<CodecI8, SEG>(block, x, y, flags, n, seg_len, (int)start, ws, nt, st);
    case TCS_I32:
      return launch_scan_block<CodecI32, SEG>(block, x, y, flags, n, seg_len, (int)start, ws, nt, st);
    default:
      return TCS_ERR_UNSUPPORTED_DTYPE;
  }
}

}  // namespace tcs

extern "C" {

size_t tcs_scan_workspace_bytes(int64_t n, int S, int tiles_per_wg) {
  if (n <= 0 || S <= 0) return tcs::lookback_ws_bytes(0);
  const int64_t tile = tcs::tile_elems_for(S, tiles_per_wg);
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
