// Stable binary split on gfx950: the elements with mask[i] == 1 first, in
// order, then all others, in order (the reference's run_split /
// run_split_ind, src/torch/torch_split.h:25-141, src/kernels/kernel_split.h;
// tests/test_split.py:52-66 compare the two parts with masked_select of
// mask == 1 and mask == 0).
//
// Two launches, no host sync: a count pass writes K = #(mask == 1) to the
// workspace, then one single-pass kernel places every element.  A tile's
// kept elements start at the look-back prefix C of kept counts over the
// tiles before it; its other elements start at K + (t0 - C), t0 being the
// tile's first index.  Each wave stages its kept run and its other run in
// LDS and writes both with coalesced stores.
#include "tcs_common.hpp"
#include "tcs_device.hpp"

namespace tcs {

template <int BYTES> struct SplitBits;
template <> struct SplitBits<1> { using T = uint8_t; };
template <> struct SplitBits<2> { using T = uint16_t; };
template <> struct SplitBits<4> { using T = uint32_t; };
template <> struct SplitBits<8> { using T = uint64_t; };

constexpr int kSplitNW = 8, kSplitG = 8;

// Groups per wave: about 24-32 KiB of input per tile.
template <int XB, bool IDX>
struct SplitGeom {
  static constexpr int kInBytes = XB + 1 + (IDX ? 4 : 0);
  static constexpr int J = kInBytes <= 3 ? 4 : 2;
  static constexpr int64_t kTile = (int64_t)kSplitNW * J * kWave * kSplitG;
};
constexpr int64_t kSplitTileMin = (int64_t)kSplitNW * 2 * kWave * kSplitG;  // 8192

// IDX: also permute idx_in (int32) the same way.
template <int XB, bool IDX>
__global__ __launch_bounds__(kSplitNW * kWave) void split_kernel(
    const typename SplitBits<XB>::T* __restrict__ x, const int8_t* __restrict__ mask,
    const int32_t* __restrict__ idx_in, typename SplitBits<XB>::T* __restrict__ out_x,
    int32_t* __restrict__ out_idx, const int64_t* __restrict__ kept_total, int64_t n, void* ws,
    uint32_t num_tiles) {
  using XT = typename SplitBits<XB>::T;
  using Geom = SplitGeom<XB, IDX>;
  constexpr int G = kSplitG, J = Geom::J, NW = kSplitNW;
  constexpr int64_t TILE = Geom::kTile;
  constexpr int64_t GSPAN = (int64_t)kWave * G, WSEG = GSPAN * J;
  __shared__ int s_wval[NW], s_wexcl[NW];
  __shared__ int64_t s_base;
  __shared__ LaunchTag s_tag;
  __shared__ XT s_stage[NW][WSEG];           // kept run, then the other run
  __shared__ int32_t s_istage[IDX ? NW : 1][IDX ? WSEG : 1];

  if (threadIdx.x == 0) s_tag = launch_arrive(ws);
  const int lane = __lane_id();
  const int wid = threadIdx.x / kWave;
  const uint32_t tile = blockIdx.x;
  const int64_t t0 = (int64_t)tile * TILE;
  const int64_t w0 = t0 + (int64_t)wid * WSEG;        // the wave's first element
  const int64_t wbase = w0 + (int64_t)lane * G;
  const bool full = t0 + TILE <= n;

  uint32_t mb[J];
  Packed<XT, G> xv[J];
  Packed<int32_t, G> iv[IDX ? J : 1];
#pragma unroll
  for (int j = 0; j < J; ++j) {
    const int64_t gb = wbase + j * GSPAN;
    Packed<int8_t, G> m;
    if (full) {
      load_items<int8_t, G, true>(mask, gb, n, m);
      load_items<XT, G, true>(x, gb, n, xv[j]);
      if constexpr (IDX) load_items<int32_t, G, true>(idx_in, gb, n, iv[j]);
    } else {
      load_items<int8_t, G, false>(mask, gb, n, m);
      load_items<XT, G, false>(x, gb, n, xv[j]);
      if constexpr (IDX) load_items<int32_t, G, false>(idx_in, gb, n, iv[j]);
    }
    uint32_t b = 0;
#pragma unroll
    for (int i = 0; i < G; ++i) b |= (m[i] == 1 ? 1u : 0u) << i;  // items past n load as 0
    mb[j] = b;
  }

  int gex[J];
  int wtot = 0;
#pragma unroll
  for (int j = 0; j < J; ++j) {
    const int c = __builtin_popcount(mb[j]);
    const int incl = wave_incl_scan(c);
    gex[j] = wtot + incl - c;
    wtot += readlane(incl, kWave - 1);
  }
  if (lane == 0) s_wval[wid] = wtot;
  __syncthreads();

  if (wid == 0) {
    int a = lane < NW ? s_wval[lane] : 0;
    const int incl = wave_incl_scan(a);
    const int agg = readlane(incl, NW - 1);
    if (lane < NW) s_wexcl[lane] = incl - a;
    auto fb = [&](int64_t u) {
      const int64_t u0 = u * TILE;
      const int64_t u1 = u0 + TILE < n ? u0 + TILE : n;
      int c = 0;
      for (int64_t p = u0 + lane; p < u1; p += kWave) c += mask[p] == 1 ? 1 : 0;
      return FlagVal<int>{u == 0 ? 1 : 0, wave_sum(c)};
    };
    uint64_t* desc = ws_desc(ws);
    const uint32_t tag = s_tag.tag;
    int carry = 0;
    if (tile == 0) {
      if (lane == 0 && num_tiles > 1) desc_publish<int>(desc, 0, agg, tag, kFlagIncl);
    } else {
      if (lane == 0) desc_publish<int>(desc, tile, agg, tag, kFlagAgg);
      carry = lookback<int>(desc, tile, tag, s_tag.polls, fb);
      if (lane == 0) desc_publish<int>(desc, tile, carry + agg, tag, kFlagIncl);
    }
    if (lane == 0) s_base = carry;
  }
  __syncthreads();

  // Kept: [kbase, kbase + wtot).  Others: [rbase, rbase + wvalid - wtot).
  const int64_t kbase = s_base + s_wexcl[wid];
  const int64_t rbase = kept_total[0] + (w0 - kbase);
  const int64_t wend = w0 + WSEG < n ? w0 + WSEG : n;
  const int wvalid = wend > w0 ? (int)(wend - w0) : 0;
  XT* stage = s_stage[wid];
#pragma unroll
  for (int j = 0; j < J; ++j) {
    int kpos = gex[j];
    const int e0 = j * (int)GSPAN + lane * G;  // wave-relative index of item 0
#pragma unroll
    for (int i = 0; i < G; ++i) {
      const bool kept = (mb[j] >> i) & 1u;
      const int e = e0 + i;
      // Others before e in the wave = e - kept before e.
      const int pos = kept ? kpos : wtot + (e - kpos);
      if (e < wvalid) {
        stage[pos] = xv[j][i];
        if constexpr (IDX) s_istage[wid][pos] = iv[j][i];
      }
      kpos += kept ? 1 : 0;
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  for (int k = lane; k < wvalid; k += kWave) {
    const int64_t dst = k < wtot ? kbase + k : rbase + (k - wtot);
    __builtin_nontemporal_store(stage[k], out_x + dst);
    if constexpr (IDX) __builtin_nontemporal_store(s_istage[wid][k], out_idx + dst);
  }
}

template <int XB, bool IDX>
static int launch_split(const void* x, const int8_t* mask, const int32_t* idx_in, void* out_x, int32_t* out_idx,
                        const int64_t* kept, int64_t n, void* ws, hipStream_t st) {
  constexpr int64_t kTile = SplitGeom<XB, IDX>::kTile;
  const uint32_t tiles = (uint32_t)((n + kTile - 1) / kTile);
  using XT = typename SplitBits<XB>::T;
  hipLaunchKernelGGL((split_kernel<XB, IDX>), dim3(tiles), dim3(kSplitNW * kWave), 0, st,
                     static_cast<const XT*>(x), mask, idx_in, static_cast<XT*>(out_x), out_idx, kept, n, ws,
                     tiles);
  TCS_CHECK_LAUNCH();
  return TCS_OK;
}

}  // namespace tcs

extern "C" {

size_t tcs_split_workspace_bytes(int64_t n) {
  // The count pass's partials and the split's descriptors share the space
  // after the header (they are used one after the other); K sits after it.
  return ((tcs_compress_workspace_bytes(n) + 255) & ~(size_t)255) + 256;
}

int tcs_split(const void* x, int x_elem_bytes, const int8_t* mask, const int32_t* idx_in, void* out_x,
              int32_t* out_idx, int64_t n, void* workspace, size_t workspace_bytes, void* stream) {
  using namespace tcs;
  if (n < 0 || (idx_in && !out_idx)) return TCS_ERR_INVALID_ARG;
  if (n == 0) return TCS_OK;
  if (!x || !mask || !out_x) return TCS_ERR_INVALID_ARG;
  if (!aligned16(x) || !aligned16(mask) || (idx_in && !aligned16(idx_in))) return TCS_ERR_INVALID_ARG;
  if (n >= ((int64_t)1 << 31)) return TCS_ERR_INVALID_ARG;
  if (!workspace || workspace_bytes < tcs_split_workspace_bytes(n)) return TCS_ERR_WORKSPACE;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const size_t kept_off = (tcs_compress_workspace_bytes(n) + 255) & ~(size_t)255;
  auto* kept = reinterpret_cast<int64_t*>(static_cast<char*>(workspace) + kept_off);
  const int rc = tcs_count_mask(mask, n, kept, workspace, kept_off, stream);
  if (rc != TCS_OK) return rc;
  const bool with_idx = idx_in != nullptr;
  switch (x_elem_bytes * 2 + (with_idx ? 1 : 0)) {
    case 2: return launch_split<1, false>(x, mask, nullptr, out_x, nullptr, kept, n, workspace, st);
    case 3: return launch_split<1, true>(x, mask, idx_in, out_x, out_idx, kept, n, workspace, st);
    case 4: return launch_split<2, false>(x, mask, nullptr, out_x, nullptr, kept, n, workspace, st);
    case 5: return launch_split<2, true>(x, mask, idx_in, out_x, out_idx, kept, n, workspace, st);
    case 8: return launch_split<4, false>(x, mask, nullptr, out_x, nullptr, kept, n, workspace, st);
    case 9: return launch_split<4, true>(x, mask, idx_in, out_x, out_idx, kept, n, workspace, st);
    case 16: return launch_split<8, false>(x, mask, nullptr, out_x, nullptr, kept, n, workspace, st);
    case 17: return launch_split<8, true>(x, mask, idx_in, out_x, out_idx, kept, n, workspace, st);
    default: return TCS_ERR_UNSUPPORTED_DTYPE;
  }
}

}  // extern "C"
