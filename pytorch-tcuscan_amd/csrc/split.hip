// Stable binary split on gfx950: the elements with mask[i] == 1 first, in
// order, then all others, in order (the reference's run_split /
// run_split_ind, src/torch/torch_split.h:25-141, src/kernels/kernel_split.h;
// tests/test_split.py:52-66 compare the two parts with masked_select of
// mask == 1 and mask == 0).
//
// Two launches, no host sync: the count pass (tcs_count_mask) writes
// K = #(mask == 1) and the kept count of every 2048-element chunk to the
// workspace, then every wave of one kernel places one chunk on its own:
// the chunk's kept elements start at kb, the kept count before the chunk
// (its block's offset plus the chunk counts before it), its other elements
// at K + (c0 - kb), c0 being the chunk's first index.  The wave stages its
// kept run and then its other run in LDS and writes each with line-aligned
// 16-byte stores.
#include "tcs_counted.hpp"

namespace tcs {

// IDX: also permute idx_in (int32) the same way.
template <int XB, bool IDX>
__global__ __launch_bounds__(kCountBlock) void split_counted_kernel(
    const typename Bits<XB>::T* __restrict__ x, const int8_t* __restrict__ mask,
    const int32_t* __restrict__ idx_in, typename Bits<XB>::T* __restrict__ out_x,
    int32_t* __restrict__ out_idx, const int64_t* __restrict__ kept_total, int64_t n,
    const int64_t* __restrict__ block_off, const uint64_t* __restrict__ chunk_cnt) {
  using XT = typename Bits<XB>::T;
  constexpr int G = 8, J = (int)(kChunk / (kWave * G));  // 4 groups of 8 per lane
  constexpr int kWaves = kCountBlock / kWave;
  using ST = typename Bits<(IDX && XB < 4) ? 4 : XB>::T;
  __shared__ ST s_stage[kWaves][kChunk];  // kept run, then the other run
  const int lane = __lane_id();
  const int wid = threadIdx.x / kWave;
  // Chunks from the last one down: the count pass read the mask front to
  // back, so the caches hold its end, which this order reads first.
  const int64_t nch = (n + kChunk - 1) / kChunk;
  const int64_t ch = nch - 1 - ((int64_t)blockIdx.x * kWaves + wid);
  if (ch < 0) return;
  const int64_t c0 = ch * kChunk;
  const int64_t part = chunk_base_part(ch, block_off, chunk_cnt, lane);
  const int64_t ktot = kept_total[0];

  const bool full = c0 + kChunk <= n;
  uint32_t mb[J];
  Packed<XT, G> xv[J];
  Packed<int32_t, G> iv[IDX ? J : 1];
#pragma unroll
  for (int j = 0; j < J; ++j) {
    const int64_t gb = c0 + j * (kWave * G) + lane * G;
    Packed<int8_t, G> m;
    if (full) {
      // The mask was read by the count pass just before: plain loads, which
      // find part of it in the caches; values and indices only this pass
      // reads.
      load_items<int8_t, G, true>(mask, gb, n, m);
      load_nt(x, gb, xv[j]);
      if constexpr (IDX) load_nt(idx_in, gb, iv[j]);
    } else {
      load_items<int8_t, G, false>(mask, gb, n, m);
      load_items<XT, G, false>(x, gb, n, xv[j]);
      if constexpr (IDX) load_items<int32_t, G, false>(idx_in, gb, n, iv[j]);
    }
    // Four mask bytes per word (items past n load as 0: not kept).
    mb[j] = ones_bits4(m.w[0]) | ones_bits4(m.w[1]) << 4;
  }
  const int64_t kb = wave_sum(part);
  int gex[J];
  int wtot = 0;
#pragma unroll
  for (int j = 0; j < J; ++j) {
    const int c = __builtin_popcount(mb[j]);
    const int incl = wave_incl_scan(c);
    gex[j] = wtot + incl - c;
    wtot += readlane(incl, kWave - 1);
  }
  const int valid = full ? (int)kChunk : (int)(n - c0);
  const int64_t rb = ktot + (c0 - kb);
  ST* stage = s_stage[wid];
  // Kept element e of the chunk goes to stage[kept before e], any other to
  // stage[wtot + (e - kept before e)]; then both runs go out.
  auto place = [&](auto out, auto value_of) {
#pragma unroll
    for (int j = 0; j < J; ++j) {
      int kpos = gex[j];
      const int e0 = j * kWave * G + lane * G;
#pragma unroll
      for (int i = 0; i < G; ++i) {
        const bool kept = (mb[j] >> i) & 1u;
        const int e = e0 + i;
        const int pos = kept ? kpos : wtot + (e - kpos);
        if (full || e < valid) stage[pos] = (ST)value_of(j, i);
        kpos += kept ? 1 : 0;
      }
    }
    wave_lds_sync();
    flush_run(stage, out, kb, wtot, lane);
    flush_run(stage + wtot, out, rb, valid - wtot, lane);
    wave_lds_sync();
  };
  place(out_x, [&](int j, int i) { return xv[j][i]; });
  if constexpr (IDX) place(out_idx, [&](int j, int i) { return (uint32_t)iv[j][i]; });
}

template <int XB, bool IDX>
static int launch_split(const void* x, const int8_t* mask, const int32_t* idx_in, void* out_x, int32_t* out_idx,
                        const int64_t* kept, int64_t n, void* ws, hipStream_t st) {
  using XT = typename Bits<XB>::T;
  const CountedLayout l = counted_layout(n);
  const int64_t nch = (n + kChunk - 1) / kChunk;
  const int64_t blocks = (nch + kCountBlock / kWave - 1) / (kCountBlock / kWave);
  const char* w = static_cast<const char*>(ws);
  hipLaunchKernelGGL((split_counted_kernel<XB, IDX>), dim3((uint32_t)blocks), dim3(kCountBlock), 0, st,
                     static_cast<const XT*>(x), mask, idx_in, static_cast<XT*>(out_x), out_idx, kept, n,
                     reinterpret_cast<const int64_t*>(w + l.block_off),
                     reinterpret_cast<const uint64_t*>(w + l.chunk_cnt));
  TCS_CHECK_LAUNCH();
  return TCS_OK;
}

}  // namespace tcs

extern "C" {

size_t tcs_split_workspace_bytes(int64_t n) {
  // The count pass's block counts, block offsets and chunk counts after the
  // header (the counted layout); K sits after them.
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
