// Top-k on gfx950 by radix select, and the reference's sampled estimator
// of the k-th largest value.
//
// tcs_topk replaces run_topk_fp16 / run_topk_int16 (src/torch/torch_topk.h:
// 69-207; the reference narrows [x_min, x_max] with parallel splits).  Here
// the k-th largest order key T is found one 8-bit digit at a time, from the
// top: a histogram of digit r over the keys whose higher digits equal the
// digits chosen so far, then one workgroup picks the bin holding the k-th
// largest.  State passes between launches in the workspace, so no result
// reaches the host until the end.  One ordered compaction (look-back on
// packed (greater, equal) counts) then takes every key above T and the
// first `need` keys equal to T in input order, and the stable radix sort
// orders those k, largest first (ties by input position).  Reads: one pass
// per key byte plus one.
//
// tcs_topk_pivot follows run_topk_pivot_fp16 (src/torch/torch_topk.h:
// 23-60, src/kernels/kernel_topk_pivot.h): per block, over its first eight
// tiles of num_samples * 32 elements, the 8th largest of each 32, then the
// 4th largest of those per tile, then the largest over the tiles.  At least
// 32 elements of the tile are >= that estimate.
#include "tcs_common.hpp"
#include "tcs_device.hpp"
#include "tcs_sortkey.hpp"

namespace tcs {

constexpr int kTkThreads = 512, kTkMaxGroups = 1024, kTkPickThreads = 1024;
constexpr int kTkNW = 8, kTkG = 8;

template <int DT>
struct TopkGeom {
  using Raw = typename SortKey<DT>::Raw;
  static constexpr int J = sizeof(Raw) <= 2 ? 4 : 2;
  static constexpr int64_t kTile = (int64_t)kTkNW * J * kWave * kTkG;
};
constexpr int64_t kTkTileMin = (int64_t)kTkNW * 2 * kWave * kTkG;  // 8192

// ---------------------------------------------------------------------------
// Digit histogram among the keys matching the chosen higher digits.
// state[0]: key digits chosen so far; state[1]: how many of the keys
// matching them are still to be taken.  64-bit words, high half zero.
// ---------------------------------------------------------------------------
template <int DT>
__global__ __launch_bounds__(kTkThreads) void topk_hist_kernel(const typename SortKey<DT>::Raw* __restrict__ x,
                                                               int64_t n, int64_t chunk, uint32_t flip, int r,
                                                               const uint64_t* __restrict__ state,
                                                               uint64_t* __restrict__ partial) {
  using KT = SortKey<DT>;
  using Raw = typename KT::Raw;
  constexpr int NWV = kTkThreads / kWave, C = 4, NH = NWV * C;
  constexpr int V = 16 / (int)sizeof(Raw);
  // 257 words per copy: the copies of one bin fall in different LDS banks.
  __shared__ uint32_t h[NH][257];
  const int tid = threadIdx.x, wid = tid / kWave;
  const int hc = wid * C + (__lane_id() % C);
  for (int k = tid; k < NH * 257; k += kTkThreads) (&h[0][0])[k] = 0;
  __syncthreads();
  const int hs = 8 * (r + 1);  // bits above digit r
  const uint32_t want = hs >= KT::kBits ? 0u : (uint32_t)(state[0] >> hs);
  const int64_t c0 = (int64_t)blockIdx.x * chunk;
  const int64_t c1 = c0 + chunk < n ? c0 + chunk : n;
  // Four 16-byte loads in flight per thread per step.
  constexpr int U = 4;
  constexpr int64_t STEP = (int64_t)kTkThreads * V;
  for (int64_t b0 = c0 + (int64_t)tid * V; b0 < c1; b0 += U * STEP) {
    Raw v[U][V];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t b = b0 + u * STEP;
      if (b + V <= c1) {
        load_items<Raw, V, true>(x, b, c1, v[u]);
      } else {
        load_items<Raw, V, false>(x, b, c1, v[u]);
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t b = b0 + u * STEP;
#pragma unroll
      for (int i = 0; i < V; ++i) {
        const uint32_t k = KT::key(v[u][i]) ^ flip;
        const bool match = b + i < c1 && (hs >= KT::kBits || (k >> hs) == want);
        if (match) atomicAdd(&h[hc][(k >> (8 * r)) & 0xffu], 1u);
      }
    }
  }
  __syncthreads();
  for (int e = tid; e < 256; e += kTkThreads) {
    uint32_t s = 0;
#pragma unroll
    for (int w = 0; w < NH; ++w) s += h[w][e];
    partial[(size_t)blockIdx.x * 256 + e] = s;
  }
}

// Picks digit r of the k-th largest key: the bin b where the count of keys
// in bins >= b first reaches the number still to take.
__global__ __launch_bounds__(kTkPickThreads) void topk_pick_kernel(const uint64_t* __restrict__ partial,
                                                                   int groups, int r, int first, int64_t k,
                                                                   uint64_t* __restrict__ state) {
  const int tid = threadIdx.x, d = tid & 255, q = tid >> 8;
  uint32_t s = 0;
  int g = q;
  for (; g + 28 < groups; g += 32) {  // eight loads in flight
    uint64_t v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = partial[(size_t)(g + 4 * u) * 256 + d];
#pragma unroll
    for (int u = 0; u < 8; ++u) s += (uint32_t)v[u];
  }
  for (; g < groups; g += 4) s += (uint32_t)partial[(size_t)g * 256 + d];
  __shared__ uint32_t sr[4][256];
  __shared__ uint32_t ws[4];
  __shared__ uint64_t s_state[2];
  sr[q][d] = s;
  if (tid == 0) {
    s_state[0] = first ? 0 : state[0];
    s_state[1] = first ? (uint64_t)k : state[1];
  }
  __syncthreads();
  // suffix[b] = keys in bins >= b: an inclusive scan over b' = 255 - b.
  uint32_t cnt = 0, incl = 0;
  if (tid < 256) {
    cnt = sr[0][255 - tid] + sr[1][255 - tid] + sr[2][255 - tid] + sr[3][255 - tid];
    incl = wave_incl_scan(cnt);
    if (__lane_id() == kWave - 1) ws[tid / kWave] = incl;
  }
  __syncthreads();
  if (tid < 256) {
    uint32_t before = 0;
    for (int w = 0; w < tid / kWave; ++w) before += ws[w];
    const uint64_t suffix = before + incl;  // keys in bins >= 255 - tid
    const uint64_t need = s_state[1];
    if (suffix >= need && suffix - cnt < need) {
      const uint32_t b = 255u - (uint32_t)tid;
      state[0] = s_state[0] | ((uint64_t)b << (8 * r));
      state[1] = need - (suffix - cnt);
    }
  }
}

// ---------------------------------------------------------------------------
// Ordered selection: keys above T go to [0, k - need) in input order, the
// first `need` keys equal to T to [k - need, k).
// ---------------------------------------------------------------------------
template <int DT>
__global__ __launch_bounds__(kTkNW * kWave) void topk_select_kernel(
    const typename SortKey<DT>::Raw* __restrict__ x, int64_t n, uint32_t flip,
    const uint64_t* __restrict__ state, int64_t k, typename SortKey<DT>::Raw* __restrict__ cand_v,
    int32_t* __restrict__ cand_i, void* ws, uint32_t num_tiles) {
  using KT = SortKey<DT>;
  using Raw = typename KT::Raw;
  constexpr int G = kTkG, J = TopkGeom<DT>::J, NW = kTkNW;
  constexpr int64_t TILE = TopkGeom<DT>::kTile;
  constexpr int64_t GSPAN = (int64_t)kWave * G, WSEG = GSPAN * J;
  __shared__ int64_t s_wval[NW], s_wexcl[NW];
  __shared__ int64_t s_base;
  __shared__ uint32_t s_tag;
  if (threadIdx.x == 0) s_tag = launch_arrive(ws);
  const uint32_t T = (uint32_t)state[0];
  const int64_t need = (int64_t)state[1];
  const int64_t c_gt = k - need;
  const int lane = __lane_id(), wid = threadIdx.x / kWave;
  const uint32_t tile = blockIdx.x;
  const int64_t t0 = (int64_t)tile * TILE;
  const int64_t wbase = t0 + (int64_t)wid * WSEG + (int64_t)lane * G;

  // Packed counts: greater << 32 | equal (each below 2^31).
  auto packed_of = [&](Raw v, bool valid) -> int64_t {
    const uint32_t kk = KT::key(v) ^ flip;
    return !valid ? 0 : (kk > T ? ((int64_t)1 << 32) : (kk == T ? 1 : 0));
  };
  uint32_t gtb[J], eqb[J];
  Packed<Raw, G> xv[J];
  int64_t gex[J];
  int64_t wtot = 0;
#pragma unroll
  for (int j = 0; j < J; ++j) {
    const int64_t gb = wbase + j * GSPAN;
    if (t0 + TILE <= n) {
      load_items<Raw, G, true>(x, gb, n, xv[j]);
    } else {
      load_items<Raw, G, false>(x, gb, n, xv[j]);
    }
    uint32_t g = 0, e = 0;
#pragma unroll
    for (int i = 0; i < G; ++i) {
      const uint32_t kk = KT::key(xv[j][i]) ^ flip;
      const bool valid = gb + i < n;
      g |= (valid && kk > T ? 1u : 0u) << i;
      e |= (valid && kk == T ? 1u : 0u) << i;
    }
    gtb[j] = g;
    eqb[j] = e;
    const int64_t c = ((int64_t)__builtin_popcount(g) << 32) | __builtin_popcount(e);
    const int64_t incl = wave_incl_scan(c);
    gex[j] = wtot + incl - c;
    wtot += readlane(incl, kWave - 1);
  }
  if (lane == 0) s_wval[wid] = wtot;
  __syncthreads();
  if (wid == 0) {
    const int64_t a = lane < NW ? s_wval[lane] : 0;
    const int64_t incl = wave_incl_scan(a);
    const int64_t agg = readlane(incl, NW - 1);
    if (lane < NW) s_wexcl[lane] = incl - a;
    auto fb = [&](int64_t u) {
      const int64_t u0 = u * TILE;
      const int64_t u1 = u0 + TILE < n ? u0 + TILE : n;
      int64_t c = 0;
      for (int64_t p = u0 + lane; p < u1; p += kWave) c += packed_of(x[p], true);
      return FlagVal<int64_t>{u == 0 ? 1 : 0, wave_sum(c)};
    };
    uint64_t* desc = ws_desc(ws);
    const uint32_t tag = s_tag;
    int64_t carry = 0;
    if (tile == 0) {
      if (lane == 0 && num_tiles > 1) desc_publish<int64_t>(desc, 0, agg, tag, kFlagIncl);
    } else {
      if (lane == 0) desc_publish<int64_t>(desc, tile, agg, tag, kFlagAgg);
      carry = lookback<int64_t>(desc, tile, tag, fb);
      if (lane == 0) desc_publish<int64_t>(desc, tile, carry + agg, tag, kFlagIncl);
    }
    if (lane == 0) s_base = carry;
  }
  __syncthreads();
  const int64_t base = s_base + s_wexcl[wid];
#pragma unroll
  for (int j = 0; j < J; ++j) {
    const int64_t before = base + gex[j];
    int64_t gpos = before >> 32, epos = before & 0xffffffffll;
    const int64_t gb = wbase + j * GSPAN;
#pragma unroll
    for (int i = 0; i < G; ++i) {
      if ((gtb[j] >> i) & 1u) {
        cand_v[gpos] = xv[j][i];
        cand_i[gpos] = (int32_t)(gb + i);
        ++gpos;
      } else if ((eqb[j] >> i) & 1u) {
        if (epos < need) {
          cand_v[c_gt + epos] = xv[j][i];
          cand_i[c_gt + epos] = (int32_t)(gb + i);
        }
        ++epos;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Sampled estimate of the k-th largest fp16 value (one per block).
// ---------------------------------------------------------------------------
constexpr int kPivInner = 8, kPivOuter = 4, kPivIters = 8;

// Descending rank of `key` among the lanes [first, first + count) of its
// wave (ties broken by lane index); lanes outside compare as absent.
__device__ __forceinline__ int rank_desc(uint32_t key, int first, int count) {
  const int lane = __lane_id();
  int r = 0;
  for (int j = 0; j < count; ++j) {
    const uint32_t o = (uint32_t)__shfl((int)key, first + j);
    r += (o > key || (o == key && first + j < lane)) ? 1 : 0;
  }
  return r;
}

__global__ __launch_bounds__(256) void topk_pivot_kernel(const uint16_t* __restrict__ x, int64_t n, int num_samples,
                                                         int64_t tiles_per_block, uint16_t* __restrict__ out) {
  using KT = SortKey<TCS_F16>;
  const int tid = threadIdx.x, lane = __lane_id(), wid = tid / kWave;
  const int64_t tile_len = (int64_t)num_samples * 32;
  const int64_t t_begin = (int64_t)blockIdx.x * tiles_per_block;
  __shared__ uint32_t s_gk[64];
  __shared__ uint16_t s_gr[64];
  __shared__ uint32_t s_tk;
  __shared__ uint16_t s_tr;
  uint32_t best_k = 0;
  uint16_t best_r = 0xfc00u;  // -inf
  for (int it = 0; it < kPivIters && it < tiles_per_block; ++it) {
    const int64_t base = (t_begin + it) * tile_len;
    if (base >= n) break;
    // Two groups of 32 per wave, four waves: eight groups per round.
    for (int g0 = wid * 2; g0 < num_samples; g0 += 8) {
      const int g = g0 + (lane >> 5);
      const int64_t p = base + (int64_t)g * 32 + (lane & 31);
      const uint16_t raw = (g < num_samples && p < n) ? x[p] : (uint16_t)0xfc00u;
      const uint32_t key = KT::key(raw);
      const int r = rank_desc(key, lane & 32, 32);
      if (g < num_samples && r == kPivInner - 1) {
        s_gk[g] = key;
        s_gr[g] = raw;
      }
    }
    __syncthreads();
    if (wid == 0) {
      const bool in = lane < num_samples;
      const uint32_t key = in ? s_gk[lane] : 0u;
      const int r = rank_desc(key, 0, num_samples);
      if (in && r == kPivOuter - 1) {
        s_tk = key;
        s_tr = s_gr[lane];
      }
    }
    __syncthreads();
    if (s_tk > best_k || it == 0) {
      best_k = s_tk;
      best_r = s_tr;
    }
    __syncthreads();
  }
  if (tid == 0) out[blockIdx.x] = best_r;
}

inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

inline int topk_elem_bytes(int dt) {
  switch (dt) {
    case TCS_I8: return 1;
    case TCS_F16: case TCS_BF16: case TCS_I16: return 2;
    case TCS_F32: case TCS_I32: case TCS_U32: return 4;
    default: return 0;
  }
}

struct TopkLayout {
  size_t state_off, part_off, own_total;
  size_t cand_i_off, sort_scratch_off, scratch;
  int groups;
  int64_t chunk;
};

inline TopkLayout topk_layout(int64_t n, int64_t k, int dt) {
  TopkLayout L{};
  const uint64_t tiles = (uint64_t)((n + kTkTileMin - 1) / kTkTileMin);
  int64_t groups = (n + 65535) / 65536;
  if (groups > kTkMaxGroups) groups = kTkMaxGroups;
  if (groups < 1) groups = 1;
  int64_t chunk = (n + groups - 1) / groups;
  chunk = (chunk + 63) & ~(int64_t)63;
  L.groups = (int)groups;
  L.chunk = chunk;
  L.state_off = lookback_ws_bytes(tiles);
  L.part_off = align256(L.state_off + 256);
  L.own_total = align256(L.part_off + (size_t)groups * 256 * 8);
  L.cand_i_off = align256((size_t)k * topk_elem_bytes(dt));
  L.sort_scratch_off = align256(L.cand_i_off + (size_t)k * 4);
  L.scratch = L.sort_scratch_off + tcs_sort_scratch_bytes(k, dt);
  return L;
}

template <int DT>
static int topk_run(const void* x, int64_t n, int64_t k, bool largest, void* out_v, int32_t* out_i, void* ws,
                    size_t ws_bytes, void* scratch, size_t scratch_bytes, hipStream_t st) {
  using KT = SortKey<DT>;
  using Raw = typename KT::Raw;
  constexpr int P = KT::kBits / 8;
  const TopkLayout L = topk_layout(n, k, DT);
  char* w = static_cast<char*>(ws);
  auto* state = reinterpret_cast<uint64_t*>(w + L.state_off);
  auto* partial = reinterpret_cast<uint64_t*>(w + L.part_off);
  char* sc = static_cast<char*>(scratch);
  auto* cand_v = reinterpret_cast<Raw*>(sc);
  auto* cand_i = reinterpret_cast<int32_t*>(sc + L.cand_i_off);
  const uint32_t flip = largest ? 0u : (KT::kBits == 32 ? 0xffffffffu : ((1u << KT::kBits) - 1u));
  const Raw* xr = static_cast<const Raw*>(x);
  for (int r = P - 1; r >= 0; --r) {
    hipLaunchKernelGGL((topk_hist_kernel<DT>), dim3(L.groups), dim3(kTkThreads), 0, st, xr, n, L.chunk, flip, r,
                       state, partial);
    TCS_CHECK_LAUNCH();
    hipLaunchKernelGGL(topk_pick_kernel, dim3(1), dim3(kTkPickThreads), 0, st, partial, L.groups, r,
                       r == P - 1 ? 1 : 0, k, state);
    TCS_CHECK_LAUNCH();
  }
  constexpr int64_t kTile = TopkGeom<DT>::kTile;
  const uint32_t tiles = (uint32_t)((n + kTile - 1) / kTile);
  hipLaunchKernelGGL((topk_select_kernel<DT>), dim3(tiles), dim3(kTkNW * kWave), 0, st, xr, n, flip, state, k,
                     cand_v, cand_i, ws, tiles);
  TCS_CHECK_LAUNCH();
  // Order the k selected: largest first (smallest first for !largest), ties
  // in input order.  The sort reuses this workspace (same header and epoch).
  return tcs_sort(cand_v, DT, cand_i, out_v, out_i, k, largest ? 1 : 0, ws, ws_bytes, sc + L.sort_scratch_off,
                  scratch_bytes - L.sort_scratch_off, st);
}

}  // namespace tcs

extern "C" {

size_t tcs_topk_workspace_bytes(int64_t n, int64_t k, int dtype) {
  using namespace tcs;
  if (n < 0 || k < 0 || topk_elem_bytes(dtype) == 0) return 0;
  const size_t own = topk_layout(n, k, dtype).own_total;
  const size_t srt = tcs_sort_workspace_bytes(k, dtype);
  return own > srt ? own : srt;
}

size_t tcs_topk_scratch_bytes(int64_t n, int64_t k, int dtype) {
  using namespace tcs;
  if (n < 0 || k < 0 || topk_elem_bytes(dtype) == 0) return 0;
  return topk_layout(n, k, dtype).scratch;
}

int tcs_topk(const void* x, int dtype, int64_t n, int64_t k, int largest, void* out_vals, int32_t* out_idx,
             void* workspace, size_t workspace_bytes, void* scratch, size_t scratch_bytes, void* stream) {
  using namespace tcs;
  if (topk_elem_bytes(dtype) == 0) return TCS_ERR_UNSUPPORTED_DTYPE;
  if (n < 0 || k < 0 || k > n || n >= ((int64_t)1 << 31)) return TCS_ERR_INVALID_ARG;
  if (k == 0) return TCS_OK;
  if (!x || !out_vals || !out_idx || !aligned16(x)) return TCS_ERR_INVALID_ARG;
  if (!workspace || workspace_bytes < tcs_topk_workspace_bytes(n, k, dtype)) return TCS_ERR_WORKSPACE;
  if (!scratch || scratch_bytes < tcs_topk_scratch_bytes(n, k, dtype)) return TCS_ERR_WORKSPACE;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const bool lg = largest != 0;
  switch (dtype) {
    case TCS_F16: return topk_run<TCS_F16>(x, n, k, lg, out_vals, out_idx, workspace, workspace_bytes, scratch, scratch_bytes, st);
    case TCS_BF16: return topk_run<TCS_BF16>(x, n, k, lg, out_vals, out_idx, workspace, workspace_bytes, scratch, scratch_bytes, st);
    case TCS_F32: return topk_run<TCS_F32>(x, n, k, lg, out_vals, out_idx, workspace, workspace_bytes, scratch, scratch_bytes, st);
    case TCS_I8: return topk_run<TCS_I8>(x, n, k, lg, out_vals, out_idx, workspace, workspace_bytes, scratch, scratch_bytes, st);
    case TCS_I16: return topk_run<TCS_I16>(x, n, k, lg, out_vals, out_idx, workspace, workspace_bytes, scratch, scratch_bytes, st);
    case TCS_I32: return topk_run<TCS_I32>(x, n, k, lg, out_vals, out_idx, workspace, workspace_bytes, scratch, scratch_bytes, st);
    default: return topk_run<TCS_U32>(x, n, k, lg, out_vals, out_idx, workspace, workspace_bytes, scratch, scratch_bytes, st);
  }
}

int64_t tcs_topk_pivot_blocks(int64_t n, int num_samples) {
  if (n <= 0 || num_samples <= 0) return 0;
  const int64_t tile = (int64_t)num_samples * 32;
  const int64_t tiles = (n + tile - 1) / tile;
  return tiles < 64 ? tiles : 64;
}

int tcs_topk_pivot(const void* x, int64_t n, int num_samples, void* out, int64_t out_len, void* stream) {
  using namespace tcs;
  if (n <= 0 || num_samples < kPivOuter || num_samples > 64 || !x || !out) return TCS_ERR_INVALID_ARG;
  const int64_t blocks = tcs_topk_pivot_blocks(n, num_samples);
  if (out_len < blocks) return TCS_ERR_INVALID_ARG;
  const int64_t tile = (int64_t)num_samples * 32;
  const int64_t tiles = (n + tile - 1) / tile;
  const int64_t per_block = (tiles + blocks - 1) / blocks;
  hipLaunchKernelGGL(topk_pivot_kernel, dim3((uint32_t)blocks), dim3(256), 0, static_cast<hipStream_t>(stream),
                     static_cast<const uint16_t*>(x), n, num_samples, per_block, static_cast<uint16_t*>(out));
  TCS_CHECK_LAUNCH();
  return TCS_OK;
}

}  // extern "C"
