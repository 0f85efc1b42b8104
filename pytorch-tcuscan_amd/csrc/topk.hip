// Top-k on gfx950 by radix select, and the reference's sampled estimator
// of the k-th largest value.
//
// tcs_topk replaces run_topk_fp16 / run_topk_int16 (src/torch/torch_topk.h:
// 69-207; the reference narrows [x_min, x_max] with parallel splits).  Here
// the k-th largest order key T is found one 8-bit digit at a time, from the
// top: a histogram of digit r over the keys whose higher digits equal the
// digits chosen so far, then one workgroup picks the bin holding the k-th
// largest.  State passes between launches in the workspace, so no result
// reaches the host until the end.  An ordered selection then takes every
// key above T and the first `need` keys equal to T in input order (per-tile
// counts, their scan, and placement by the tiles that hold any of them),
// and the stable radix sort orders those k, largest first (ties by input
// position).  Reads: one pass per key byte plus one, plus the tiles that
// hold selected keys.
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
// Placement walks the tiles with a fixed grid: most tiles place nothing
// when k is small, and a workgroup per tile would cost a launch each.
constexpr uint32_t kTkPlaceGrid = 2048;

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
  // Four 16-byte loads in flight per thread per step; steps that lie wholly
  // inside the chunk skip the per-element bounds checks.
  constexpr int U = 4;
  constexpr int64_t STEP = (int64_t)kTkThreads * V;
  auto count = [&](const Raw (&v)[V]) {
#pragma unroll
    for (int i = 0; i < V; ++i) {
      const uint32_t k = KT::key(v[i]) ^ flip;
      if (hs >= KT::kBits || (k >> hs) == want) atomicAdd(&h[hc][(k >> (8 * r)) & 0xffu], 1u);
    }
  };
  int64_t b0 = c0 + (int64_t)tid * V;
  const int64_t full_end = c0 + ((c1 - c0) / (U * STEP)) * (U * STEP);  // steps [c0, full_end) are whole
  for (; b0 < full_end; b0 += U * STEP) {
    Raw v[U][V];
#pragma unroll
    for (int u = 0; u < U; ++u) load_items<Raw, V, true>(x, b0 + u * STEP, c1, v[u]);
#pragma unroll
    for (int u = 0; u < U; ++u) count(v[u]);
  }
  for (; b0 < c1; b0 += STEP) {
    Raw v[V];
    load_items<Raw, V, false>(x, b0, c1, v);
#pragma unroll
    for (int i = 0; i < V; ++i) {
      const uint32_t k = KT::key(v[i]) ^ flip;
      const bool match = b0 + i < c1 && (hs >= KT::kBits || (k >> hs) == want);
      if (match) atomicAdd(&h[hc][(k >> (8 * r)) & 0xffu], 1u);
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

// First level of the digit-count sum: workgroup q adds the partials of
// groups [q * per, (q + 1) * per) into psum[q] (64-bit, high half zero), so
// that the pick reads kTkSumGroups rows instead of one per histogram group.
constexpr int kTkSumGroups = 32;
__global__ __launch_bounds__(256) void topk_sum_kernel(const uint64_t* __restrict__ partial, int groups, int per,
                                                       uint64_t* __restrict__ psum) {
  const int d = threadIdx.x;
  const int g0 = blockIdx.x * per;
  const int g1 = g0 + per < groups ? g0 + per : groups;
  uint32_t s = 0;
  int g = g0;
  for (; g + 8 <= g1; g += 8) {  // eight loads in flight
    uint64_t v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = partial[(size_t)(g + u) * 256 + d];
#pragma unroll
    for (int u = 0; u < 8; ++u) s += (uint32_t)v[u];
  }
  for (; g < g1; ++g) s += (uint32_t)partial[(size_t)g * 256 + d];
  psum[(size_t)blockIdx.x * 256 + d] = s;
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
// first `need` keys equal to T to [k - need, k).  Three launches with no
// look-back: per-tile counts, one scan of them, then placement, in which a
// tile that places nothing (most of them when k is small) exits after
// reading two words.
// ---------------------------------------------------------------------------

// Loads tile `tile` and sets, per lane group, the bits of the keys above T
// (gtb) and equal to T (eqb); FULL: the tile lies wholly inside [0, n).
// A tile holds at most 16384 keys, so a count
// packed as greater << 16 | equal never carries between the halves.
template <int DT, bool FULL>
__device__ __forceinline__ void topk_classify(const typename SortKey<DT>::Raw* __restrict__ x, int64_t n,
                                              uint32_t flip, uint32_t T, uint32_t tile, int wid, int lane,
                                              Packed<typename SortKey<DT>::Raw, kTkG> (&xv)[TopkGeom<DT>::J],
                                              uint32_t (&gtb)[TopkGeom<DT>::J], uint32_t (&eqb)[TopkGeom<DT>::J]) {
  using KT = SortKey<DT>;
  using Raw = typename KT::Raw;
  constexpr int G = kTkG, J = TopkGeom<DT>::J;
  constexpr int64_t TILE = TopkGeom<DT>::kTile;
  static_assert(TILE <= 16384, "packed 16-bit tile counts");
  constexpr int64_t GSPAN = (int64_t)kWave * G, WSEG = GSPAN * J;
  const int64_t wbase = (int64_t)tile * TILE + (int64_t)wid * WSEG + (int64_t)lane * G;
#pragma unroll
  for (int j = 0; j < J; ++j) {
    const int64_t gb = wbase + j * GSPAN;
    load_items<Raw, G, FULL>(x, gb, n, xv[j]);
    uint32_t g = 0, e = 0;
#pragma unroll
    for (int i = 0; i < G; ++i) {
      const uint32_t kk = KT::key(xv[j][i]) ^ flip;
      const bool valid = FULL || gb + i < n;
      g |= (valid && kk > T ? 1u : 0u) << i;
      e |= (valid && kk == T ? 1u : 0u) << i;
    }
    gtb[j] = g;
    eqb[j] = e;
  }
}

template <int DT>
__device__ __forceinline__ void topk_classify_any(const typename SortKey<DT>::Raw* __restrict__ x, int64_t n,
                                                  uint32_t flip, uint32_t T, uint32_t tile, int wid, int lane,
                                                  Packed<typename SortKey<DT>::Raw, kTkG> (&xv)[TopkGeom<DT>::J],
                                                  uint32_t (&gtb)[TopkGeom<DT>::J],
                                                  uint32_t (&eqb)[TopkGeom<DT>::J]) {
  if ((int64_t)(tile + 1) * TopkGeom<DT>::kTile <= n) {
    topk_classify<DT, true>(x, n, flip, T, tile, wid, lane, xv, gtb, eqb);
  } else {
    topk_classify<DT, false>(x, n, flip, T, tile, wid, lane, xv, gtb, eqb);
  }
}

// tcnt[tile] = (keys above T) << 32 | (keys equal to T) in the tile.
template <int DT>
__global__ __launch_bounds__(kTkNW * kWave) void topk_tilecount_kernel(
    const typename SortKey<DT>::Raw* __restrict__ x, int64_t n, uint32_t flip,
    const uint64_t* __restrict__ state, uint64_t* __restrict__ tcnt) {
  constexpr int J = TopkGeom<DT>::J, NW = kTkNW;
  __shared__ int s_w[NW];
  const int lane = __lane_id(), wid = threadIdx.x / kWave;
  Packed<typename SortKey<DT>::Raw, kTkG> xv[J];
  uint32_t gtb[J], eqb[J];
  topk_classify_any<DT>(x, n, flip, (uint32_t)state[0], blockIdx.x, wid, lane, xv, gtb, eqb);
  int c = 0;
#pragma unroll
  for (int j = 0; j < J; ++j) c += (__builtin_popcount(gtb[j]) << 16) | __builtin_popcount(eqb[j]);
  c = wave_sum(c);
  if (lane == 0) s_w[wid] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) t += s_w[w];
    tcnt[blockIdx.x] = ((uint64_t)((uint32_t)t >> 16) << 32) | ((uint32_t)t & 0xffffu);
  }
}

// In place: tcnt[t] = sum of tcnt over the tiles before t, for t in
// [0, tiles]; one workgroup, 16 consecutive tiles per thread per round.
// The packed halves sum without a carry (each total is below 2^31).
constexpr int kTkScanThreads = 1024, kTkScanPer = 16;
__global__ __launch_bounds__(kTkScanThreads) void topk_tilescan_kernel(uint64_t* __restrict__ tcnt,
                                                                       uint32_t tiles) {
  constexpr int NWV = kTkScanThreads / kWave;
  constexpr int64_t ROUND = (int64_t)kTkScanThreads * kTkScanPer;
  __shared__ int64_t s_w[NWV];
  const int tid = threadIdx.x, lane = __lane_id(), wid = tid / kWave;
  int64_t carry = 0;
  for (int64_t r0 = 0; r0 < tiles; r0 += ROUND) {
    const int64_t t0 = r0 + (int64_t)tid * kTkScanPer;
    int64_t v[kTkScanPer];
    int64_t sum = 0;
#pragma unroll
    for (int q = 0; q < kTkScanPer; ++q) {
      v[q] = t0 + q < tiles ? (int64_t)tcnt[t0 + q] : 0;
      sum += v[q];
    }
    const int64_t incl = wave_incl_scan(sum);
    if (lane == kWave - 1) s_w[wid] = incl;
    __syncthreads();
    int64_t before = carry, total = carry;
#pragma unroll
    for (int w = 0; w < NWV; ++w) {
      const int64_t sw = s_w[w];
      before += w < wid ? sw : 0;
      total += sw;
    }
    before += incl - sum;
#pragma unroll
    for (int q = 0; q < kTkScanPer; ++q) {
      if (t0 + q < tiles) tcnt[t0 + q] = (uint64_t)before;
      before += v[q];
    }
    carry = total;
    __syncthreads();  // s_w is rewritten by the next round
  }
  if (tid == 0) tcnt[tiles] = (uint64_t)carry;
}

template <int DT>
__global__ __launch_bounds__(kTkNW * kWave) void topk_place_kernel(
    const typename SortKey<DT>::Raw* __restrict__ x, int64_t n, uint32_t flip,
    const uint64_t* __restrict__ state, int64_t k, const uint64_t* __restrict__ tcnt,
    typename SortKey<DT>::Raw* __restrict__ cand_v, int32_t* __restrict__ cand_i, uint32_t num_tiles) {
  constexpr int G = kTkG, J = TopkGeom<DT>::J, NW = kTkNW;
  constexpr int64_t TILE = TopkGeom<DT>::kTile;
  constexpr int64_t GSPAN = (int64_t)kWave * G, WSEG = GSPAN * J;
  __shared__ int s_wval[NW];
  const int64_t need = (int64_t)state[1];
  const int64_t c_gt = k - need;
  const int lane = __lane_id(), wid = threadIdx.x / kWave;
  for (uint32_t tile = blockIdx.x; tile < num_tiles; tile += gridDim.x) {
    const uint64_t e0 = tcnt[tile], e1 = tcnt[tile + 1];
    const int64_t gt0 = (int64_t)(e0 >> 32), eq0 = (int64_t)(e0 & 0xffffffffu);
    const int64_t gth = (int64_t)(e1 >> 32) - gt0, eqh = (int64_t)(e1 & 0xffffffffu) - eq0;
    if (gth == 0 && (eqh == 0 || eq0 >= need)) continue;  // the same for the whole workgroup

    Packed<typename SortKey<DT>::Raw, G> xv[J];
    uint32_t gtb[J], eqb[J];
    topk_classify_any<DT>(x, n, flip, (uint32_t)state[0], tile, wid, lane, xv, gtb, eqb);
    int gex[J];
    int wtot = 0;
#pragma unroll
    for (int j = 0; j < J; ++j) {
      const int c = (__builtin_popcount(gtb[j]) << 16) | __builtin_popcount(eqb[j]);
      const int incl = wave_incl_scan(c);
      gex[j] = wtot + incl - c;
      wtot += readlane(incl, kWave - 1);
    }
    if (lane == 0) s_wval[wid] = wtot;
    __syncthreads();
    int wex = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) wex += w < wid ? s_wval[w] : 0;
    __syncthreads();  // s_wval is rewritten for the next tile
    const int64_t wbase = (int64_t)tile * TILE + (int64_t)wid * WSEG + (int64_t)lane * G;
#pragma unroll
    for (int j = 0; j < J; ++j) {
      const int before = wex + gex[j];
      int64_t gpos = gt0 + (before >> 16), epos = eq0 + (before & 0xffff);
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
  size_t state_off, part_off, psum_off, own_total;
  size_t cand_i_off, tcnt_off, sort_scratch_off, scratch;
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
  L.psum_off = align256(L.part_off + (size_t)groups * 256 * 8);
  L.own_total = align256(L.psum_off + (size_t)kTkSumGroups * 256 * 8);
  L.cand_i_off = align256((size_t)k * topk_elem_bytes(dt));
  L.tcnt_off = align256(L.cand_i_off + (size_t)k * 4);
  L.sort_scratch_off = align256(L.tcnt_off + (size_t)(tiles + 1) * 8);
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
  auto* psum = reinterpret_cast<uint64_t*>(w + L.psum_off);
  const int sum_per = (L.groups + kTkSumGroups - 1) / kTkSumGroups;
  const int sum_groups = (L.groups + sum_per - 1) / sum_per;
  char* sc = static_cast<char*>(scratch);
  auto* cand_v = reinterpret_cast<Raw*>(sc);
  auto* cand_i = reinterpret_cast<int32_t*>(sc + L.cand_i_off);
  const uint32_t flip = largest ? 0u : (KT::kBits == 32 ? 0xffffffffu : ((1u << KT::kBits) - 1u));
  const Raw* xr = static_cast<const Raw*>(x);
  for (int r = P - 1; r >= 0; --r) {
    hipLaunchKernelGGL((topk_hist_kernel<DT>), dim3(L.groups), dim3(kTkThreads), 0, st, xr, n, L.chunk, flip, r,
                       state, partial);
    TCS_CHECK_LAUNCH();
    hipLaunchKernelGGL(topk_sum_kernel, dim3(sum_groups), dim3(256), 0, st, partial, L.groups, sum_per, psum);
    TCS_CHECK_LAUNCH();
    hipLaunchKernelGGL(topk_pick_kernel, dim3(1), dim3(kTkPickThreads), 0, st, psum, sum_groups, r,
                       r == P - 1 ? 1 : 0, k, state);
    TCS_CHECK_LAUNCH();
  }
  constexpr int64_t kTile = TopkGeom<DT>::kTile;
  const uint32_t tiles = (uint32_t)((n + kTile - 1) / kTile);
  auto* tcnt = reinterpret_cast<uint64_t*>(sc + L.tcnt_off);
  hipLaunchKernelGGL((topk_tilecount_kernel<DT>), dim3(tiles), dim3(kTkNW * kWave), 0, st, xr, n, flip, state, tcnt);
  TCS_CHECK_LAUNCH();
  hipLaunchKernelGGL(topk_tilescan_kernel, dim3(1), dim3(kTkScanThreads), 0, st, tcnt, tiles);
  TCS_CHECK_LAUNCH();
  const uint32_t place_grid = tiles < kTkPlaceGrid ? tiles : kTkPlaceGrid;
  hipLaunchKernelGGL((topk_place_kernel<DT>), dim3(place_grid), dim3(kTkNW * kWave), 0, st, xr, n, flip, state, k,
                     tcnt, cand_v, cand_i, tiles);
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
