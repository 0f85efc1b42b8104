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
// position).  The per-tile counts of keys of two or more bytes come from a
// record the last digit round keeps (keys above the higher digits, low
// bytes of the few matching them), so reads are one pass per key byte plus
// the tiles that hold selected keys (one-byte keys: one more pass for the
// tile counts).
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

// The last digit round of a key of two or more bytes also records, per tile
// (one unrolled step of a histogram workgroup covers one tile), the number
// of keys whose higher digits are above the chosen ones and the low bytes of
// up to kTkListMax keys whose higher digits equal them.  Once the last digit
// is chosen, each tile's counts above and at the threshold follow from that
// record without reading the keys again; a tile with more matching keys than
// were recorded is counted from its keys (topk_tilefix_kernel).
constexpr int kTkListMax = 16;
// Placement reads only the sub-chunks of a tile that hold selected keys
// (their record's mask bits): a selected key is above or matches the
// higher digits.
constexpr int64_t kTkSubChunk = 2048;
// Tiles per histogram chunk: chunks are whole tiles, at most
// ceil(n / kTkMaxGroups) <= 2^21 keys for n < 2^31, so 256 of 8192.
constexpr int kTkMaxChunkTiles = 256;

inline int64_t topk_tile(int key_bytes) { return (int64_t)kTkNW * (key_bytes <= 2 ? 4 : 2) * kWave * kTkG; }

// ---------------------------------------------------------------------------
// Digit histogram among the keys matching the chosen higher digits.
// state[0]: key digits chosen so far; state[1]: how many of the keys
// matching them are still to be taken.  64-bit words, high half zero.
// ---------------------------------------------------------------------------
// FUSE (the last round, r = 0, of keys of two or more bytes): also record
// per tile the keys above the chosen higher digits, which of its
// kTkSubChunk-key sub-chunks hold any key above or matching them, and the
// low bytes of up to kTkListMax keys matching them: tstat[tile] = above <<
// 32 | sub-chunk mask << 16 | matching, tlist[tile * kTkListMax + j] = low
// byte (any order).
template <int DT, bool FUSE>
__global__ __launch_bounds__(kTkThreads) void topk_hist_kernel(const typename SortKey<DT>::Raw* __restrict__ x,
                                                               int64_t n, int64_t chunk, uint32_t flip, int r,
                                                               const uint64_t* __restrict__ state,
                                                               uint64_t* __restrict__ partial,
                                                               uint64_t* __restrict__ tstat,
                                                               uint8_t* __restrict__ tlist) {
  using KT = SortKey<DT>;
  using Raw = typename KT::Raw;
  constexpr int NWV = kTkThreads / kWave, C = 4, NH = NWV * C;
  constexpr int V = 16 / (int)sizeof(Raw);
  // Four 16-byte loads in flight per thread per step; steps that lie wholly
  // inside the chunk skip the per-element bounds checks.
  constexpr int U = 4;
  constexpr int64_t STEP = (int64_t)kTkThreads * V;
  constexpr int64_t TILE = TopkGeom<DT>::kTile;
  static_assert(!FUSE || (U * STEP == TILE && KT::kBits > 8), "one whole step per tile");
  static_assert(TILE < 65536, "packed 16-bit tile counts");
  static_assert(TILE / kTkSubChunk <= 16, "sub-chunk mask in 16 bits");
  constexpr int NT = FUSE ? kTkMaxChunkTiles : 1;
  // 257 words per copy: the copies of one bin fall in different LDS banks.
  __shared__ uint32_t h[NH][257];
  __shared__ uint32_t s_am[NT];                         // per tile: above << 16 | matching
  __shared__ uint32_t s_sub[NT];                        // per tile: sub-chunks holding either
  __shared__ uint8_t s_list[NT][FUSE ? kTkListMax : 1];  // low bytes of matching keys
  const int tid = threadIdx.x, wid = tid / kWave;
  const int hc = wid * C + (__lane_id() % C);
  for (int k = tid; k < NH * 257; k += kTkThreads) (&h[0][0])[k] = 0;
  if constexpr (FUSE)
    for (int k = tid; k < NT; k += kTkThreads) s_am[k] = s_sub[k] = 0;
  __syncthreads();
  const int hs = 8 * (r + 1);  // bits above digit r
  const uint32_t want = hs >= KT::kBits ? 0u : (uint32_t)(state[0] >> hs);
  const int64_t c0 = (int64_t)blockIdx.x * chunk;
  const int64_t c1 = c0 + chunk < n ? c0 + chunk : n;
  auto count = [&](const Raw (&v)[V]) {
#pragma unroll
    for (int i = 0; i < V; ++i) {
      const uint32_t k = KT::key(v[i]) ^ flip;
      if (hs >= KT::kBits || (k >> hs) == want) atomicAdd(&h[hc][(k >> (8 * r)) & 0xffu], 1u);
    }
  };
  // FUSE: histogram one step's keys (Q groups of V, valid: in range) and add
  // them to the record of local tile lt; off: the tile offset of group 0's
  // first key (group u starts u * STEP later, inside one sub-chunk).  The
  // low bytes are taken group by group, so only one group's keys are live
  // at a time.
  auto count_record = [&](auto& v, int lt, int64_t off, auto valid) {
    constexpr int Q = (int)(sizeof(v) / sizeof(v[0]));
    uint32_t above = 0, subm = 0;
#pragma unroll
    for (int u = 0; u < Q; ++u) {
      uint32_t eqm = 0, abv = 0;
#pragma unroll
      for (int i = 0; i < V; ++i) {
        const uint32_t k = KT::key(v[u][i]) ^ flip;
        const bool ok = valid(u, i);
        const bool eq = ok && (k >> 8) == want;
        if (eq) atomicAdd(&h[hc][k & 0xffu], 1u);
        abv += ok && (k >> 8) > want ? 1u : 0u;
        eqm |= (eq ? 1u : 0u) << i;
      }
      above += abv;
      if ((eqm | abv) != 0) subm |= 1u << (int)((off + u * STEP) / kTkSubChunk);
      if (eqm != 0) {
        uint32_t slot = atomicAdd(&s_am[lt], (uint32_t)__builtin_popcount(eqm)) & 0xffffu;
        if (slot < (uint32_t)kTkListMax) {
#pragma unroll
          for (int i = 0; i < V; ++i) {
            if ((eqm >> i) & 1u) {
              if (slot < (uint32_t)kTkListMax) s_list[lt][slot] = (uint8_t)((KT::key(v[u][i]) ^ flip) & 0xffu);
              ++slot;
            }
          }
        }
      }
    }
    if (above != 0) atomicAdd(&s_am[lt], above << 16);
    if (subm != 0) atomicOr(&s_sub[lt], subm);
  };
  int64_t b0 = c0 + (int64_t)tid * V;
  const int64_t full_end = c0 + ((c1 - c0) / (U * STEP)) * (U * STEP);  // steps [c0, full_end) are whole
  for (; b0 < full_end; b0 += U * STEP) {
    Raw v[U][V];
#pragma unroll
    for (int u = 0; u < U; ++u) load_items<Raw, V, true>(x, b0 + u * STEP, c1, v[u]);
    if constexpr (FUSE) {
      // Chunks are whole tiles and one step is one tile.
      count_record(v, (int)((b0 - c0) / TILE), (b0 - c0) % TILE, [](int, int) { return true; });
    } else {
#pragma unroll
      for (int u = 0; u < U; ++u) count(v[u]);
    }
  }
  for (; b0 < c1; b0 += STEP) {
    Raw v[1][V];
    load_items<Raw, V, false>(x, b0, c1, v[0]);
    if constexpr (FUSE) {
      // [full_end, c1) is less than one tile: the chunk's last.
      count_record(v, (int)((full_end - c0) / TILE), b0 - full_end, [&](int, int i) { return b0 + i < c1; });
    } else {
#pragma unroll
      for (int i = 0; i < V; ++i) {
        const uint32_t k = KT::key(v[0][i]) ^ flip;
        const bool match = b0 + i < c1 && (hs >= KT::kBits || (k >> hs) == want);
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
  if constexpr (FUSE) {
    const int64_t t0 = c0 / TILE;
    const int ntl = c1 > c0 ? (int)((c1 - c0 + TILE - 1) / TILE) : 0;
    for (int e = tid; e < ntl; e += kTkThreads) {
      const uint32_t am = s_am[e];
      tstat[t0 + e] = ((uint64_t)(am >> 16) << 32) | (s_sub[e] << 16) | (am & 0xffffu);
    }
    for (int e = tid; e < ntl * kTkListMax; e += kTkThreads)
      tlist[t0 * kTkListMax + e] = s_list[e / kTkListMax][e % kTkListMax];
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

// tcnt[tile] as topk_tilecount_kernel writes it, from the record of the
// last digit round: keys above T are those whose higher digits are above
// T's plus the recorded matching keys whose low byte is above T's, keys
// equal to T the recorded ones with T's low byte.  A tile with more matching
// keys than were recorded is counted from its keys.  One wave per tile.
template <int DT>
__global__ __launch_bounds__(256) void topk_tilefix_kernel(const typename SortKey<DT>::Raw* __restrict__ x,
                                                           int64_t n, uint32_t flip,
                                                           const uint64_t* __restrict__ state,
                                                           const uint64_t* __restrict__ tstat,
                                                           const uint8_t* __restrict__ tlist,
                                                           uint64_t* __restrict__ tcnt, uint32_t tiles) {
  using KT = SortKey<DT>;
  using Raw = typename KT::Raw;
  constexpr int64_t TILE = TopkGeom<DT>::kTile;
  constexpr int V = 16 / (int)sizeof(Raw);
  const int lane = __lane_id();
  const uint32_t tile = blockIdx.x * (256 / kWave) + threadIdx.x / kWave;
  if (tile >= tiles) return;
  const uint32_t T = (uint32_t)state[0];
  const uint64_t rec = tstat[tile];
  const uint32_t m = (uint32_t)rec & 0xffffu;
  uint32_t g, e;
  if (m <= (uint32_t)kTkListMax) {
    const bool in = lane < (int)m;
    const uint32_t low = in ? (uint32_t)tlist[(size_t)tile * kTkListMax + lane] : 0u;
    g = (uint32_t)(rec >> 32) + (uint32_t)__popcll(__ballot(in && low > (T & 0xffu)));
    e = (uint32_t)__popcll(__ballot(in && low == (T & 0xffu)));
  } else {
    const int64_t t0 = (int64_t)tile * TILE;
    const int64_t t1 = t0 + TILE < n ? t0 + TILE : n;
    uint32_t gl = 0, el = 0;
    for (int64_t b = t0 + (int64_t)lane * V; b < t1; b += (int64_t)kWave * V) {
      Raw v[V];
      load_items<Raw, V>(x, b, t1, v);
#pragma unroll
      for (int i = 0; i < V; ++i) {
        const uint32_t kk = KT::key(v[i]) ^ flip;
        const bool ok = b + i < t1;
        gl += ok && kk > T ? 1u : 0u;
        el += ok && kk == T ? 1u : 0u;
      }
    }
    g = (uint32_t)wave_sum((int)gl);
    e = (uint32_t)wave_sum((int)el);
  }
  if (lane == 0) tcnt[tile] = ((uint64_t)g << 32) | e;
}

// In place: tcnt[t] = sum of tcnt over the tiles before t, for t in
// [0, tiles]; one workgroup, 16384 tiles per round.  Wave w takes the
// round's tiles [w * 1024, (w + 1) * 1024), lane l tiles w * 1024 + 64q + l
// (coalesced loads and stores: 16 consecutive tiles per thread put every
// lane of a load in a different line, 17.8 us per scan of 16384 tiles),
// scanned as 16 wave scans in order.  The packed halves sum without a carry
// (each total is below 2^31).
constexpr int kTkScanThreads = 1024, kTkScanPer = 16;
__global__ __launch_bounds__(kTkScanThreads) void topk_tilescan_kernel(uint64_t* __restrict__ tcnt,
                                                                       uint32_t tiles) {
  constexpr int NWV = kTkScanThreads / kWave;
  constexpr int64_t WSPAN = (int64_t)kWave * kTkScanPer;
  constexpr int64_t ROUND = (int64_t)kTkScanThreads * kTkScanPer;
  __shared__ int64_t s_w[NWV];
  const int tid = threadIdx.x, lane = __lane_id(), wid = tid / kWave;
  int64_t carry = 0;
  for (int64_t r0 = 0; r0 < tiles; r0 += ROUND) {
    const int64_t wb = r0 + (int64_t)wid * WSPAN + lane;
    int64_t v[kTkScanPer];
#pragma unroll
    for (int q = 0; q < kTkScanPer; ++q) v[q] = wb + q * kWave < tiles ? (int64_t)tcnt[wb + q * kWave] : 0;
    int64_t run = 0;  // the wave's tiles before row q
#pragma unroll
    for (int q = 0; q < kTkScanPer; ++q) {
      const int64_t incl = wave_incl_scan(v[q]);
      const int64_t row = readlane(incl, kWave - 1);
      v[q] = run + incl - v[q];
      run += row;
    }
    if (lane == 0) s_w[wid] = run;
    __syncthreads();
    int64_t before = carry, total = carry;
#pragma unroll
    for (int w = 0; w < NWV; ++w) {
      const int64_t sw = s_w[w];
      before += w < wid ? sw : 0;
      total += sw;
    }
#pragma unroll
    for (int q = 0; q < kTkScanPer; ++q)
      if (wb + q * kWave < tiles) tcnt[wb + q * kWave] = (uint64_t)(before + v[q]);
    carry = total;
    __syncthreads();  // s_w is rewritten by the next round
  }
  if (tid == 0) tcnt[tiles] = (uint64_t)carry;
}

// One wave per tile, no workgroup barrier: wave w of the grid takes tiles
// w, w + W, ... (W waves in all), reads the counts of 64 of them at once
// and ballots which place anything; for each such tile it reads the keys
// one 2048-key sub-chunk (four 512-key groups) at a time, only the
// sub-chunks its record marks when there is one (tstat non-null), and
// places them, carrying the tile's running offsets through one wave scan
// per group.  Lane l holds keys 8l..8l+7 of a group, so the wave's (group,
// lane, item) order is input order.  2^28 fp16, k = 2048: about 1,940
// tiles place; reading all of each took 31-39 us of the placement's 45-58
// (scattered 32 KB reads run far below the streaming rate), reading only
// the first 1 KB of each 4-5 us.
constexpr int kTkPlaceThreads = 256;
template <int DT>
__global__ __launch_bounds__(kTkPlaceThreads) void topk_place_kernel(
    const typename SortKey<DT>::Raw* __restrict__ x, int64_t n, uint32_t flip,
    const uint64_t* __restrict__ state, int64_t k, const uint64_t* __restrict__ tcnt,
    const uint64_t* __restrict__ tstat, typename SortKey<DT>::Raw* __restrict__ cand_v,
    int32_t* __restrict__ cand_i, uint32_t num_tiles) {
  using KT = SortKey<DT>;
  using Raw = typename KT::Raw;
  constexpr int G = kTkG;
  constexpr int64_t TILE = TopkGeom<DT>::kTile;
  constexpr int64_t GSPAN = (int64_t)kWave * G;
  constexpr int B = (int)(kTkSubChunk / GSPAN);  // groups per sub-chunk
  constexpr int NS = (int)(TILE / kTkSubChunk);  // sub-chunks per tile
  static_assert(TILE % kTkSubChunk == 0 && kTkSubChunk % GSPAN == 0, "whole sub-chunks of whole groups");
  constexpr uint32_t kAll = (1u << NS) - 1u;
  static_assert(GSPAN < 65536, "packed 16-bit group counts");
  const int64_t need = (int64_t)state[1];
  const int64_t c_gt = k - need;
  const uint32_t T = (uint32_t)state[0];
  const int lane = __lane_id();
  const uint32_t W = gridDim.x * (kTkPlaceThreads / kWave);
  const uint32_t gw = blockIdx.x * (kTkPlaceThreads / kWave) + threadIdx.x / kWave;
  if (gw >= num_tiles) return;
  const uint32_t per = (num_tiles - gw + W - 1) / W;
  for (uint32_t j0 = 0; j0 < per; j0 += kWave) {
    int64_t lgt0 = 0, leq0 = 0;
    uint32_t lsub = kAll;
    bool places = false;
    if (j0 + lane < per) {
      const uint32_t t = gw + (j0 + lane) * W;
      const uint64_t e0 = tcnt[t], e1 = tcnt[t + 1];
      if (tstat) lsub = (uint32_t)(tstat[t] >> 16) & kAll;
      lgt0 = (int64_t)(e0 >> 32);
      leq0 = (int64_t)(e0 & 0xffffffffu);
      const int64_t gth = (int64_t)(e1 >> 32) - lgt0, eqh = (int64_t)(e1 & 0xffffffffu) - leq0;
      places = !(gth == 0 && (eqh == 0 || leq0 >= need));
    }
    uint64_t mask = __ballot(places);
    while (mask != 0) {
      const int b = __builtin_ctzll(mask);
      mask &= mask - 1;
      const uint32_t tile = gw + (j0 + (uint32_t)b) * W;
      int64_t gpos = readlane(lgt0, b), epos = readlane(leq0, b);  // the wave's running offsets
      uint32_t subs = (uint32_t)readlane((int)lsub, b);
      const int64_t t0 = (int64_t)tile * TILE;
      auto batch = [&](int r, auto full_c) {
        constexpr bool FULL = decltype(full_c)::value;
        Packed<Raw, G> xv[B];
#pragma unroll
        for (int q = 0; q < B; ++q) load_items<Raw, G, FULL>(x, t0 + (r + q) * GSPAN + lane * G, n, xv[q]);
#pragma unroll
        for (int q = 0; q < B; ++q) {
          const int64_t gb = t0 + (r + q) * GSPAN + lane * G;
          uint32_t gtb = 0, eqb = 0;
#pragma unroll
          for (int i = 0; i < G; ++i) {
            const uint32_t kk = KT::key(xv[q][i]) ^ flip;
            const bool valid = FULL || gb + i < n;
            gtb |= (valid && kk > T ? 1u : 0u) << i;
            eqb |= (valid && kk == T ? 1u : 0u) << i;
          }
          // Most groups of a placing tile hold no selected key: skip their
          // scan (2^28 fp16, k = 2048: the scan and placement loop of every
          // group made each placing wave a chain of ~5,400 VALU
          // instructions, 53 us per call).
          if (__ballot((gtb | eqb) != 0) == 0) continue;
          const int c = (__builtin_popcount(gtb) << 16) | __builtin_popcount(eqb);
          const int incl = wave_incl_scan(c);
          const int tot = readlane(incl, kWave - 1);
          if ((gtb | eqb) != 0) {
            const int ex = incl - c;
            int64_t gp = gpos + (ex >> 16), ep = epos + (ex & 0xffff);
#pragma unroll
            for (int i = 0; i < G; ++i) {
              if ((gtb >> i) & 1u) {
                cand_v[gp] = xv[q][i];
                cand_i[gp] = (int32_t)(gb + i);
                ++gp;
              } else if ((eqb >> i) & 1u) {
                if (ep < need) {
                  cand_v[c_gt + ep] = xv[q][i];
                  cand_i[c_gt + ep] = (int32_t)(gb + i);
                }
                ++ep;
              }
            }
          }
          gpos += tot >> 16;
          epos += tot & 0xffff;
        }
      };
      // Sub-chunks in order; the ones skipped hold no selected key.
      const bool full_tile = t0 + TILE <= n;
      while (subs != 0) {
        const int c = __builtin_ctz(subs);
        subs &= subs - 1;
        if (full_tile) {
          batch(c * B, std::integral_constant<bool, true>{});
        } else if (t0 + c * kTkSubChunk < n) {
          batch(c * B, std::integral_constant<bool, false>{});
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

// Orders k <= kTkSmallSort selected keys in one workgroup: a bitonic sort
// of 64-bit words (inverted order key << 32 | candidate index), so keys go
// largest first by the order key after flip and equal keys keep candidate
// order, which is input order: the order the stable radix sort gives, in one
// launch instead of five.  The workgroup sorts P = 1024 E words; lane l of
// wave w holds positions w * 64E + 64e + l, e < E, in registers, so the
// compare-exchange steps of stride below 64 are lane shuffles, strides 64 to
// 32E are swaps inside a lane, and only the larger strides go through LDS
// with barriers (2^28 fp16, k = 2048: 36.8 us for the radix sort's five
// launches, 21.7 us with every step in LDS).
constexpr int kTkSmallSort = 4096, kTkSortThreads = 1024;
template <int DT, int E>
__global__ __launch_bounds__(kTkSortThreads) void topk_small_sort_kernel(
    const typename SortKey<DT>::Raw* __restrict__ cand_v, const int32_t* __restrict__ cand_i, int k, uint32_t flip,
    typename SortKey<DT>::Raw* __restrict__ out_v, int32_t* __restrict__ out_i) {
  using KT = SortKey<DT>;
  constexpr int P = kTkSortThreads * E, SEG = kWave * E;
  static_assert(E == 2 || E == 4, "two or four words per lane");
  __shared__ uint64_t s[P];
  const int tid = threadIdx.x, lane = __lane_id();
  const int seg = (tid / kWave) * SEG + lane;  // position of word 0
  uint64_t a[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int pos = seg + e * kWave;
    a[e] = pos < k ? ((uint64_t)(~(KT::key(cand_v[pos]) ^ flip)) << 32) | (uint32_t)pos : ~0ull;
  }
  // Swaps inside the lane: words e and e + D (stride 64 D).
  auto cas_regs = [&](int size, auto d_c) {
    constexpr int D = decltype(d_c)::value;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      if ((e & D) == 0) {
        const bool up = ((seg + e * kWave) & size) == 0;
        const uint64_t x0 = a[e], x1 = a[e + D];
        if ((x0 > x1) == up) {
          a[e] = x1;
          a[e + D] = x0;
        }
      }
    }
  };
  for (int size = 2; size <= P; size <<= 1) {
    int stride = size >> 1;
    if (stride >= SEG) {
#pragma unroll
      for (int e = 0; e < E; ++e) s[seg + e * kWave] = a[e];
      __syncthreads();
      for (; stride >= SEG; stride >>= 1) {
        for (int t = tid; t < P / 2; t += kTkSortThreads) {
          const int lo = 2 * t - (t & (stride - 1));  // the pair (lo, lo + stride), bit `stride` of lo clear
          const bool up = (lo & size) == 0;            // this bitonic block ascends
          const uint64_t x0 = s[lo], x1 = s[lo + stride];
          if ((x0 > x1) == up) {
            s[lo] = x1;
            s[lo + stride] = x0;
          }
        }
        __syncthreads();
      }
      // Each thread reads back only its own positions, which no thread
      // touches again before the barrier of the next LDS phase.
#pragma unroll
      for (int e = 0; e < E; ++e) a[e] = s[seg + e * kWave];
    }
    if constexpr (E == 4) {
      if (stride == 2 * kWave) {
        cas_regs(size, std::integral_constant<int, 2>{});
        stride >>= 1;
      }
    }
    if (stride == kWave) {
      cas_regs(size, std::integral_constant<int, 1>{});
      stride >>= 1;
    }
    for (; stride > 0; stride >>= 1) {
      const bool lo = (lane & stride) == 0;
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const bool up = ((seg + e * kWave) & size) == 0;
        const uint32_t blo = (uint32_t)__shfl_xor((int)(uint32_t)a[e], stride);
        const uint32_t bhi = (uint32_t)__shfl_xor((int)(uint32_t)(a[e] >> 32), stride);
        const uint64_t b = ((uint64_t)bhi << 32) | blo;
        const uint64_t mn = a[e] < b ? a[e] : b, mx = a[e] < b ? b : a[e];
        a[e] = lo == up ? mn : mx;
      }
    }
  }
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int pos = seg + e * kWave;
    if (pos < k) {
      const uint32_t c = (uint32_t)a[e];
      out_v[pos] = cand_v[c];
      out_i[pos] = cand_i[c];
    }
  }
}

// Placement walks the tiles with the resident grid of topk_place_kernel:
// most tiles place nothing when k is small, and each that does costs a few
// dependent round trips to memory, which the resident waves pay in
// parallel.  2^28 fp16, k = 2048: a workgroup of 8 waves per tile (block
// scans between the waves) took 42 us with 2048 workgroups of which about
// 512 fit at once and 52 us with the resident grid.
template <int DT>
static uint32_t topk_place_grid() {
  static std::atomic<int> cached[kMaxCachedDevices];
  return (uint32_t)device_cached(cached, [](int dev) {
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, topk_place_kernel<DT>, kTkPlaceThreads, 0) !=
            hipSuccess ||
        per_cu < 1)
      per_cu = 1;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) cus = 256;
    return per_cu * cus;
  });
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
  size_t cand_i_off, tcnt_off, tstat_off, tlist_off, sort_scratch_off, scratch;
  int groups;
  int64_t chunk;
};

inline TopkLayout topk_layout(int64_t n, int64_t k, int dt) {
  TopkLayout L{};
  const uint64_t tiles = (uint64_t)((n + kTkTileMin - 1) / kTkTileMin);
  int64_t groups = (n + 65535) / 65536;
  if (groups > kTkMaxGroups) groups = kTkMaxGroups;
  if (groups < 1) groups = 1;
  // Histogram chunks are whole tiles (the last round records per tile).
  const int64_t tile = topk_tile(topk_elem_bytes(dt));
  int64_t chunk = (n + groups - 1) / groups;
  chunk = (chunk + tile - 1) / tile * tile;
  groups = n > 0 ? (n + chunk - 1) / chunk : 1;
  L.groups = (int)groups;
  L.chunk = chunk;
  L.state_off = lookback_ws_bytes(tiles);
  L.part_off = align256(L.state_off + 256);
  L.psum_off = align256(L.part_off + (size_t)groups * 256 * 8);
  L.own_total = align256(L.psum_off + (size_t)kTkSumGroups * 256 * 8);
  L.cand_i_off = align256((size_t)k * topk_elem_bytes(dt));
  L.tcnt_off = align256(L.cand_i_off + (size_t)k * 4);
  L.tstat_off = align256(L.tcnt_off + (size_t)(tiles + 1) * 8);
  L.tlist_off = align256(L.tstat_off + (size_t)tiles * 8);
  L.sort_scratch_off = align256(L.tlist_off + (size_t)tiles * kTkListMax);
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
  constexpr int64_t kTile = TopkGeom<DT>::kTile;
  static_assert(kTile == (int64_t)kTkNW * (sizeof(Raw) <= 2 ? 4 : 2) * kWave * kTkG, "topk_tile");
  if (L.chunk % kTile != 0 || L.chunk / kTile > kTkMaxChunkTiles) return TCS_ERR_INVALID_ARG;
  const uint32_t tiles = (uint32_t)((n + kTile - 1) / kTile);
  auto* tcnt = reinterpret_cast<uint64_t*>(sc + L.tcnt_off);
  auto* tstat = reinterpret_cast<uint64_t*>(sc + L.tstat_off);
  auto* tlist = reinterpret_cast<uint8_t*>(sc + L.tlist_off);
  constexpr bool kFuse = P > 1;  // keys of one byte: the tile counts read the keys
  for (int r = P - 1; r >= 0; --r) {
    if (kFuse && r == 0) {
      hipLaunchKernelGGL((topk_hist_kernel<DT, kFuse>), dim3(L.groups), dim3(kTkThreads), 0, st, xr, n, L.chunk,
                         flip, r, state, partial, tstat, tlist);
    } else {
      hipLaunchKernelGGL((topk_hist_kernel<DT, false>), dim3(L.groups), dim3(kTkThreads), 0, st, xr, n, L.chunk,
                         flip, r, state, partial, nullptr, nullptr);
    }
    TCS_CHECK_LAUNCH();
    hipLaunchKernelGGL(topk_sum_kernel, dim3(sum_groups), dim3(256), 0, st, partial, L.groups, sum_per, psum);
    TCS_CHECK_LAUNCH();
    hipLaunchKernelGGL(topk_pick_kernel, dim3(1), dim3(kTkPickThreads), 0, st, psum, sum_groups, r,
                       r == P - 1 ? 1 : 0, k, state);
    TCS_CHECK_LAUNCH();
  }
  if constexpr (kFuse) {
    hipLaunchKernelGGL((topk_tilefix_kernel<DT>), dim3((tiles + 3) / 4), dim3(256), 0, st, xr, n, flip, state, tstat,
                       tlist, tcnt, tiles);
  } else {
    hipLaunchKernelGGL((topk_tilecount_kernel<DT>), dim3(tiles), dim3(kTkNW * kWave), 0, st, xr, n, flip, state,
                       tcnt);
  }
  TCS_CHECK_LAUNCH();
  hipLaunchKernelGGL(topk_tilescan_kernel, dim3(1), dim3(kTkScanThreads), 0, st, tcnt, tiles);
  TCS_CHECK_LAUNCH();
  const uint32_t resident = topk_place_grid<DT>();
  const uint32_t wave_groups = (tiles + kTkPlaceThreads / kWave - 1) / (kTkPlaceThreads / kWave);  // a wave per tile
  const uint32_t place_grid = wave_groups < resident ? wave_groups : resident;
  hipLaunchKernelGGL((topk_place_kernel<DT>), dim3(place_grid), dim3(kTkPlaceThreads), 0, st, xr, n, flip, state, k,
                     tcnt, kFuse ? tstat : nullptr, cand_v, cand_i, tiles);
  TCS_CHECK_LAUNCH();
  // Order the k selected: largest first (smallest first for !largest), ties
  // in input order.  The radix sort reuses this workspace (same header and
  // epoch).
  if (k <= kTkSmallSort) {
    if (k <= kTkSortThreads * 2) {
      hipLaunchKernelGGL((topk_small_sort_kernel<DT, 2>), dim3(1), dim3(kTkSortThreads), 0, st, cand_v, cand_i,
                         (int)k, flip, static_cast<Raw*>(out_v), out_i);
    } else {
      hipLaunchKernelGGL((topk_small_sort_kernel<DT, 4>), dim3(1), dim3(kTkSortThreads), 0, st, cand_v, cand_i,
                         (int)k, flip, static_cast<Raw*>(out_v), out_i);
    }
    TCS_CHECK_LAUNCH();
    return TCS_OK;
  }
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
