// Stable LSD radix sort on gfx950 with 8-bit digits, returning the sorted
// values and their original positions (the reference's run_radix_sort,
// src/torch/torch_sort.h:21-81, src/radix_sort.cpp: a 1-bit split per bit,
// 16 passes, after an order-preserving re-encoding of fp16).
//
// Order: torch.sort(stable=True): ascending (or descending) by value, ties in
// input order; floats order -0.0 and +0.0 as equal and every NaN after +inf
// (before everything when descending).  Values move unchanged, so the
// output holds the input's exact bits; the order key is recomputed from
// them in every pass.
//
// Launches: one histogram pass over the keys counts the digits of every
// digit position at once; a small kernel turns the counts into the
// first index of each digit's bucket; then one kernel per digit position.
// A pass kernel ranks each tile's keys by digit (per wave: one ballot per
// digit bit gives each lane the lanes holding the same digit, so ranks are
// stable without atomics), publishes the tile's 256 digit counts, finds
// how many keys of each digit the tiles before it hold through a per-digit
// decoupled look-back, reorders the tile in LDS by digit and writes each
// digit's run contiguously.
#include "tcs_common.hpp"
#include "tcs_device.hpp"
#include "tcs_sortkey.hpp"

namespace tcs {

constexpr int kSortNW = 8, kSortItems = 16;
constexpr int64_t kSortTile = (int64_t)kSortNW * kWave * kSortItems;  // 8192
constexpr int kHistThreads = 512, kHistMaxGroups = 1024;
constexpr int kMaxPasses = 4;

// ---------------------------------------------------------------------------
// Digit counts of every pass: per-workgroup partials (no atomics in HBM),
// then one workgroup per pass sums them in a fixed order and scans them.
// ---------------------------------------------------------------------------
template <int DT>
__global__ __launch_bounds__(kHistThreads) void sort_hist_kernel(const typename SortKey<DT>::Raw* __restrict__ x,
                                                                 int64_t n, int64_t chunk, uint32_t flip,
                                                                 uint64_t* __restrict__ partial) {
  using KT = SortKey<DT>;
  using Raw = typename KT::Raw;
  constexpr int P = KT::kBits / 8;
  constexpr int NWV = kHistThreads / kWave;
  constexpr int V = 16 / (int)sizeof(Raw);
  // C copies of the counts per wave, lane l adding into copy l % C: lanes
  // whose keys share a digit (skewed data) then contend C times less.
  constexpr int C = 4 / P > 0 ? 4 / P : 1;
  constexpr int NH = NWV * C;
  // One word of padding per copy: the copies of one bin fall in different
  // LDS banks.
  __shared__ uint32_t h[NH][P * 256 + 1];
  const int tid = threadIdx.x, wid = tid / kWave;
  const int hc = wid * C + (__lane_id() % C);
  for (int k = tid; k < NH * (P * 256 + 1); k += kHistThreads) (&h[0][0])[k] = 0;
  __syncthreads();
  const int64_t c0 = (int64_t)blockIdx.x * chunk;
  const int64_t c1 = c0 + chunk < n ? c0 + chunk : n;
  // Four 16-byte loads in flight per thread per step; steps that lie wholly
  // inside the chunk skip the per-element bounds checks.
  constexpr int U = 4;
  constexpr int64_t STEP = (int64_t)kHistThreads * V;
  auto add = [&](Raw r) {
    const uint32_t k = KT::key(r) ^ flip;
#pragma unroll
    for (int p = 0; p < P; ++p) atomicAdd(&h[hc][p * 256 + ((k >> (8 * p)) & 0xffu)], 1u);
  };
  int64_t b0 = c0 + (int64_t)tid * V;
  const int64_t full_end = c0 + ((c1 - c0) / (U * STEP)) * (U * STEP);  // steps [c0, full_end) are whole
  for (; b0 < full_end; b0 += U * STEP) {
    Raw r[U][V];
#pragma unroll
    for (int u = 0; u < U; ++u) load_items<Raw, V, true>(x, b0 + u * STEP, c1, r[u]);
#pragma unroll
    for (int u = 0; u < U; ++u) {
#pragma unroll
      for (int i = 0; i < V; ++i) add(r[u][i]);
    }
  }
  for (; b0 < c1; b0 += STEP) {
    Raw r[V];
    load_items<Raw, V, false>(x, b0, c1, r);
#pragma unroll
    for (int i = 0; i < V; ++i) {
      if (b0 + i < c1) add(r[i]);
    }
  }
  __syncthreads();
  for (int e = tid; e < P * 256; e += kHistThreads) {
    uint32_t s = 0;
#pragma unroll
    for (int w = 0; w < NH; ++w) s += h[w][e];
    partial[(size_t)blockIdx.x * P * 256 + e] = s;  // 64-bit: high half zero
  }
}

// First level of the digit-count sum: workgroup (q, p) adds pass p's
// partials of groups [q * per, (q + 1) * per) into psum (same layout as the
// partials, kSortSumGroups rows; 64-bit words, high half zero).
constexpr int kSortSumGroups = 32;
__global__ __launch_bounds__(256) void sort_sum_kernel(const uint64_t* __restrict__ partial, int groups, int per,
                                                       int passes, uint64_t* __restrict__ psum) {
  const int d = threadIdx.x, q = blockIdx.x, p = blockIdx.y;
  const int g0 = q * per;
  const int g1 = g0 + per < groups ? g0 + per : groups;
  uint32_t s = 0;
  int g = g0;
  for (; g + 8 <= g1; g += 8) {  // eight loads in flight
    uint64_t v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = partial[((size_t)(g + u) * passes + p) * 256 + d];
#pragma unroll
    for (int u = 0; u < 8; ++u) s += (uint32_t)v[u];
  }
  for (; g < g1; ++g) s += (uint32_t)partial[((size_t)g * passes + p) * 256 + d];
  psum[((size_t)q * passes + p) * 256 + d] = s;
}

// gbase[p * 256 + d] = number of keys whose digit p is below d.  One
// workgroup per pass: four threads per digit sum interleaved groups (eight
// loads in flight each), then the 256 totals are scanned.
constexpr int kBucketThreads = 1024;
__global__ __launch_bounds__(kBucketThreads) void sort_bucket_kernel(const uint64_t* __restrict__ partial,
                                                                     int groups, int passes,
                                                                     uint64_t* __restrict__ gbase) {
  const int p = blockIdx.x, tid = threadIdx.x, d = tid & 255, r = tid >> 8;
  uint32_t s = 0;
  int g = r;
  for (; g + 28 < groups; g += 32) {
    uint64_t v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = partial[((size_t)(g + 4 * u) * passes + p) * 256 + d];
#pragma unroll
    for (int u = 0; u < 8; ++u) s += (uint32_t)v[u];
  }
  for (; g < groups; g += 4) s += (uint32_t)partial[((size_t)g * passes + p) * 256 + d];
  __shared__ uint32_t sr[4][256];
  __shared__ uint32_t ws[4];
  sr[r][d] = s;
  __syncthreads();
  uint32_t t = 0, incl = 0;
  if (tid < 256) {
    t = sr[0][d] + sr[1][d] + sr[2][d] + sr[3][d];
    incl = wave_incl_scan(t);
    if (__lane_id() == kWave - 1) ws[tid / kWave] = incl;
  }
  __syncthreads();
  if (tid < 256) {
    uint32_t before = 0;
    for (int w = 0; w < tid / kWave; ++w) before += ws[w];
    gbase[p * 256 + d] = (uint64_t)(before + incl - t);
  }
}

// ---------------------------------------------------------------------------
// One digit position.  Tile t holds elements [t * 8192, t * 8192 + 8192):
// wave w, item i, lane l is element t * 8192 + w * 1024 + i * 64 + l, so the
// (item, lane) order of a wave is the input order.
// ---------------------------------------------------------------------------
template <int DT, bool GEN_IDX>
__global__ __launch_bounds__(kSortNW * kWave) void sort_pass_kernel(
    const typename SortKey<DT>::Raw* __restrict__ kin, const int32_t* __restrict__ vin,
    typename SortKey<DT>::Raw* __restrict__ kout, int32_t* __restrict__ vout, const uint64_t* __restrict__ gbase,
    int shift, uint32_t flip, int64_t n, void* ws, uint32_t num_tiles) {
  using KT = SortKey<DT>;
  using Raw = typename KT::Raw;
  constexpr int NW = kSortNW, IT = kSortItems;
  __shared__ uint32_t s_whist[NW][256];  // per wave: digit counts, then their exclusive prefix over waves
  __shared__ uint32_t s_excl[256];       // first tile-local position of each digit
  __shared__ int32_t s_gofs[256];        // output index of tile-local position 0 for each digit
  __shared__ uint32_t s_wsum[4];
  __shared__ LaunchTag s_tag;
  __shared__ Raw s_keys[kSortTile];
  __shared__ int32_t s_vals[kSortTile];
  // Digits beside the reordered keys for keys of up to 2 bytes; 4-byte keys
  // recompute them (the extra 8 KiB would leave one workgroup per CU).
  constexpr bool kDigLds = sizeof(Raw) <= 2;
  __shared__ uint8_t s_dig[kDigLds ? kSortTile : 1];

  const int tid = threadIdx.x, lane = __lane_id(), wid = tid / kWave;
  if (tid == 0) s_tag = launch_arrive(ws);
  for (int k = tid; k < NW * 256; k += NW * kWave) (&s_whist[0][0])[k] = 0;
  const uint32_t tile = blockIdx.x;
  const int64_t t0 = (int64_t)tile * kSortTile;
  const int64_t wb = t0 + (int64_t)wid * (kWave * IT);
  auto digit_of = [&](Raw r) { return (int)(((KT::key(r) ^ flip) >> shift) & 0xffu); };

  // A tile wholly inside [0, n) takes the unchecked path (every tile but
  // the last).
  const bool full = t0 + kSortTile <= n;
  Raw kr[IT];
  int32_t vv[IT];
  uint32_t dp[IT];  // digit << 16 | position within the wave's digit bucket
  auto load_rank = [&](auto full_c) {
    constexpr bool FULL = decltype(full_c)::value;
#pragma unroll
    for (int i = 0; i < IT; ++i) {
      const int64_t p = wb + i * kWave + lane;
      const bool valid = FULL || p < n;
      kr[i] = valid ? kin[p] : Raw(0);
      if constexpr (GEN_IDX) {
        vv[i] = (int32_t)p;
      } else {
        vv[i] = valid ? vin[p] : 0;
      }
    }
    __syncthreads();

    // Stable rank of every item among the wave's items of the same digit.
    const uint64_t lt = (1ull << lane) - 1ull;
#pragma unroll
    for (int i = 0; i < IT; ++i) {
      const bool valid = FULL || wb + i * kWave + lane < n;
      const int d = digit_of(kr[i]);
      uint64_t peers = FULL ? ~0ull : __ballot(valid);
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        const bool bit = (d >> b) & 1;
        const uint64_t m = __ballot(bit);
        peers &= bit ? m : ~m;
      }
      const uint32_t rank = (uint32_t)__popcll(peers & lt);
      uint32_t base = 0;
      if (valid) {
        base = s_whist[wid][d];
        if (rank == 0) s_whist[wid][d] = base + (uint32_t)__popcll(peers);
      }
      dp[i] = ((uint32_t)d << 16) | (base + rank);
    }
  };
  if (full) {
    load_rank(std::integral_constant<bool, true>{});
  } else {
    load_rank(std::integral_constant<bool, false>{});
  }
  __syncthreads();

  uint64_t* desc = ws_desc(ws);
  const uint32_t tag = s_tag.tag;
  const int max_polls = s_tag.polls;
  uint32_t cnt = 0, incl = 0;
  if (tid < 256) {
    const int d = tid;
    uint32_t run = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      const uint32_t c = s_whist[w][d];
      s_whist[w][d] = run;
      run += c;
    }
    cnt = run;
    // Publish this tile's count of digit d at once (tile 0's is inclusive).
    if (num_tiles > 1)
      desc_publish<uint32_t>(desc, (uint64_t)tile * 256 + d, cnt, tag, tile == 0 ? kFlagIncl : kFlagAgg);
    incl = wave_incl_scan(cnt);
    if (lane == kWave - 1) s_wsum[wid] = incl;
  }
  __syncthreads();
  if (tid < 256) {
    const int d = tid;
    uint32_t before = 0;
    for (int w = 0; w < wid; ++w) before += s_wsum[w];
    const uint32_t excl = before + incl - cnt;
    s_excl[d] = excl;
    // Keys of digit d in the tiles before this one: walk back to the
    // nearest inclusive descriptor, one predecessor at a time (reading
    // several at once measured slower: the extra device-scope loads cost
    // more than the walk saves).  A predecessor that has not published
    // after the poll limit (kFallbackPolls by default) is counted from its keys here.
    uint32_t prefix = 0;
    for (int64_t q = (int64_t)tile - 1; q >= 0; --q) {
      uint32_t v = 0;
      int flag = 0;
      int polls = 0;
      while (!desc_try_read<uint32_t>(desc, (uint64_t)q * 256 + d, tag, v, flag)) {
        if (polls++ >= max_polls) {
          const int64_t u0 = q * kSortTile;
          const int64_t u1 = u0 + kSortTile < n ? u0 + kSortTile : n;
          v = 0;
          for (int64_t e = u0; e < u1; ++e) v += digit_of(kin[e]) == d ? 1u : 0u;
          flag = q == 0 ? (int)kFlagIncl : (int)kFlagAgg;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
      prefix += v;
      if (flag == (int)kFlagIncl) break;
    }
    if (tile > 0 && tile + 1 < num_tiles) desc_publish<uint32_t>(desc, (uint64_t)tile * 256 + d, prefix + cnt, tag, kFlagIncl);
    s_gofs[d] = (int32_t)((uint32_t)gbase[d] + prefix) - (int32_t)excl;
  }
  __syncthreads();

  // Reorder the tile by digit in LDS (stable), then write each digit's run.
#pragma unroll
  for (int i = 0; i < IT; ++i) {
    if (full || wb + i * kWave + lane < n) {
      const int d = (int)(dp[i] >> 16);
      const uint32_t pos = s_excl[d] + s_whist[wid][d] + (dp[i] & 0xffffu);
      s_keys[pos] = kr[i];
      s_vals[pos] = vv[i];
      if constexpr (kDigLds) s_dig[pos] = (uint8_t)d;
    }
  }
  __syncthreads();
  const int tvalid = full ? (int)kSortTile : (int)(n - t0);
  for (int pos = tid; pos < tvalid; pos += NW * kWave) {
    const Raw r = s_keys[pos];
    const int d = kDigLds ? (int)s_dig[pos] : digit_of(r);
    const int64_t dst = (int64_t)s_gofs[d] + pos;
    kout[dst] = r;
    vout[dst] = s_vals[pos];
  }
}

// Workspace: header, 256 descriptors per tile, bucket starts, histogram
// partials and their first-level sums (zeroed once and reusable, like every
// workspace; counts are kept in 64-bit words with a zero high half, which
// never read as a valid descriptor when a later call uses the same bytes
// for descriptors).
// Scratch: one temporary copy of keys and positions (any contents).
struct SortLayout {
  size_t desc_off, gbase_off, part_off, psum_off, total;
  size_t tmpv_off, scratch;
  int groups;
  int64_t chunk;
};

inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

inline int sort_passes(int dt) {
  switch (dt) {
    case TCS_I8: return 1;
    case TCS_F16: case TCS_BF16: case TCS_I16: return 2;
    case TCS_F32: case TCS_I32: case TCS_U32: return 4;
    default: return 0;
  }
}

inline int sort_elem_bytes(int dt) { return dt == TCS_I8 ? 1 : (sort_passes(dt) == 2 ? 2 : 4); }

inline SortLayout sort_layout(int64_t n, int dt) {
  SortLayout L{};
  const int P = sort_passes(dt);
  const uint64_t tiles = (uint64_t)((n + kSortTile - 1) / kSortTile);
  int64_t groups = (n + 65535) / 65536;
  if (groups > kHistMaxGroups) groups = kHistMaxGroups;
  if (groups < 1) groups = 1;
  int64_t chunk = (n + groups - 1) / groups;
  chunk = (chunk + 63) & ~(int64_t)63;  // 16-byte aligned chunk starts
  L.groups = (int)groups;
  L.chunk = chunk;
  L.desc_off = kHeaderBytes;
  L.gbase_off = align256(L.desc_off + (size_t)tiles * 256 * 8);
  L.part_off = align256(L.gbase_off + (size_t)kMaxPasses * 256 * 8);
  L.psum_off = align256(L.part_off + (size_t)groups * P * 256 * 8);
  L.total = align256(L.psum_off + (size_t)kSortSumGroups * P * 256 * 8);
  L.tmpv_off = align256((size_t)n * sort_elem_bytes(dt));
  L.scratch = align256(L.tmpv_off + (size_t)n * 4);
  return L;
}

template <int DT>
static int sort_run(const void* x, const int32_t* idx_in, void* out_x, int32_t* out_idx, int64_t n, bool desc,
                    void* ws, void* scratch, hipStream_t st) {
  using KT = SortKey<DT>;
  using Raw = typename KT::Raw;
  constexpr int P = KT::kBits / 8;
  const SortLayout L = sort_layout(n, DT);
  char* w = static_cast<char*>(ws);
  auto* gbase = reinterpret_cast<uint64_t*>(w + L.gbase_off);
  auto* partial = reinterpret_cast<uint64_t*>(w + L.part_off);
  auto* tk = reinterpret_cast<Raw*>(scratch);
  auto* tv = reinterpret_cast<int32_t*>(static_cast<char*>(scratch) + L.tmpv_off);
  const uint32_t flip = desc ? (KT::kBits == 32 ? 0xffffffffu : ((1u << KT::kBits) - 1u)) : 0u;
  hipLaunchKernelGGL((sort_hist_kernel<DT>), dim3(L.groups), dim3(kHistThreads), 0, st, static_cast<const Raw*>(x),
                     n, L.chunk, flip, partial);
  TCS_CHECK_LAUNCH();
  auto* psum = reinterpret_cast<uint64_t*>(w + L.psum_off);
  const int sum_per = (L.groups + kSortSumGroups - 1) / kSortSumGroups;
  const int sum_groups = (L.groups + sum_per - 1) / sum_per;
  hipLaunchKernelGGL(sort_sum_kernel, dim3(sum_groups, P), dim3(256), 0, st, partial, L.groups, sum_per, P, psum);
  TCS_CHECK_LAUNCH();
  hipLaunchKernelGGL(sort_bucket_kernel, dim3(P), dim3(kBucketThreads), 0, st, psum, sum_groups, P, gbase);
  TCS_CHECK_LAUNCH();
  const uint32_t tiles = (uint32_t)((n + kSortTile - 1) / kSortTile);
  const Raw* kin = static_cast<const Raw*>(x);
  const int32_t* vin = idx_in;
  for (int p = 0; p < P; ++p) {
    // The last pass writes the outputs; earlier ones alternate so that it does.
    const bool to_out = ((P - 1 - p) % 2) == 0;
    Raw* kout = to_out ? static_cast<Raw*>(out_x) : tk;
    int32_t* vout = to_out ? out_idx : tv;
    if (p == 0 && !idx_in) {
      hipLaunchKernelGGL((sort_pass_kernel<DT, true>), dim3(tiles), dim3(kSortNW * kWave), 0, st, kin, vin, kout,
                         vout, gbase + p * 256, 8 * p, flip, n, ws, tiles);
    } else {
      hipLaunchKernelGGL((sort_pass_kernel<DT, false>), dim3(tiles), dim3(kSortNW * kWave), 0, st, kin, vin, kout,
                         vout, gbase + p * 256, 8 * p, flip, n, ws, tiles);
    }
    TCS_CHECK_LAUNCH();
    kin = kout;
    vin = vout;
  }
  return TCS_OK;
}

}  // namespace tcs

extern "C" {

size_t tcs_sort_workspace_bytes(int64_t n, int dtype) {
  if (n < 0 || tcs::sort_passes(dtype) == 0) return 0;
  return tcs::sort_layout(n, dtype).total;
}

size_t tcs_sort_scratch_bytes(int64_t n, int dtype) {
  if (n < 0 || tcs::sort_passes(dtype) == 0) return 0;
  return tcs::sort_layout(n, dtype).scratch;
}

int tcs_sort(const void* x, int dtype, const int32_t* idx_in, void* out_x, int32_t* out_idx, int64_t n,
             int descending, void* workspace, size_t workspace_bytes, void* scratch, size_t scratch_bytes,
             void* stream) {
  using namespace tcs;
  if (sort_passes(dtype) == 0) return TCS_ERR_UNSUPPORTED_DTYPE;
  if (n < 0 || n >= ((int64_t)1 << 31)) return TCS_ERR_INVALID_ARG;
  if (n == 0) return TCS_OK;
  if (!x || !out_x || !out_idx) return TCS_ERR_INVALID_ARG;
  if (!aligned16(x) || (idx_in && !aligned16(idx_in))) return TCS_ERR_INVALID_ARG;
  const SortLayout L = sort_layout(n, dtype);
  if (!workspace || workspace_bytes < L.total || !scratch || scratch_bytes < L.scratch) return TCS_ERR_WORKSPACE;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const bool d = descending != 0;
  switch (dtype) {
    case TCS_F16: return sort_run<TCS_F16>(x, idx_in, out_x, out_idx, n, d, workspace, scratch, st);
    case TCS_BF16: return sort_run<TCS_BF16>(x, idx_in, out_x, out_idx, n, d, workspace, scratch, st);
    case TCS_F32: return sort_run<TCS_F32>(x, idx_in, out_x, out_idx, n, d, workspace, scratch, st);
    case TCS_I8: return sort_run<TCS_I8>(x, idx_in, out_x, out_idx, n, d, workspace, scratch, st);
    case TCS_I16: return sort_run<TCS_I16>(x, idx_in, out_x, out_idx, n, d, workspace, scratch, st);
    case TCS_I32: return sort_run<TCS_I32>(x, idx_in, out_x, out_idx, n, d, workspace, scratch, st);
    default: return sort_run<TCS_U32>(x, idx_in, out_x, out_idx, n, d, workspace, scratch, st);
  }
}

}  // extern "C"
