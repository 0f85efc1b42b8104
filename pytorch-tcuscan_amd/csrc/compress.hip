// Stream compaction and CSR gather on gfx950.
//
// Compaction keeps x[i] where mask[i] == 1 exactly (the reference's
// CompareScalar EQ 1, src/kernels/kernel_compress.h:167-168), in order.
// One pass: each tile counts its kept elements (lane-serial, then DPP wave
// scans and a block scan), the tile's output offset arrives through the
// decoupled look-back on counts, each wave stages its kept elements in LDS
// in output order and writes its contiguous output range with coalesced
// stores (lane k of a store writes element k of a 64-element run).  Optionally the kept positions (int32) or the
// matching entries of an index array are written beside the values
// (run_compress_ind / _no_arange).  A separate count pass sizes the outputs
// when the caller does not know the count (the reference syncs on a reduce
// for the same purpose, src/torch/torch_compress.h:62-67).
//
// CSR gather: z[i] = vals[i] * x[cols[i]] rounded once to the input dtype
// (fp16/fp32: the exact product rounded to nearest even, as torch does;
// int16: two's-complement wrap), following
// src/kernels/kernel_csr_gather.h:122-137.  Columns outside [0, x_len)
// yield 0 rather than an out-of-bounds read.
#include "tcs_counted.hpp"

namespace tcs {

constexpr int kCmpNW = 8, kCmpG = 8;
// The smallest tile of any compaction (1 group per wave); workspaces are
// sized for this many tiles.
constexpr int64_t kCmpTile = (int64_t)kCmpNW * 1 * kWave * kCmpG;  // 4096

// Which elements are kept: mask[i] == 1 (KeepMask), or a comparison of
// x[i] with a pivot in x's own type (KeepCmp: the filters and `where`, with
// the pivot rounded to that type on the host, as the reference rounds it to
// half, src/compare.cpp greater_equal_fp16).  A NaN passes only "!=".
struct KeepMask {
  static constexpr bool kFromX = false;
  const int8_t* mask;
  template <class XT>
  __device__ __forceinline__ bool at(const XT* __restrict__, int64_t p) const { return mask[p] == 1; }
};

template <int DT, int CMP>  // DT: TCS_F16 / TCS_BF16 / TCS_F32; CMP: a tcs_cmp value
struct KeepCmp {
  static constexpr bool kFromX = true;
  float pivot;
  template <class XT>
  __device__ __forceinline__ bool of(XT bits) const {
    float v;
    if constexpr (DT == TCS_F16) {
      v = (float)__builtin_bit_cast(_Float16, (uint16_t)bits);
    } else if constexpr (DT == TCS_BF16) {
      v = __uint_as_float((uint32_t)bits << 16);
    } else {
      v = __uint_as_float((uint32_t)bits);
    }
    if constexpr (CMP == TCS_CMP_LT) return v < pivot;
    if constexpr (CMP == TCS_CMP_GT) return v > pivot;
    if constexpr (CMP == TCS_CMP_EQ) return v == pivot;
    if constexpr (CMP == TCS_CMP_LE) return v <= pivot;
    if constexpr (CMP == TCS_CMP_GE) return v >= pivot;
    return v != pivot;
  }
  template <class XT>
  __device__ __forceinline__ bool at(const XT* __restrict__ x, int64_t p) const { return of(x[p]); }
};

// IDX: 0 = values only, 1 = gather idx_in, 2 = positions (int32),
// 3 = positions + 1 (int32; the row ends of flag-delimited segments),
// 4 = positions only (int32, no values: `where`).
//
// Tile geometry: NW waves x J groups x 64 lanes x G items.  J is 8 for
// inputs of up to 2 bytes per element and 4 above (fewer, larger tiles
// amortise the per-tile look-back: fp32 values + mask + int32 indices,
// 0.615 ms per 2^28 at J = 2 vs 0.588 at 4), while one wave's LDS staging
// stays within 8 KiB.
template <int XB, int IB, int IDX, bool FROMX>
struct CmpGeom {
  static constexpr int kInBytes = XB + (FROMX ? 0 : 1) + (IDX == 1 ? IB : 0);
  static constexpr int kStageBytes = IDX == 0 ? XB : (IDX == 4 ? IB : (XB > IB ? XB : IB));
  static constexpr int kByInput = kInBytes <= 2 ? 8 : 4;
  static constexpr int kByLds = 16 / kStageBytes;
  static constexpr int J = kByInput < kByLds ? kByInput : kByLds;
  static constexpr int64_t kTile = (int64_t)kCmpNW * J * kWave * kCmpG;
  static_assert(kTile >= kCmpTile, "workspaces are sized for 4096-element tiles");
};

template <int XB, int IB, int IDX, class Keep>
__global__ __launch_bounds__(kCmpNW * kWave) void compress_kernel(
    const typename Bits<XB>::T* __restrict__ x, const typename Bits<IB>::T* __restrict__ idx_in,
    Keep keep, typename Bits<XB>::T* __restrict__ out_x,
    typename Bits<IB>::T* __restrict__ out_idx, int64_t n, int64_t cap, void* ws, uint32_t num_tiles) {
  using XT = typename Bits<XB>::T;
  using IT = typename Bits<IB>::T;
  using Geom = CmpGeom<XB, IB, IDX, Keep::kFromX>;
  constexpr int G = kCmpG, J = Geom::J, NW = kCmpNW;
  constexpr int64_t TILE = Geom::kTile;
  constexpr int64_t GSPAN = (int64_t)kWave * G, WSEG = GSPAN * J;
  using ST = typename Bits<Geom::kStageBytes>::T;
  __shared__ int s_wval[NW], s_wexcl[NW];
  __shared__ int64_t s_base;
  __shared__ LaunchTag s_tag;
  __shared__ ST s_stage[NW][J * G * kWave];  // one wave's kept elements, in order

  if (threadIdx.x == 0) s_tag = launch_arrive(ws);
  const int lane = __lane_id();
  const int wid = threadIdx.x / kWave;
  const uint32_t tile = blockIdx.x;
  const int64_t t0 = (int64_t)tile * TILE;
  const int64_t wbase = t0 + (int64_t)wid * WSEG + (int64_t)lane * G;
  const bool full = t0 + TILE <= n;

  uint32_t mb[J];  // bit i: item i of group j is kept
  Packed<XT, G> xv[J];
  IT iv[IDX == 1 ? J : 1][G];
#pragma unroll
  for (int j = 0; j < J; ++j) {
    const int64_t gb = wbase + j * GSPAN;
    Packed<int8_t, G> m;
    if (full) {
      if constexpr (!Keep::kFromX) load_items<int8_t, G, true>(keep.mask, gb, n, m);
      load_items<XT, G, true>(x, gb, n, xv[j]);
      if constexpr (IDX == 1) load_items<IT, G, true>(idx_in, gb, n, iv[j]);
    } else {
      if constexpr (!Keep::kFromX) load_items<int8_t, G, false>(keep.mask, gb, n, m);
      load_items<XT, G, false>(x, gb, n, xv[j]);
      if constexpr (IDX == 1) load_items<IT, G, false>(idx_in, gb, n, iv[j]);
    }
    uint32_t b = 0;
#pragma unroll
    for (int i = 0; i < G; ++i) {
      bool k;
      if constexpr (Keep::kFromX) {
        k = (full || gb + i < n) && keep.of(xv[j][i]);
      } else {
        k = m[i] == 1;  // items past n load as 0
      }
      b |= (k ? 1u : 0u) << i;
    }
    mb[j] = b;
  }

  // Kept counts: per group, the lane's exclusive offset inside the wave.
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
      for (int64_t p = u0 + lane; p < u1; p += kWave) c += keep.at(x, p) ? 1 : 0;
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
    // Counts are 32-bit: n < 2^31 per call (checked on the host).
    if (lane == 0) s_base = carry;
  }
  __syncthreads();

  // This wave's kept elements go to [base, base + wtot), capped at cap.
  const int64_t base = s_base + s_wexcl[wid];
  ST* stage = s_stage[wid];
  const int64_t room = cap - base;
  const int wn = room <= 0 ? 0 : (room < wtot ? (int)room : wtot);
  auto flush = [&](auto* out) {
    using OT = typename std::remove_pointer<decltype(out)>::type;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    for (int k = lane; k < wn; k += kWave) __builtin_nontemporal_store((OT)stage[k], out + base + k);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  };
  if constexpr (IDX != 4) {
#pragma unroll
    for (int j = 0; j < J; ++j) {
      int pos = gex[j];
#pragma unroll
      for (int i = 0; i < G; ++i)
        if ((mb[j] >> i) & 1u) stage[pos++] = (ST)xv[j][i];
    }
    flush(out_x);
  }
  if constexpr (IDX != 0) {
#pragma unroll
    for (int j = 0; j < J; ++j) {
      int pos = gex[j];
      const int64_t gb = wbase + j * GSPAN;
#pragma unroll
      for (int i = 0; i < G; ++i) {
        if ((mb[j] >> i) & 1u) {
          if constexpr (IDX == 1) stage[pos] = (ST)iv[j][i];
          if constexpr (IDX == 2 || IDX == 4) stage[pos] = (ST)(IT)(gb + i);
          if constexpr (IDX == 3) stage[pos] = (ST)(IT)(gb + i + 1);
          ++pos;
        }
      }
    }
    flush(out_idx);
  }
}

// Filters (a comparison with x decides what is kept) run persistently.
// Tile geometry: NW waves x J groups x 64 lanes x G items, about 16-36 KiB
// of input per tile; the kernel holds two tiles in registers and one
// wave's kept elements of one tile in LDS (at most 8 KiB per wave).  Mask
// compactions measured faster with one tile per workgroup
// (compress_kernel): 0.68 vs 0.78 ms for run_compress_ind at 2^28, while
// the fp16 filter runs 0.32 ms here vs 0.385 there.  The fp16 filter's
// 16K-element tiles measured far better than 32K ones (0.315 vs 1.12 ms).
template <int XB, int IB, int IDX, bool FROMX>
struct FltGeom {
  static constexpr int kInBytes = XB + (FROMX ? 0 : 1) + (IDX == 1 ? IB : 0);
  static constexpr int kStageBytes = IDX == 0 ? XB : (IDX == 4 ? IB : (XB > IB ? XB : IB));
  static constexpr int kByInput = kInBytes <= 2 ? 4 : (kInBytes <= 9 ? 2 : 1);
  static constexpr int kByLds = 16 / kStageBytes;
  static constexpr int J = kByInput < kByLds ? kByInput : kByLds;
  static constexpr int64_t kTile = (int64_t)kCmpNW * J * kWave * kCmpG;
  static_assert(kTile >= kCmpTile, "workspaces are sized for 4096-element tiles");
};

// One tile as a data wave holds it: values, kept bits, the mask bytes or
// the index entries it came with.
template <class XT, class IT, int J, int G, bool MASK, bool IDX>
struct CmpRegs {
  Packed<XT, G> xv[J];
  Packed<int8_t, G> mk[MASK ? J : 1];
  Packed<IT, G> iv[IDX ? J : 1];
  uint32_t mb[J];  // bit i: item i of group j is kept
};

// LDS state of a tile in flight (two alternate).
template <int NW>
struct CmpSlot {
  int wval[NW];   // kept per data wave
  int wexcl[NW];  // kept in the waves before
  int agg;        // kept in the tile
  int carry;      // kept in the tiles before
  int count;      // data waves that have counted the tile
};

// Persistent filter compaction with NW data waves and one look-back wave.
// Workgroups draw tiles from the ticket counters of the workspace header
// exactly as the scan does (workgroup b draws from counter b % K, which
// hands out tiles k, k + K, ...; one draw ahead), so every tile below the
// smallest undrawn one is held by a running workgroup or done.  One step
// per tile t, with the next tile tn going into the other register set:
// while the look-back wave walks back for t's carry (and publishes t's
// inclusive count), the data waves load and count tn and the last of them
// to finish publishes tn's aggregate at once (an LDS counter, no
// barrier); after one barrier the data waves stage t's kept elements in
// LDS in output order and write them with coalesced stores.
template <int XB, int IB, int IDX, class Keep>
__global__ __launch_bounds__((kCmpNW + 1) * kWave) void filter_kernel(
    const typename Bits<XB>::T* __restrict__ x, const typename Bits<IB>::T* __restrict__ idx_in,
    Keep keep, typename Bits<XB>::T* __restrict__ out_x,
    typename Bits<IB>::T* __restrict__ out_idx, int64_t n, int64_t cap, void* ws, uint32_t num_tiles) {
  using XT = typename Bits<XB>::T;
  using IT = typename Bits<IB>::T;
  using Geom = FltGeom<XB, IB, IDX, Keep::kFromX>;
  constexpr int G = kCmpG, J = Geom::J, NW = kCmpNW;
  constexpr bool kMask = !Keep::kFromX;
  constexpr int64_t TILE = Geom::kTile;
  constexpr int64_t GSPAN = (int64_t)kWave * G, WSEG = GSPAN * J;
  using ST = typename Bits<Geom::kStageBytes>::T;
  using Regs = CmpRegs<XT, IT, J, G, kMask, IDX == 1>;
  __shared__ CmpSlot<NW> slot[2];
  __shared__ uint32_t s_draw[2];
  __shared__ LaunchTag s_tag;
  __shared__ ST s_stage[NW][J * G * kWave];  // one wave's kept elements, in order

  const int lane = __lane_id();
  const int wid = threadIdx.x / kWave;  // wave NW is the look-back wave
  const bool lb_wave = wid == NW;
  const uint32_t T = num_tiles;
  const uint32_t K = gridDim.x < (uint32_t)kTicketCounters ? gridDim.x : (uint32_t)kTicketCounters;
  const uint32_t grp = blockIdx.x % K;
  uint64_t* ctr = ws_slot(ws, (int)grp);
  const uint64_t grp_tiles = ((uint64_t)T - grp + K - 1) / K;
  const uint64_t grp_wgs = (gridDim.x - grp + K - 1) / K;
  const uint64_t last_draw = grp_tiles + grp_wgs - 1;
  auto draw = [&]() -> uint32_t {
    const uint64_t v = __hip_atomic_fetch_add(ctr, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (v == last_draw) __hip_atomic_store(ctr, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return v >= grp_tiles ? T : (uint32_t)(v * K + grp);
  };

  auto load = [&](Regs& r, uint32_t t) {
    const int64_t t0 = (int64_t)t * TILE;
    const int64_t wbase = t0 + (int64_t)wid * WSEG + (int64_t)lane * G;
    if (t0 + TILE <= n) {
#pragma unroll
      for (int j = 0; j < J; ++j) {
        const int64_t gb = wbase + j * GSPAN;
        if constexpr (kMask) load_items<int8_t, G, true>(keep.mask, gb, n, r.mk[j]);
        load_items<XT, G, true>(x, gb, n, r.xv[j]);
        if constexpr (IDX == 1) load_items<IT, G, true>(idx_in, gb, n, r.iv[j]);
      }
    } else {
#pragma unroll
      for (int j = 0; j < J; ++j) {
        const int64_t gb = wbase + j * GSPAN;
        if constexpr (kMask) load_items<int8_t, G, false>(keep.mask, gb, n, r.mk[j]);
        load_items<XT, G, false>(x, gb, n, r.xv[j]);
        if constexpr (IDX == 1) load_items<IT, G, false>(idx_in, gb, n, r.iv[j]);
      }
    }
  };
  // The last data wave to count a tile: the waves' exclusive offsets and
  // the tile's aggregate.
  auto finalize = [&](CmpSlot<NW>& sl, uint32_t t) {
    const int a = lane < NW ? sl.wval[lane] : 0;
    const int incl = wave_incl_scan(a);
    const int agg = readlane(incl, NW - 1);
    if (lane < NW) sl.wexcl[lane] = incl - a;
    if (lane == 0) {
      sl.agg = agg;
      sl.count = 0;
      uint64_t* desc = ws_desc(ws);
      if (t == 0) {
        if (T > 1) desc_publish<int>(desc, 0, agg, s_tag.tag, kFlagIncl);
      } else {
        desc_publish<int>(desc, t, agg, s_tag.tag, kFlagAgg);
      }
    }
  };
  // Kept bits, per-group exclusive offsets inside the wave, wave total.
  auto count = [&](Regs& r, int (&gex)[J], CmpSlot<NW>& sl, uint32_t t) {
    const int64_t wbase = (int64_t)t * TILE + (int64_t)wid * WSEG + (int64_t)lane * G;
    // Kept bits; a tile wholly inside [0, n) needs no bounds checks.
    auto bits = [&](auto full_c) {
      constexpr bool FULL = decltype(full_c)::value;
#pragma unroll
      for (int j = 0; j < J; ++j) {
        const int64_t gb = wbase + j * GSPAN;
        uint32_t b = 0;
#pragma unroll
        for (int i = 0; i < G; ++i) {
          bool k;
          if constexpr (kMask) {
            k = r.mk[j][i] == 1;  // items past n load as 0
          } else {
            k = (FULL || gb + i < n) && keep.of(r.xv[j][i]);
          }
          b |= (k ? 1u : 0u) << i;
        }
        r.mb[j] = b;
      }
    };
    if ((int64_t)(t + 1) * TILE <= n) {
      bits(std::integral_constant<bool, true>{});
    } else {
      bits(std::integral_constant<bool, false>{});
    }
    int wtot = 0;
#pragma unroll
    for (int j = 0; j < J; ++j) {
      const int c = __builtin_popcount(r.mb[j]);
      const int incl = wave_incl_scan(c);
      gex[j] = wtot + incl - c;
      wtot += readlane(incl, kWave - 1);
    }
    int last = 0;
    if (lane == 0) {
      sl.wval[wid] = wtot;
      last = __hip_atomic_fetch_add(&sl.count, 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP) == NW - 1;
    }
    if (readlane(last, 0)) finalize(sl, t);
  };
  // Look-back wave: kept elements in the tiles before t.  Counts are 32-bit:
  // n < 2^31 per call (checked on the host).
  auto walk = [&](CmpSlot<NW>& sl, uint32_t t) {
    auto fb = [&](int64_t u) {
      const int64_t u0 = u * TILE;
      const int64_t u1 = u0 + TILE < n ? u0 + TILE : n;
      int c = 0;
      for (int64_t p = u0 + lane; p < u1; p += kWave) c += keep.at(x, p) ? 1 : 0;
      return FlagVal<int>{u == 0 ? 1 : 0, wave_sum(c)};
    };
    int carry = 0;
    if (t > 0) {
      carry = lookback<int>(ws_desc(ws), t, s_tag.tag, s_tag.polls, fb);
      if (lane == 0) desc_publish<int>(ws_desc(ws), t, carry + sl.agg, s_tag.tag, kFlagIncl);
    }
    if (lane == 0) sl.carry = carry;
  };
  // Stage this wave's kept elements of t in output order, then write them.
  auto write = [&](const Regs& r, const int (&gex)[J], const CmpSlot<NW>& sl, uint32_t t) {
    const int64_t base = (int64_t)sl.carry + sl.wexcl[wid];
    const int wtot = sl.wval[wid];
    const int64_t wbase = (int64_t)t * TILE + (int64_t)wid * WSEG + (int64_t)lane * G;
    ST* stage = s_stage[wid];
    const int64_t room = cap - base;
    const int wn = room <= 0 ? 0 : (room < wtot ? (int)room : wtot);
    auto flush = [&](auto* out) {
      using OT = typename std::remove_pointer<decltype(out)>::type;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      for (int k = lane; k < wn; k += kWave) __builtin_nontemporal_store((OT)stage[k], out + base + k);
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    };
    if constexpr (IDX != 4) {
#pragma unroll
      for (int j = 0; j < J; ++j) {
        int pos = gex[j];
#pragma unroll
        for (int i = 0; i < G; ++i)
          if ((r.mb[j] >> i) & 1u) stage[pos++] = (ST)r.xv[j][i];
      }
      flush(out_x);
    }
    if constexpr (IDX != 0) {
#pragma unroll
      for (int j = 0; j < J; ++j) {
        int pos = gex[j];
        const int64_t gb = wbase + j * GSPAN;
#pragma unroll
        for (int i = 0; i < G; ++i) {
          if ((r.mb[j] >> i) & 1u) {
            if constexpr (IDX == 1) stage[pos] = (ST)r.iv[j][i];
            if constexpr (IDX == 2 || IDX == 4) stage[pos] = (ST)(IT)(gb + i);
            if constexpr (IDX == 3) stage[pos] = (ST)(IT)(gb + i + 1);
            ++pos;
          }
        }
      }
      flush(out_idx);
    }
  };

  if (threadIdx.x < 2) slot[threadIdx.x].count = 0;
  if (lb_wave && lane == 0) {
    s_tag = launch_arrive(ws);
    const uint32_t d0 = draw();
    s_draw[0] = d0;
    s_draw[1] = d0 < T ? draw() : T;
  }
  __syncthreads();
  uint32_t t = s_draw[0];
  uint32_t tn = s_draw[1];
  if (t >= T) return;

  Regs reg[2];
  int gex[2][J];
  if (!lb_wave) {
    load(reg[0], t);
    count(reg[0], gex[0], slot[0], t);
  }
  __syncthreads();

  auto step = [&](auto parity) -> bool {
    constexpr int P = decltype(parity)::value;
    constexpr int Q = 1 - P;
    if (lb_wave) {
      walk(slot[P], t);
      if (lane == 0) s_draw[P] = tn < T ? draw() : T;
    } else if (tn < T) {
      load(reg[Q], tn);
      count(reg[Q], gex[Q], slot[Q], tn);
    }
    __syncthreads();
    if (!lb_wave) write(reg[P], gex[P], slot[P], t);
    t = tn;
    tn = s_draw[P];
    return t < T;
  };
  for (;;) {
    if (!step(std::integral_constant<int, 0>{})) break;
    if (!step(std::integral_constant<int, 1>{})) break;
  }
}

// Counted compaction: the layout and helpers are in tcs_counted.hpp.

// Block b (256 threads, 64K elements): wave w counts chunks w, w + 4, ...,
// each lane 32 mask bytes of a chunk, four chunks' loads in flight at a
// time.  Plain loads: the compaction that follows reads the mask again and
// finds part of it in the caches (count + compaction of fp32 values with
// int32 indices at 2^28: 0.584 ms with nontemporal mask loads in both
// passes, 0.547 ms with plain ones).
__global__ __launch_bounds__(kCountBlock) void count_mask_chunks_kernel(const int8_t* __restrict__ mask,
                                                                        int64_t n, int64_t* __restrict__ partial,
                                                                        uint64_t* __restrict__ chunk_cnt) {
  const int lane = __lane_id();
  const int wid = threadIdx.x / kWave;
  constexpr int kWaves = kCountBlock / kWave;
  constexpr int kPerWave = kChunksPerBlock / kWaves;  // 8
  constexpr int kBatch = 4;
  const int64_t b0 = (int64_t)blockIdx.x * kCountChunk;
  uint32_t mine = 0;  // lane i keeps the count of chunk wid + 4i
  uint32_t tot = 0;
#pragma unroll
  for (int i0 = 0; i0 < kPerWave; i0 += kBatch) {
    u32x4 v[kBatch][2];
#pragma unroll
    for (int i = 0; i < kBatch; ++i) {
      const int64_t base = b0 + (int64_t)(wid + kWaves * (i0 + i)) * kChunk + lane * 32;
      if (base + 32 <= n) {
        const u32x4* p = reinterpret_cast<const u32x4*>(mask + base);
        v[i][0] = p[0];
        v[i][1] = p[1];
      } else {
        uint32_t w[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          uint32_t word = 0;
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int64_t e = base + q * 4 + r;
            word |= (e < n ? (uint32_t)(uint8_t)mask[e] : 0u) << (8 * r);
          }
          w[q] = word;
        }
        v[i][0] = u32x4{w[0], w[1], w[2], w[3]};
        v[i][1] = u32x4{w[4], w[5], w[6], w[7]};
      }
    }
#pragma unroll
    for (int i = 0; i < kBatch; ++i) {
      uint32_t k = 0;
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int q = 0; q < 4; ++q) k += ones_in_word(v[i][h][q]);
      const uint32_t c = (uint32_t)wave_sum((int)k);
      tot += c;
      mine = lane == i0 + i ? c : mine;
    }
  }
  if (lane < kPerWave) {
    // Chunks past n are not written and never read.
    const int64_t ch = (int64_t)blockIdx.x * kChunksPerBlock + wid + kWaves * lane;
    if (ch * kChunk < n) chunk_cnt[ch] = mine;
  }
  __shared__ uint32_t s[kWaves];
  if (lane == 0) s[wid] = tot;
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t t = 0;
#pragma unroll
    for (int k = 0; k < kWaves; ++k) t += s[k];
    partial[blockIdx.x] = t;
  }
}

// The chunk counts of a comparison (x cmp pivot) in the same layout: block
// b (256 threads, 64K elements), wave w counts chunks w, w + 4, ..., each
// lane 32 consecutive elements of a chunk; two chunks' loads in flight at
// a time (plain loads, for the compaction's second read of x).  Elements
// past n are not counted.
template <class XT, class Keep>
__global__ __launch_bounds__(kCountBlock) void count_keep_chunks_kernel(const XT* __restrict__ x, Keep keep,
                                                                        int64_t n, int64_t* __restrict__ partial,
                                                                        uint64_t* __restrict__ chunk_cnt) {
  const int lane = __lane_id();
  const int wid = threadIdx.x / kWave;
  constexpr int kWaves = kCountBlock / kWave;
  constexpr int kPerWave = kChunksPerBlock / kWaves;  // 8
  constexpr int kBatch = 2;
  constexpr int kItems = (int)(kChunk / kWave);  // 32 per lane
  const int64_t b0 = (int64_t)blockIdx.x * kCountChunk;
  uint32_t mine = 0;  // lane i keeps the count of chunk wid + 4i
  uint32_t tot = 0;
#pragma unroll
  for (int i0 = 0; i0 < kPerWave; i0 += kBatch) {
    Packed<XT, kItems> v[kBatch];
    int lim[kBatch];  // this lane's items below n (kItems except at the tail)
#pragma unroll
    for (int i = 0; i < kBatch; ++i) {
      const int64_t base = b0 + (int64_t)(wid + kWaves * (i0 + i)) * kChunk + lane * kItems;
      if (base + kItems <= n) {
        load_items<XT, kItems, true>(x, base, n, v[i]);
        lim[i] = kItems;
      } else {
        load_items<XT, kItems, false>(x, base, n, v[i]);
        lim[i] = base >= n ? 0 : (int)(n - base);
      }
    }
#pragma unroll
    for (int i = 0; i < kBatch; ++i) {
      uint32_t k = 0;
      if (lim[i] == kItems) {
#pragma unroll
        for (int q = 0; q < kItems; ++q) k += keep.of(v[i][q]) ? 1u : 0u;
      } else {
#pragma unroll
        for (int q = 0; q < kItems; ++q) k += (q < lim[i] && keep.of(v[i][q])) ? 1u : 0u;
      }
      const uint32_t c = (uint32_t)wave_sum((int)k);
      tot += c;
      mine = lane == i0 + i ? c : mine;
    }
  }
  if (lane < kPerWave) {
    const int64_t ch = (int64_t)blockIdx.x * kChunksPerBlock + wid + kWaves * lane;
    if (ch * kChunk < n) chunk_cnt[ch] = mine;
  }
  __shared__ uint32_t s[kWaves];
  if (lane == 0) s[wid] = tot;
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t t = 0;
#pragma unroll
    for (int k = 0; k < kWaves; ++k) t += s[k];
    partial[blockIdx.x] = t;
  }
}

// One workgroup: exclusive offsets of the block counts and their total, in
// rounds of 4096 counts (four consecutive per thread, so each wave reads
// and writes 2 KiB contiguously).
constexpr int kScanBlock = 1024;
__global__ __launch_bounds__(kScanBlock) void count_scan_kernel(const int64_t* __restrict__ partial, int64_t np,
                                                                int64_t* __restrict__ block_off,
                                                                int64_t* __restrict__ out) {
  constexpr int kWaves = kScanBlock / kWave, kPer = 4;
  __shared__ int64_t s_w[kWaves];
  __shared__ int64_t s_carry;
  const int lane = __lane_id(), wid = threadIdx.x / kWave;
  if (threadIdx.x == 0) s_carry = 0;
  __syncthreads();
  for (int64_t r0 = 0; r0 < np; r0 += (int64_t)kScanBlock * kPer) {
    const int64_t k0 = r0 + (int64_t)threadIdx.x * kPer;
    int64_t v[kPer];
    int64_t t = 0;
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
      v[q] = k0 + q < np ? partial[k0 + q] : 0;
      t += v[q];
    }
    const int64_t incl = wave_incl_scan(t);
    if (lane == kWave - 1) s_w[wid] = incl;
    __syncthreads();
    int64_t before = s_carry, all = 0;
#pragma unroll
    for (int q = 0; q < kWaves; ++q) {
      before += q < wid ? s_w[q] : 0;
      all += s_w[q];
    }
    int64_t run = before + incl - t;
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
      if (k0 + q < np) block_off[k0 + q] = run;
      run += v[q];
    }
    __syncthreads();  // every thread has read s_w and s_carry
    if (threadIdx.x == 0) s_carry += all;
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = s_carry;
}

// Wave-per-chunk compaction using the counts left by the count pass.  Keep
// is KeepMask (the mask array) or KeepCmp (x compared with a pivot, in
// registers: the filters).
template <int XB, int IB, int IDX, class Keep>
__global__ __launch_bounds__(kCountBlock) void compress_counted_kernel(
    const typename Bits<XB>::T* __restrict__ x, const typename Bits<IB>::T* __restrict__ idx_in,
    Keep keep, typename Bits<XB>::T* __restrict__ out_x,
    typename Bits<IB>::T* __restrict__ out_idx, int64_t n, int64_t cap, const int64_t* __restrict__ block_off,
    const uint64_t* __restrict__ chunk_cnt) {
  using XT = typename Bits<XB>::T;
  using IT = typename Bits<IB>::T;
  constexpr int G = 8, J = (int)(kChunk / (kWave * G));  // 4 groups of 8 per lane
  constexpr int kWaves = kCountBlock / kWave;
  constexpr int kStageBytes = IDX == 0 ? XB : IDX == 4 ? IB : (XB > IB ? XB : IB);
  using ST = typename Bits<kStageBytes>::T;
  __shared__ ST s_stage[kWaves][kChunk];
  const int lane = __lane_id();
  const int wid = threadIdx.x / kWave;
  // Chunks from the last one down: the count pass read its input front to
  // back, so the caches hold its end, which this order reads first (filters
  // of 2^28 fp16: 0.2355 -> 0.2243 ms per call; mask compactions unchanged).
  const int64_t nch = (n + kChunk - 1) / kChunk;
  const int64_t ch = nch - 1 - ((int64_t)blockIdx.x * kWaves + wid);
  if (ch < 0) return;
  const int64_t c0 = ch * kChunk;
  const int64_t part = chunk_base_part(ch, block_off, chunk_cnt, lane);

  const bool full = c0 + kChunk <= n;
  uint32_t mb[J];
  Packed<XT, G> xv[J];
  Packed<IT, G> iv[IDX == 1 ? J : 1];
#pragma unroll
  for (int j = 0; j < J; ++j) {
    const int64_t gb = c0 + j * (kWave * G) + lane * G;
    Packed<int8_t, G> m;
    if (full) {
      // Values and indices that only this pass reads: nontemporal loads
      // (0.573 -> 0.533 ms for fp32 values + int32 indices at 2^28).  The
      // mask, and a filter's x, were read by the count pass just before:
      // plain loads, which find part of them in the caches.
      if constexpr (!Keep::kFromX) load_items<int8_t, G, true>(keep.mask, gb, n, m);
      if constexpr (Keep::kFromX) {
        load_items<XT, G, true>(x, gb, n, xv[j]);
      } else {
        load_nt(x, gb, xv[j]);
      }
      if constexpr (IDX == 1) load_nt(idx_in, gb, iv[j]);
    } else {
      if constexpr (!Keep::kFromX) load_items<int8_t, G, false>(keep.mask, gb, n, m);
      load_items<XT, G, false>(x, gb, n, xv[j]);
      if constexpr (IDX == 1) load_items<IT, G, false>(idx_in, gb, n, iv[j]);
    }
    uint32_t b = 0;
    if constexpr (Keep::kFromX) {
#pragma unroll
      for (int i = 0; i < G; ++i) b |= (keep.of(xv[j][i]) && (full || gb + i < n) ? 1u : 0u) << i;
    } else {
      // Four mask bytes per word (items past n load as 0).
#pragma unroll
      for (int w = 0; w < G / 4; ++w) b |= ones_bits4(m.w[w]) << (4 * w);
    }
    mb[j] = b;
  }
  const int64_t base = wave_sum(part);
  int gex[J];
  int wtot = 0;
#pragma unroll
  for (int j = 0; j < J; ++j) {
    const int c = __builtin_popcount(mb[j]);
    const int incl = wave_incl_scan(c);
    gex[j] = wtot + incl - c;
    wtot += readlane(incl, kWave - 1);
  }
  ST* stage = s_stage[wid];
  const int64_t room = cap - base;
  // A workspace that does not hold this input's counts (no count pass just
  // before) can give any base: nothing outside [0, cap) is written.
  const int wn = (room <= 0 || base < 0) ? 0 : (room < wtot ? (int)room : wtot);
  auto flush = [&](auto* out) {
    wave_lds_sync();
    flush_run(stage, out, base, wn, lane);
    wave_lds_sync();
  };
  if constexpr (IDX != 4) {
#pragma unroll
    for (int j = 0; j < J; ++j) {
      int pos = gex[j];
#pragma unroll
      for (int i = 0; i < G; ++i)
        if ((mb[j] >> i) & 1u) stage[pos++] = (ST)xv[j][i];
    }
    flush(out_x);
  }
  if constexpr (IDX != 0) {
#pragma unroll
    for (int j = 0; j < J; ++j) {
      int pos = gex[j];
      const int64_t gb = c0 + j * (kWave * G) + lane * G;
#pragma unroll
      for (int i = 0; i < G; ++i) {
        if ((mb[j] >> i) & 1u) {
          if constexpr (IDX == 1) stage[pos] = (ST)iv[j][i];
          if constexpr (IDX == 2 || IDX == 4) stage[pos] = (ST)(IT)(gb + i);
          if constexpr (IDX == 3) stage[pos] = (ST)(IT)(gb + i + 1);
          ++pos;
        }
      }
    }
    flush(out_idx);
  }
}

template <int XB, int IB, int IDX, class Keep>
static int launch_counted(const void* x, const void* idx_in, Keep keep, void* out_x, void* out_idx,
                          int64_t n, int64_t cap, void* ws, hipStream_t st) {
  using XT = typename Bits<XB>::T;
  using IT = typename Bits<IB>::T;
  const CountedLayout l = counted_layout(n);
  const int64_t nch = (n + kChunk - 1) / kChunk;
  const int64_t blocks = (nch + kCountBlock / kWave - 1) / (kCountBlock / kWave);
  char* w = static_cast<char*>(ws);
  hipLaunchKernelGGL((compress_counted_kernel<XB, IB, IDX, Keep>), dim3((uint32_t)blocks), dim3(kCountBlock), 0, st,
                     static_cast<const XT*>(x), static_cast<const IT*>(idx_in), keep, static_cast<XT*>(out_x),
                     static_cast<IT*>(out_idx), n, cap, reinterpret_cast<const int64_t*>(w + l.block_off),
                     reinterpret_cast<const uint64_t*>(w + l.chunk_cnt));
  TCS_CHECK_LAUNCH();
  return TCS_OK;
}

template <int XB>
static int counted_idx(int idx_mode, int idx_bytes, const void* x, const void* idx_in, const int8_t* mask,
                       void* out_x, void* out_idx, int64_t n, int64_t cap, void* ws, hipStream_t st) {
  const KeepMask keep{mask};
  if (idx_mode == 0) return launch_counted<XB, 4, 0>(x, nullptr, keep, out_x, nullptr, n, cap, ws, st);
  if (idx_mode == 2) return launch_counted<XB, 4, 2>(x, nullptr, keep, out_x, out_idx, n, cap, ws, st);
  if (idx_mode == 3) return launch_counted<XB, 4, 3>(x, nullptr, keep, out_x, out_idx, n, cap, ws, st);
  if (idx_bytes == 4) return launch_counted<XB, 4, 1>(x, idx_in, keep, out_x, out_idx, n, cap, ws, st);
  if (idx_bytes == 8) return launch_counted<XB, 8, 1>(x, idx_in, keep, out_x, out_idx, n, cap, ws, st);
  return TCS_ERR_UNSUPPORTED_DTYPE;
}

template <class T, int KIND>
__device__ __forceinline__ T gather_mul(T v, T g);
template <>
__device__ __forceinline__ uint16_t gather_mul<uint16_t, 0>(uint16_t v, uint16_t g) {  // fp16
  const float p = (float)__builtin_bit_cast(_Float16, v) * (float)__builtin_bit_cast(_Float16, g);
  return __builtin_bit_cast(uint16_t, (_Float16)p);
}
template <>
__device__ __forceinline__ uint16_t gather_mul<uint16_t, 1>(uint16_t v, uint16_t g) {  // int16
  return (uint16_t)((int32_t)(int16_t)v * (int32_t)(int16_t)g);
}
template <>
__device__ __forceinline__ uint32_t gather_mul<uint32_t, 2>(uint32_t v, uint32_t g) {  // fp32
  return __float_as_uint(__uint_as_float(v) * __uint_as_float(g));
}

template <class T, int KIND>
__global__ __launch_bounds__(256) void csr_gather_kernel(const T* __restrict__ vals,
                                                         const int32_t* __restrict__ cols,
                                                         const T* __restrict__ x, T* __restrict__ z,
                                                         int64_t nnz, int64_t x_len) {
  constexpr int G = 8;
  const int64_t b = ((int64_t)blockIdx.x * 256 + threadIdx.x) * G;
  if (b >= nnz) return;
  T v[G], o[G];
  int32_t c[G];
  if (b + G <= nnz) {
    load_items<T, G, true>(vals, b, nnz, v);
    load_items<int32_t, G, true>(cols, b, nnz, c);
  } else {
    load_items<T, G, false>(vals, b, nnz, v);
    load_items<int32_t, G, false>(cols, b, nnz, c);
  }
#pragma unroll
  for (int i = 0; i < G; ++i) {
    const T g = (uint32_t)c[i] < (uint64_t)x_len ? x[c[i]] : T(0);
    o[i] = gather_mul<T, KIND>(v[i], g);
  }
  if (b + G <= nnz) {
    store_items<T, G, true>(z, b, nnz, o);
  } else {
    store_items<T, G, false>(z, b, nnz, o);
  }
}

template <int XB, int IB, int IDX, class Keep = KeepMask>
static int launch_compress(const void* x, const void* idx_in, Keep keep, void* out_x,
                           void* out_idx, int64_t n, int64_t cap, void* ws, uint32_t,
                           hipStream_t st) {
  using XT = typename Bits<XB>::T;
  using IT = typename Bits<IB>::T;
  // The workspace holds descriptors for n / kCmpTile tiles, at least as many
  // as either kernel's tiles need.
  if constexpr (!Keep::kFromX) {
    constexpr int64_t kTile = CmpGeom<XB, IB, IDX, false>::kTile;
    const uint32_t tiles = (uint32_t)((n + kTile - 1) / kTile);
    hipLaunchKernelGGL((compress_kernel<XB, IB, IDX, Keep>), dim3(tiles), dim3(kCmpNW * kWave), 0, st,
                       static_cast<const XT*>(x), static_cast<const IT*>(idx_in), keep, static_cast<XT*>(out_x),
                       static_cast<IT*>(out_idx), n, cap, ws, tiles);
  } else {
    // Persistent grid: the resident workgroups, at most one per tile.
    constexpr int64_t kTile = FltGeom<XB, IB, IDX, true>::kTile;
    const uint32_t tiles = (uint32_t)((n + kTile - 1) / kTile);
    static std::atomic<int> cached[kMaxCachedDevices];
    uint32_t grid = (uint32_t)device_cached(cached, [](int dev) {
      int b = 0, cus = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, filter_kernel<XB, IB, IDX, Keep>, (kCmpNW + 1) * kWave,
                                                       0) != hipSuccess ||
          b < 1)
        b = 1;
      if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
        cus = 256;
      return b * cus;
    });
    if (grid > tiles) grid = tiles;
    hipLaunchKernelGGL((filter_kernel<XB, IB, IDX, Keep>), dim3(grid), dim3((kCmpNW + 1) * kWave), 0, st,
                       static_cast<const XT*>(x), static_cast<const IT*>(idx_in), keep, static_cast<XT*>(out_x),
                       static_cast<IT*>(out_idx), n, cap, ws, tiles);
  }
  TCS_CHECK_LAUNCH();
  return TCS_OK;
}

// Pivot rounded to the input type (through float first, as a float pivot
// argument is), so the device compares in that type.
static float pivot_in_type(double pivot, int dt) {
  const float f = (float)pivot;
  if (dt == TCS_F16) return (float)(_Float16)f;
  if (dt == TCS_BF16) {
    uint32_t b = __builtin_bit_cast(uint32_t, f);
    if ((b & 0x7f800000u) != 0x7f800000u) b += 0x7fffu + ((b >> 16) & 1u);  // round to nearest even
    return __builtin_bit_cast(float, b & 0xffff0000u);
  }
  return f;
}

template <int DT, int CMP>
static int filter_launch(const void* x, float pv, void* out_vals, int32_t* out_pos, int64_t n, int64_t cap,
                         void* ws, uint32_t tiles, hipStream_t st) {
  constexpr int XB = DT == TCS_F32 ? 4 : 2;
  const KeepCmp<DT, CMP> keep{pv};
  if (!out_pos) return launch_compress<XB, 4, 0>(x, nullptr, keep, out_vals, nullptr, n, cap, ws, tiles, st);
  if (!out_vals) return launch_compress<XB, 4, 4>(x, nullptr, keep, nullptr, out_pos, n, cap, ws, tiles, st);
  return launch_compress<XB, 4, 2>(x, nullptr, keep, out_vals, out_pos, n, cap, ws, tiles, st);
}

// Count pass of a filter: the kept count, and the chunk counts and block
// offsets of the counted layout for tcs_filter_counted.
template <int DT, int CMP>
static int filter_count(const void* x, float pv, int64_t n, int64_t* count_out, void* ws, hipStream_t st) {
  using XT = typename Bits<DT == TCS_F32 ? 4 : 2>::T;
  const CountedLayout l = counted_layout(n);
  const int64_t np = (n + kCountChunk - 1) / kCountChunk;
  char* w = static_cast<char*>(ws);
  if (np > 0) {
    hipLaunchKernelGGL((count_keep_chunks_kernel<XT, KeepCmp<DT, CMP>>), dim3((uint32_t)np), dim3(kCountBlock), 0,
                       st, static_cast<const XT*>(x), KeepCmp<DT, CMP>{pv}, n,
                       reinterpret_cast<int64_t*>(w + l.partial), reinterpret_cast<uint64_t*>(w + l.chunk_cnt));
    TCS_CHECK_LAUNCH();
  }
  hipLaunchKernelGGL(count_scan_kernel, dim3(1), dim3(kScanBlock), 0, st,
                     reinterpret_cast<const int64_t*>(w + l.partial), np,
                     reinterpret_cast<int64_t*>(w + l.block_off), count_out);
  TCS_CHECK_LAUNCH();
  return TCS_OK;
}

// Counted filter: values (IDX 0), positions (IDX 4) or both (IDX 2).
template <int DT, int CMP>
static int filter_counted_launch(const void* x, float pv, void* out_vals, int32_t* out_pos, int64_t n, int64_t cap,
                                 void* ws, hipStream_t st) {
  constexpr int XB = DT == TCS_F32 ? 4 : 2;
  const KeepCmp<DT, CMP> keep{pv};
  if (!out_pos) return launch_counted<XB, 4, 0>(x, nullptr, keep, out_vals, nullptr, n, cap, ws, st);
  if (!out_vals) return launch_counted<XB, 4, 4>(x, nullptr, keep, nullptr, out_pos, n, cap, ws, st);
  return launch_counted<XB, 4, 2>(x, nullptr, keep, out_vals, out_pos, n, cap, ws, st);
}

template <int DT>
static int filter_count_any(int cmp, const void* x, float pv, int64_t n, int64_t* count_out, void* ws,
                            hipStream_t st) {
  switch (cmp) {
    case TCS_CMP_LT: return filter_count<DT, TCS_CMP_LT>(x, pv, n, count_out, ws, st);
    case TCS_CMP_GT: return filter_count<DT, TCS_CMP_GT>(x, pv, n, count_out, ws, st);
    case TCS_CMP_EQ: return filter_count<DT, TCS_CMP_EQ>(x, pv, n, count_out, ws, st);
    case TCS_CMP_LE: return filter_count<DT, TCS_CMP_LE>(x, pv, n, count_out, ws, st);
    case TCS_CMP_GE: return filter_count<DT, TCS_CMP_GE>(x, pv, n, count_out, ws, st);
    default: return filter_count<DT, TCS_CMP_NE>(x, pv, n, count_out, ws, st);
  }
}

template <int XB>
static int compress_idx(int idx_mode, int idx_bytes, const void* x, const void* idx_in,
                        const int8_t* mask, void* out_x, void* out_idx, int64_t n, int64_t cap,
                        void* ws, uint32_t tiles, hipStream_t st) {
  const KeepMask keep{mask};
  if (idx_mode == 0) return launch_compress<XB, 4, 0>(x, nullptr, keep, out_x, nullptr, n, cap, ws, tiles, st);
  if (idx_mode == 2) return launch_compress<XB, 4, 2>(x, nullptr, keep, out_x, out_idx, n, cap, ws, tiles, st);
  if (idx_mode == 3) return launch_compress<XB, 4, 3>(x, nullptr, keep, out_x, out_idx, n, cap, ws, tiles, st);
  if (idx_bytes == 4) return launch_compress<XB, 4, 1>(x, idx_in, keep, out_x, out_idx, n, cap, ws, tiles, st);
  if (idx_bytes == 8) return launch_compress<XB, 8, 1>(x, idx_in, keep, out_x, out_idx, n, cap, ws, tiles, st);
  return TCS_ERR_UNSUPPORTED_DTYPE;
}

}  // namespace tcs

extern "C" {

size_t tcs_compress_workspace_bytes(int64_t n) {
  if (n <= 0) return tcs::lookback_ws_bytes(0);
  const uint64_t tiles = (uint64_t)((n + tcs::kCmpTile - 1) / tcs::kCmpTile);
  const size_t lb = tcs::lookback_ws_bytes(tiles);
  const size_t counted = tcs::counted_layout(n).end;
  return lb > counted ? lb : counted;
}

int tcs_count_mask(const int8_t* mask, int64_t n, int64_t* count_out, void* workspace,
                   size_t workspace_bytes, void* stream) {
  using namespace tcs;
  if (n < 0 || !count_out || (n > 0 && !mask)) return TCS_ERR_INVALID_ARG;
  if (n > 0 && !aligned16(mask)) return TCS_ERR_INVALID_ARG;
  if (n >= ((int64_t)1 << 31)) return TCS_ERR_INVALID_ARG;
  const CountedLayout l = counted_layout(n);
  if (!workspace || workspace_bytes < l.end) return TCS_ERR_WORKSPACE;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int64_t np = (n + kCountChunk - 1) / kCountChunk;
  char* w = static_cast<char*>(workspace);
  if (np > 0) {
    hipLaunchKernelGGL(count_mask_chunks_kernel, dim3((uint32_t)np), dim3(kCountBlock), 0, st, mask, n,
                       reinterpret_cast<int64_t*>(w + l.partial), reinterpret_cast<uint64_t*>(w + l.chunk_cnt));
    TCS_CHECK_LAUNCH();
  }
  hipLaunchKernelGGL(count_scan_kernel, dim3(1), dim3(kScanBlock), 0, st,
                     reinterpret_cast<const int64_t*>(w + l.partial), np,
                     reinterpret_cast<int64_t*>(w + l.block_off), count_out);
  TCS_CHECK_LAUNCH();
  return TCS_OK;
}

int tcs_compress_counted(const void* x, int x_elem_bytes, const void* idx_in, int idx_mode, int idx_elem_bytes,
                         const int8_t* mask, void* out_x, void* out_idx, int64_t n, int64_t out_capacity,
                         void* workspace, size_t workspace_bytes, void* stream) {
  using namespace tcs;
  if (n < 0 || out_capacity < 0 || idx_mode < 0 || idx_mode > 3) return TCS_ERR_INVALID_ARG;
  if (n == 0) return TCS_OK;
  if (!x || !mask || (out_capacity > 0 && !out_x)) return TCS_ERR_INVALID_ARG;
  if (idx_mode != 0 && out_capacity > 0 && !out_idx) return TCS_ERR_INVALID_ARG;
  if (idx_mode == 1 && !idx_in) return TCS_ERR_INVALID_ARG;
  if (idx_mode >= 2 && idx_elem_bytes != 4) return TCS_ERR_UNSUPPORTED_DTYPE;
  if (!aligned16(x) || !aligned16(mask) || (idx_mode == 1 && !aligned16(idx_in))) return TCS_ERR_INVALID_ARG;
  if (n >= ((int64_t)1 << 31)) return TCS_ERR_INVALID_ARG;
  if (!workspace || workspace_bytes < counted_layout(n).end) return TCS_ERR_WORKSPACE;
  hipStream_t st = static_cast<hipStream_t>(stream);
  switch (x_elem_bytes) {
    case 1: return counted_idx<1>(idx_mode, idx_elem_bytes, x, idx_in, mask, out_x, out_idx, n, out_capacity, workspace, st);
    case 2: return counted_idx<2>(idx_mode, idx_elem_bytes, x, idx_in, mask, out_x, out_idx, n, out_capacity, workspace, st);
    case 4: return counted_idx<4>(idx_mode, idx_elem_bytes, x, idx_in, mask, out_x, out_idx, n, out_capacity, workspace, st);
    case 8: return counted_idx<8>(idx_mode, idx_elem_bytes, x, idx_in, mask, out_x, out_idx, n, out_capacity, workspace, st);
    default: return TCS_ERR_UNSUPPORTED_DTYPE;
  }
}

int tcs_count_filter(const void* x, int x_dtype, double pivot, int cmp, int64_t n, int64_t* count_out,
                     void* workspace, size_t workspace_bytes, void* stream) {
  using namespace tcs;
  if (n < 0 || !count_out || cmp < TCS_CMP_LT || cmp > TCS_CMP_NE || (n > 0 && !x)) return TCS_ERR_INVALID_ARG;
  if (n > 0 && !aligned16(x)) return TCS_ERR_INVALID_ARG;
  if (x_dtype != TCS_F16 && x_dtype != TCS_BF16 && x_dtype != TCS_F32) return TCS_ERR_UNSUPPORTED_DTYPE;
  if (!workspace || workspace_bytes < counted_layout(n).end) return TCS_ERR_WORKSPACE;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const float pv = pivot_in_type(pivot, x_dtype);
  switch (x_dtype) {
    case TCS_F16: return filter_count_any<TCS_F16>(cmp, x, pv, n, count_out, workspace, st);
    case TCS_BF16: return filter_count_any<TCS_BF16>(cmp, x, pv, n, count_out, workspace, st);
    default: return filter_count_any<TCS_F32>(cmp, x, pv, n, count_out, workspace, st);
  }
}

int tcs_filter(const void* x, int x_dtype, double pivot, int cmp, void* out_vals, int32_t* out_pos,
               int64_t n, int64_t out_capacity, void* workspace, size_t workspace_bytes, void* stream) {
  using namespace tcs;
  if (n < 0 || out_capacity < 0 || (cmp != TCS_CMP_GE && cmp != TCS_CMP_LE)) return TCS_ERR_INVALID_ARG;
  if (x_dtype != TCS_F16 && x_dtype != TCS_BF16 && x_dtype != TCS_F32) return TCS_ERR_UNSUPPORTED_DTYPE;
  if (n == 0) return TCS_OK;
  if (!x || (!out_vals && !out_pos) || !aligned16(x)) return TCS_ERR_INVALID_ARG;
  if (n >= ((int64_t)1 << 31)) return TCS_ERR_INVALID_ARG;
  const uint64_t tiles = (uint64_t)((n + kCmpTile - 1) / kCmpTile);
  if (!workspace || workspace_bytes < lookback_ws_bytes(tiles)) return TCS_ERR_WORKSPACE;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const uint32_t nt = (uint32_t)tiles;
  const float pv = pivot_in_type(pivot, x_dtype);
  const bool ge = cmp == TCS_CMP_GE;
  switch (x_dtype) {
    case TCS_F16:
      return ge ? filter_launch<TCS_F16, TCS_CMP_GE>(x, pv, out_vals, out_pos, n, out_capacity, workspace, nt, st)
                : filter_launch<TCS_F16, TCS_CMP_LE>(x, pv, out_vals, out_pos, n, out_capacity, workspace, nt, st);
    case TCS_BF16:
      return ge ? filter_launch<TCS_BF16, TCS_CMP_GE>(x, pv, out_vals, out_pos, n, out_capacity, workspace, nt, st)
                : filter_launch<TCS_BF16, TCS_CMP_LE>(x, pv, out_vals, out_pos, n, out_capacity, workspace, nt, st);
    default:
      return ge ? filter_launch<TCS_F32, TCS_CMP_GE>(x, pv, out_vals, out_pos, n, out_capacity, workspace, nt, st)
                : filter_launch<TCS_F32, TCS_CMP_LE>(x, pv, out_vals, out_pos, n, out_capacity, workspace, nt, st);
  }
}

int tcs_filter_counted(const void* x, int x_dtype, double pivot, int cmp, void* out_vals, int32_t* out_pos,
                       int64_t n, int64_t out_capacity, void* workspace, size_t workspace_bytes, void* stream) {
  using namespace tcs;
  if (n < 0 || out_capacity < 0 || (cmp != TCS_CMP_GE && cmp != TCS_CMP_LE)) return TCS_ERR_INVALID_ARG;
  if (x_dtype != TCS_F16 && x_dtype != TCS_BF16 && x_dtype != TCS_F32) return TCS_ERR_UNSUPPORTED_DTYPE;
  if (n == 0) return TCS_OK;
  if (!x || (!out_vals && !out_pos) || !aligned16(x)) return TCS_ERR_INVALID_ARG;
  if (n >= ((int64_t)1 << 31)) return TCS_ERR_INVALID_ARG;
  if (!workspace || workspace_bytes < counted_layout(n).end) return TCS_ERR_WORKSPACE;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const float pv = pivot_in_type(pivot, x_dtype);
  const bool ge = cmp == TCS_CMP_GE;
  switch (x_dtype) {
    case TCS_F16:
      return ge ? filter_counted_launch<TCS_F16, TCS_CMP_GE>(x, pv, out_vals, out_pos, n, out_capacity, workspace, st)
                : filter_counted_launch<TCS_F16, TCS_CMP_LE>(x, pv, out_vals, out_pos, n, out_capacity, workspace, st);
    case TCS_BF16:
      return ge ? filter_counted_launch<TCS_BF16, TCS_CMP_GE>(x, pv, out_vals, out_pos, n, out_capacity, workspace, st)
                : filter_counted_launch<TCS_BF16, TCS_CMP_LE>(x, pv, out_vals, out_pos, n, out_capacity, workspace, st);
    default:
      return ge ? filter_counted_launch<TCS_F32, TCS_CMP_GE>(x, pv, out_vals, out_pos, n, out_capacity, workspace, st)
                : filter_counted_launch<TCS_F32, TCS_CMP_LE>(x, pv, out_vals, out_pos, n, out_capacity, workspace, st);
  }
}

int tcs_compress(const void* x, int x_elem_bytes, const void* idx_in, int idx_mode, int idx_elem_bytes,
                 const int8_t* mask, void* out_x, void* out_idx, int64_t n, int64_t out_capacity,
                 void* workspace, size_t workspace_bytes, void* stream) {
  using namespace tcs;
  if (n < 0 || out_capacity < 0 || idx_mode < 0 || idx_mode > 3) return TCS_ERR_INVALID_ARG;
  if (n == 0) return TCS_OK;
  if (!x || !mask || (out_capacity > 0 && !out_x)) return TCS_ERR_INVALID_ARG;
  if (idx_mode != 0 && out_capacity > 0 && !out_idx) return TCS_ERR_INVALID_ARG;
  if (idx_mode == 1 && !idx_in) return TCS_ERR_INVALID_ARG;
  if (idx_mode >= 2 && idx_elem_bytes != 4) return TCS_ERR_UNSUPPORTED_DTYPE;
  if (!aligned16(x) || !aligned16(mask) || (idx_mode == 1 && !aligned16(idx_in))) return TCS_ERR_INVALID_ARG;
  if (n >= ((int64_t)1 << 31)) return TCS_ERR_INVALID_ARG;
  const uint64_t tiles = (uint64_t)((n + kCmpTile - 1) / kCmpTile);
  if (!workspace || workspace_bytes < lookback_ws_bytes(tiles)) return TCS_ERR_WORKSPACE;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const uint32_t nt = (uint32_t)tiles;
  switch (x_elem_bytes) {
    case 1: return compress_idx<1>(idx_mode, idx_elem_bytes, x, idx_in, mask, out_x, out_idx, n, out_capacity, workspace, nt, st);
    case 2: return compress_idx<2>(idx_mode, idx_elem_bytes, x, idx_in, mask, out_x, out_idx, n, out_capacity, workspace, nt, st);
    case 4: return compress_idx<4>(idx_mode, idx_elem_bytes, x, idx_in, mask, out_x, out_idx, n, out_capacity, workspace, nt, st);
    case 8: return compress_idx<8>(idx_mode, idx_elem_bytes, x, idx_in, mask, out_x, out_idx, n, out_capacity, workspace, nt, st);
    default: return TCS_ERR_UNSUPPORTED_DTYPE;
  }
}

int tcs_csr_gather(const void* vals, int dtype, const int32_t* cols, const void* x, int64_t x_len,
                   void* z, int64_t nnz, void* stream) {
  using namespace tcs;
  if (nnz < 0 || x_len < 0) return TCS_ERR_INVALID_ARG;
  if (nnz == 0) return TCS_OK;
  if (!vals || !cols || !x || !z) return TCS_ERR_INVALID_ARG;
  if (!aligned16(vals) || !aligned16(cols) || !aligned16(z)) return TCS_ERR_INVALID_ARG;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const uint32_t blocks = (uint32_t)((nnz + 256 * 8 - 1) / (256 * 8));
  switch (dtype) {
    case TCS_F16:
      hipLaunchKernelGGL((csr_gather_kernel<uint16_t, 0>), dim3(blocks), dim3(256), 0, st,
                         static_cast<const uint16_t*>(vals), cols, static_cast<const uint16_t*>(x),
                         static_cast<uint16_t*>(z), nnz, x_len);
      break;
    case TCS_I16:
      hipLaunchKernelGGL((csr_gather_kernel<uint16_t, 1>), dim3(blocks), dim3(256), 0, st,
                         static_cast<const uint16_t*>(vals), cols, static_cast<const uint16_t*>(x),
                         static_cast<uint16_t*>(z), nnz, x_len);
      break;
    case TCS_F32:
      hipLaunchKernelGGL((csr_gather_kernel<uint32_t, 2>), dim3(blocks), dim3(256), 0, st,
                         static_cast<const uint32_t*>(vals), cols, static_cast<const uint32_t*>(x),
                         static_cast<uint32_t*>(z), nnz, x_len);
      break;
    default:
      return TCS_ERR_UNSUPPORTED_DTYPE;
  }
  TCS_CHECK_LAUNCH();
  return TCS_OK;
}

}  // extern "C"
