// Segmented sums over CSR row pointers and CSR SpMV on gfx950.
//
// y[r] = sum over row r of v[p], with row r covering [end(r-1), end(r)),
// end(-1) = 0, end(r) = ends[r] (the reference passes indptr + 1 to its
// kernels and treats row 0 as starting at 0: src/torch/torch_seg_ops.h:
// 234-237, 378-381).  v[p] = x[p] for the segmented sums and
// vals[p] * xv[cols[p]] for SpMV.
//
// Two launches.  The row pass gives every row one thread: empty rows get
// y[r] = 0 and every tile records the row holding its first element.  The
// nnz pass walks the nonzeros in tiles of NW * J * 64 * G elements with the
// same lane layout as the scans: the threads of a tile mark the row
// boundaries that fall in it in two LDS bitmaps (row heads, row ends) and
// record which row ends at each end position, so every lane reads its
// group's heads and ends with one LDS load; a segmented scan runs over the
// tile, and the carry of a row that began in an earlier tile arrives
// through the decoupled look-back.  Row sums are stored at the row ends.  The work per tile is proportional to its nonzeros, however the
// rows are distributed (lognormal rows of up to 20000 nonzeros are split
// across tiles like any other).
//
// Accumulation: fp16/fp32 values in double (fp16 segment sums are exact and
// correctly rounded to float; SpMV products are formed exactly), int8 in
// int32 (exact).  The reference rounds SpMV products to fp16
// (src/kernels/kernel_csr_gather.h:131-133); this build does not.
#include "tcs_common.hpp"

namespace tcs {

constexpr int kSegNW = 8, kSegJ = 2, kSegG = 8;
constexpr int64_t kSegTile = (int64_t)kSegNW * kSegJ * kWave * kSegG;  // 8192

template <class IdxT>
__device__ __forceinline__ int64_t row_end(const IdxT* __restrict__ ends, int64_t r, int64_t n_ends,
                                           int64_t nnz) {
  if (r < 0) return 0;
  const int64_t e = r < n_ends ? (int64_t)ends[r] : nnz;
  return e < nnz ? e : nnz;
}

// Row pass: zero empty rows, record each tile's first row.  tile_row[b] = R
// marks a tile past the last row.
template <class IdxT, class OutT>
__global__ __launch_bounds__(256) void seg_rows_kernel(const IdxT* __restrict__ ends, int64_t R,
                                                       int64_t n_ends, int64_t nnz,
                                                       OutT* __restrict__ y,
                                                       uint64_t* __restrict__ tile_row,
                                                       uint32_t num_tiles) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= R) return;
  const int64_t s = row_end(ends, r - 1, n_ends, nnz);
  const int64_t e = row_end(ends, r, n_ends, nnz);
  if (e <= s) {
    y[r] = OutT(0);
  } else {
    for (int64_t b = (s + kSegTile - 1) / kSegTile; b * kSegTile < e && b < num_tiles; ++b)
      tile_row[b] = (uint64_t)r;
  }
  if (r == R - 1) {
    const int64_t last = e > s ? e : s;
    for (int64_t b = (last + kSegTile - 1) / kSegTile; b < num_tiles; ++b) tile_row[b] = (uint64_t)R;
  }
}

// Value sources.
template <class C>
struct SrcPlain {
  using Acc = typename C::Acc;
  using In = typename C::In;
  const In* __restrict__ x;
  struct Regs { In v[kSegJ][kSegG]; };
  template <bool FULL>
  __device__ __forceinline__ void load(Regs& r, int j, int64_t gb, int64_t t1) const {
    load_items<In, kSegG, FULL>(x, gb, t1, r.v[j]);
  }
  __device__ __forceinline__ Acc value(const Regs& r, int j, int i) const { return C::conv(r.v[j][i]); }
  __device__ __forceinline__ Acc value_at(int64_t p) const { return C::conv(x[p]); }
  __device__ __forceinline__ void barrier(Regs& r) const {
#pragma unroll
    for (int j = 0; j < kSegJ; ++j) reg_barrier(r.v[j]);
  }
};

// SpMV products vals[p] * xv[cols[p]] formed exactly in double.  Columns
// outside [0, x_len) contribute 0.
template <class C>
struct SrcSpmv {
  using Acc = double;
  using In = typename C::In;
  const In* __restrict__ vals;
  const int32_t* __restrict__ cols;
  const In* __restrict__ xv;
  int64_t x_len;
  struct Regs { In v[kSegJ][kSegG]; In g[kSegJ][kSegG]; };
  // Values and column indices stream through once: nontemporal loads on
  // whole groups, so that they do not displace the gathered vector x from
  // the caches.
  template <class T>
  __device__ __forceinline__ static void load_nt(const T* __restrict__ p, int64_t base, T (&r)[kSegG]) {
    constexpr int kBytes = kSegG * (int)sizeof(T);
    static_assert(kBytes % 16 == 0, "whole 16-byte vectors");
    const u32x4* vp = reinterpret_cast<const u32x4*>(p + base);
    u32x4 tmp[kBytes / 16];
#pragma unroll
    for (int k = 0; k < kBytes / 16; ++k) tmp[k] = __builtin_nontemporal_load(vp + k);
    __builtin_memcpy(r, tmp, kBytes);
  }
  template <bool FULL>
  __device__ __forceinline__ void load(Regs& r, int j, int64_t gb, int64_t t1) const {
    int32_t c[kSegG];
    if constexpr (FULL) {
      load_nt<In>(vals, gb, r.v[j]);
      load_nt<int32_t>(cols, gb, c);
    } else {
      load_items<In, kSegG, false>(vals, gb, t1, r.v[j]);
      load_items<int32_t, kSegG, false>(cols, gb, t1, c);
    }
#pragma unroll
    for (int i = 0; i < kSegG; ++i)
      r.g[j][i] = ((uint32_t)c[i] < (uint64_t)x_len && (FULL || gb + i < t1)) ? xv[c[i]] : In(0);
  }
  __device__ __forceinline__ Acc value(const Regs& r, int j, int i) const {
    return (double)C::conv(r.v[j][i]) * (double)C::conv(r.g[j][i]);
  }
  __device__ __forceinline__ Acc value_at(int64_t p) const {
    const int32_t c = cols[p];
    if ((uint32_t)c >= (uint64_t)x_len) return 0.0;
    return (double)C::conv(vals[p]) * (double)C::conv(xv[c]);
  }
  __device__ __forceinline__ void barrier(Regs& r) const {
#pragma unroll
    for (int j = 0; j < kSegJ; ++j) {
      reg_barrier(r.v[j]);
      reg_barrier(r.g[j]);
    }
  }
};

template <class Src, class IdxT, class OutT>
__global__ __launch_bounds__(kSegNW * kWave) void seg_reduce_kernel(
    Src src, const IdxT* __restrict__ ends, int64_t R, int64_t n_ends, int64_t nnz,
    OutT* __restrict__ y, const uint64_t* __restrict__ tile_row, uint32_t num_tiles, void* ws) {
  using Acc = typename Src::Acc;
  constexpr int G = kSegG, J = kSegJ, NW = kSegNW;
  constexpr int64_t GSPAN = (int64_t)kWave * G, WSEG = GSPAN * J;
  __shared__ uint32_t s_head[kSegTile / 32];  // bit q: a row starts at t0 + q
  __shared__ uint32_t s_endb[kSegTile / 32];  // bit q: a row ends at t0 + q (inclusive)
  __shared__ uint32_t s_rowk[kSegTile];       // at an end position: r - (r_lo - 1) of that row
  __shared__ Acc s_wval[NW], s_wexcl[NW], s_carry;
  __shared__ int s_wflag[NW], s_wexclf[NW];
  __shared__ LaunchTag s_tag;

  // Every workgroup arrives, including those that return early below.
  if (threadIdx.x == 0) s_tag = launch_arrive(ws);
  const int lane = __lane_id();
  const int wid = threadIdx.x / kWave;
  const uint32_t tile = blockIdx.x;
  const int64_t t0 = (int64_t)tile * kSegTile;
  const int64_t end_all = row_end(ends, R - 1, n_ends, nnz);
  if (t0 >= end_all) return;  // no row reaches this tile (nor any later one)
  const int64_t t1 = t0 + kSegTile < end_all ? t0 + kSegTile : end_all;

  // Rows touching the tile: r_lo (holds t0) .. r_hi; their ends are
  // end(r_lo - 1 + k) for k in [0, cnt).
  int64_t r_lo = (int64_t)tile_row[tile];
  int64_t r_hi = tile + 1 < num_tiles ? (int64_t)tile_row[tile + 1] : R - 1;
  if (r_lo > R - 1) r_lo = R - 1;
  if (r_hi > R - 1) r_hi = R - 1;
  if (r_hi < r_lo) r_hi = r_lo;
  const int64_t cnt = r_hi - r_lo + 2;

  typename Src::Regs regs;
  const int64_t wbase = t0 + (int64_t)wid * WSEG + (int64_t)lane * G;
  const bool full = t0 + kSegTile <= t1;
  if (full) {
#pragma unroll
    for (int j = 0; j < J; ++j) src.template load<true>(regs, j, wbase + j * GSPAN, t1);
  } else {
#pragma unroll
    for (int j = 0; j < J; ++j) src.template load<false>(regs, j, wbase + j * GSPAN, t1);
  }

  // Mark the boundaries.  Among equal ends (empty rows) only the first, the
  // row that holds the element before it, marks them, so every position is
  // written by at most one thread.
  for (int k = threadIdx.x; k < kSegTile / 32; k += NW * kWave) {
    s_head[k] = 0;
    s_endb[k] = 0;
  }
  __syncthreads();
  for (int64_t k = threadIdx.x; k < cnt; k += NW * kWave) {
    const int64_t e = row_end(ends, r_lo - 1 + k, n_ends, nnz);
    const int64_t ep = k > 0 ? row_end(ends, r_lo - 2 + k, n_ends, nnz) : -1;
    if (k > 0 && ep == e) continue;  // an empty row: the same boundary as before
    if (e >= t0 && e < t1) {
      const int q = (int)(e - t0);
      atomicOr(&s_head[q >> 5], 1u << (q & 31));
    }
    if (e > t0 && e <= t1) {
      const int q = (int)(e - 1 - t0);
      atomicOr(&s_endb[q >> 5], 1u << (q & 31));
      s_rowk[q] = (uint32_t)k;
    }
  }
  __syncthreads();

  // Heads and ends of each lane group (G = 8 elements at an 8-aligned offset).
  uint32_t hmask[J], emask[J];
#pragma unroll
  for (int j = 0; j < J; ++j) {
    const int q = (int)(wbase + j * GSPAN - t0);
    hmask[j] = (s_head[q >> 5] >> (q & 31)) & 0xffu;
    emask[j] = (s_endb[q >> 5] >> (q & 31)) & 0xffu;
  }

  // Segmented reduction of the tile (as in the scans).
  Acc gex[J];
  int gexf[J];
  Acc wtot = Acc(0);
  int wtotf = 0;
#pragma unroll
  for (int j = 0; j < J; ++j) {
    Acc tv = Acc(0);
    int tf = 0;
#pragma unroll
    for (int i = 0; i < G; ++i) {
      const Acc v = src.value(regs, j, i);
      if (hmask[j] >> i & 1u) {
        tv = v;
        tf = 1;
      } else {
        tv = tv + v;
      }
    }
    int wf = tf;
    Acc wv = tv;
    wave_seg_incl_scan(wf, wv);
    const Acc ev = dpp_mov<kWaveShr1>(wv);
    const int ef = dpp_i32<kWaveShr1>(wf);
    gex[j] = ef ? ev : wtot + ev;
    gexf[j] = wtotf | ef;
    const Acc gt = readlane(wv, kWave - 1);
    const int gf = readlane(wf, kWave - 1);
    wtot = gf ? gt : wtot + gt;
    wtotf |= gf;
  }
  if (lane == 0) {
    s_wval[wid] = wtot;
    s_wflag[wid] = wtotf;
  }
  __syncthreads();

  if (wid == 0) {
    Acc a = lane < NW ? s_wval[lane] : Acc(0);
    int af = lane < NW ? s_wflag[lane] : 0;
    wave_seg_incl_scan(af, a);
    const Acc agg = readlane(a, NW - 1);
    const int aggf = readlane(af, NW - 1);
    const Acc xa = dpp_mov<kWaveShr1>(a);
    const int xf = dpp_i32<kWaveShr1>(af);
    if (lane < NW) {
      s_wexcl[lane] = xa;
      s_wexclf[lane] = xf;
    }
    const bool first_head = readlane((int)(hmask[0] & 1u), 0) != 0;
    // Fallback: {row head inside tile u, sum since its last head}.
    auto fb = [&](int64_t u) {
      const int64_t u0 = u * kSegTile;
      const int64_t u1 = u0 + kSegTile < end_all ? u0 + kSegTile : end_all;
      // Row holding u1 - 1: first r with end(r) > u1 - 1.
      int64_t lo = 0, hi = R - 1;
      while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (row_end(ends, mid, n_ends, nnz) <= u1 - 1) lo = mid + 1; else hi = mid;
      }
      const int64_t s = row_end(ends, lo - 1, n_ends, nnz);
      const int64_t from = s > u0 ? s : u0;
      Acc sum = Acc(0);
      for (int64_t p = from + lane; p < u1; p += kWave) sum = sum + src.value_at(p);
      return FlagVal<Acc>{s >= u0 ? 1 : 0, wave_sum(sum)};
    };
    uint64_t* desc = ws_desc(ws);
    const uint32_t tag = s_tag.tag;
    const int max_polls = s_tag.polls;
    Acc carry = Acc(0);
    if (tile == 0) {
      if (lane == 0 && num_tiles > 1) desc_publish<Acc>(desc, 0, agg, tag, kFlagIncl);
    } else if (aggf) {
      if (lane == 0) desc_publish<Acc>(desc, tile, agg, tag, kFlagIncl);
      if (!first_head) carry = lookback<Acc>(desc, tile, tag, max_polls, fb);
    } else {
      if (lane == 0) desc_publish<Acc>(desc, tile, agg, tag, kFlagAgg);
      carry = lookback<Acc>(desc, tile, tag, max_polls, fb);
      if (lane == 0) desc_publish<Acc>(desc, tile, carry + agg, tag, kFlagIncl);
    }
    if (lane == 0) s_carry = carry;
  }
  __syncthreads();

  src.barrier(regs);
  Acc cin = s_carry;
  {
    const Acc wx = s_wexcl[wid];
    cin = s_wexclf[wid] ? wx : cin + wx;
  }
#pragma unroll
  for (int j = 0; j < J; ++j) {
    if (!emask[j]) continue;
    Acc run = gexf[j] ? gex[j] : cin + gex[j];
    const int q0 = (int)(wbase + j * GSPAN - t0);
#pragma unroll
    for (int i = 0; i < G; ++i) {
      const Acc v = src.value(regs, j, i);
      if (hmask[j] >> i & 1u) {
        run = v;
      } else {
        run = run + v;
      }
      if (emask[j] >> i & 1u) {
        int64_t row = r_lo - 1 + (int64_t)s_rowk[q0 + i];
        if (row < 0) row = 0;
        if (row > R - 1) row = R - 1;
        y[row] = (OutT)run;
      }
    }
  }
}

struct SegSumPlan {
  uint32_t num_tiles;
  size_t ws_bytes;
  size_t tile_row_off;
};

inline SegSumPlan seg_plan(int64_t nnz) {
  SegSumPlan p;
  p.num_tiles = (uint32_t)((nnz + kSegTile - 1) / kSegTile);
  const size_t lb = lookback_ws_bytes(p.num_tiles);
  p.tile_row_off = lb;
  p.ws_bytes = lb + (((size_t)p.num_tiles * 8 + 255) & ~(size_t)255);
  return p;
}

template <class Src, class IdxT, class OutT>
static int launch_seg(const Src& src, const void* ends_v, int64_t R, int64_t n_ends, int64_t nnz,
                      void* y_v, void* ws, size_t ws_bytes, hipStream_t st) {
  const SegSumPlan plan = seg_plan(nnz);
  if (!ws || ws_bytes < plan.ws_bytes) return TCS_ERR_WORKSPACE;
  const auto* ends = static_cast<const IdxT*>(ends_v);
  auto* y = static_cast<OutT*>(y_v);
  auto* tile_row = reinterpret_cast<uint64_t*>(static_cast<char*>(ws) + plan.tile_row_off);
  const int64_t row_blocks = (R + 255) / 256;
  hipLaunchKernelGGL((seg_rows_kernel<IdxT, OutT>), dim3((uint32_t)row_blocks), dim3(256), 0, st,
                     ends, R, n_ends, nnz, y, tile_row, plan.num_tiles);
  TCS_CHECK_LAUNCH();
  if (plan.num_tiles == 0) return TCS_OK;
  hipLaunchKernelGGL((seg_reduce_kernel<Src, IdxT, OutT>), dim3(plan.num_tiles), dim3(kSegNW * kWave),
                     0, st, src, ends, R, n_ends, nnz, y, tile_row, plan.num_tiles, ws);
  TCS_CHECK_LAUNCH();
  return TCS_OK;
}

template <class Src, class OutT>
static int launch_seg_idx(const Src& src, const void* indptr, int idx_dtype, int64_t R,
                          int64_t n_ends, int64_t nnz, void* y, void* ws, size_t ws_bytes,
                          hipStream_t st) {
  if (idx_dtype == TCS_I32)
    return launch_seg<Src, int32_t, OutT>(src, indptr, R, n_ends, nnz, y, ws, ws_bytes, st);
  if (idx_dtype == TCS_U32)
    return launch_seg<Src, uint32_t, OutT>(src, indptr, R, n_ends, nnz, y, ws, ws_bytes, st);
  if (idx_dtype == TCS_I64)
    return launch_seg<Src, int64_t, OutT>(src, indptr, R, n_ends, nnz, y, ws, ws_bytes, st);
  return TCS_ERR_UNSUPPORTED_DTYPE;
}

}  // namespace tcs

extern "C" {

size_t tcs_seg_sum_workspace_bytes(int64_t nnz) {
  return tcs::seg_plan(nnz < 0 ? 0 : nnz).ws_bytes;
}

int tcs_seg_sum_indptr(const void* x, int x_dtype, const void* ends, int ends_dtype,
                       int64_t num_segments, int64_t num_ends, void* y, int64_t nnz,
                       void* workspace, size_t workspace_bytes, void* stream) {
  using namespace tcs;
  if (num_segments < 0 || nnz < 0 || num_ends < 0 || num_ends > num_segments)
    return TCS_ERR_INVALID_ARG;
  if (num_segments == 0) return TCS_OK;
  if (!y || !ends || (nnz > 0 && !x)) return TCS_ERR_INVALID_ARG;
  if ((nnz > 0 && !aligned16(x))) return TCS_ERR_INVALID_ARG;
  if (nnz >= ((int64_t)1 << 32) || num_segments >= ((int64_t)1 << 32) - 1) return TCS_ERR_INVALID_ARG;
  if (ends_dtype != TCS_I32 && ends_dtype != TCS_U32 && ends_dtype != TCS_I64)
    return TCS_ERR_UNSUPPORTED_DTYPE;
  hipStream_t st = static_cast<hipStream_t>(stream);
  switch (x_dtype) {
    case TCS_F16:
      return launch_seg_idx<SrcPlain<CodecF16>, float>(SrcPlain<CodecF16>{static_cast<const uint16_t*>(x)},
                                                       ends, ends_dtype, num_segments, num_ends, nnz, y,
                                                       workspace, workspace_bytes, st);
    case TCS_BF16:
      return launch_seg_idx<SrcPlain<CodecBF16>, float>(SrcPlain<CodecBF16>{static_cast<const uint16_t*>(x)},
                                                        ends, ends_dtype, num_segments, num_ends, nnz, y,
                                                        workspace, workspace_bytes, st);
    case TCS_F32:
      return launch_seg_idx<SrcPlain<CodecF32>, float>(SrcPlain<CodecF32>{static_cast<const float*>(x)},
                                                       ends, ends_dtype, num_segments, num_ends, nnz, y,
                                                       workspace, workspace_bytes, st);
    case TCS_I8:
      return launch_seg_idx<SrcPlain<CodecI8>, int>(SrcPlain<CodecI8>{static_cast<const int8_t*>(x)},
                                                    ends, ends_dtype, num_segments, num_ends, nnz, y,
                                                    workspace, workspace_bytes, st);
    default:
      return TCS_ERR_UNSUPPORTED_DTYPE;
  }
}

int tcs_spmv_csr(const void* vals, int vals_dtype, const void* indptr_ends, int ends_dtype,
                 const int32_t* cols, const void* x, int64_t x_len, float* y, int64_t num_rows,
                 int64_t nnz, void* workspace, size_t workspace_bytes, void* stream) {
  using namespace tcs;
  if (num_rows < 0 || nnz < 0 || x_len < 0) return TCS_ERR_INVALID_ARG;
  if (num_rows == 0) return TCS_OK;
  if (!y || !indptr_ends || (nnz > 0 && (!vals || !cols || !x))) return TCS_ERR_INVALID_ARG;
  if (nnz > 0 && (!aligned16(vals) || !aligned16(cols))) return TCS_ERR_INVALID_ARG;
  if (nnz >= ((int64_t)1 << 32) || num_rows >= ((int64_t)1 << 32) - 1) return TCS_ERR_INVALID_ARG;
  if (ends_dtype != TCS_I32 && ends_dtype != TCS_U32 && ends_dtype != TCS_I64)
    return TCS_ERR_UNSUPPORTED_DTYPE;
  hipStream_t st = static_cast<hipStream_t>(stream);
  switch (vals_dtype) {
    case TCS_F16: {
      SrcSpmv<CodecF16> src{static_cast<const uint16_t*>(vals), cols, static_cast<const uint16_t*>(x), x_len};
      return launch_seg_idx<SrcSpmv<CodecF16>, float>(src, indptr_ends, ends_dtype, num_rows, num_rows, nnz,
                                                      y, workspace, workspace_bytes, st);
    }
    case TCS_F32: {
      SrcSpmv<CodecF32> src{static_cast<const float*>(vals), cols, static_cast<const float*>(x), x_len};
      return launch_seg_idx<SrcSpmv<CodecF32>, float>(src, indptr_ends, ends_dtype, num_rows, num_rows, nnz,
                                                      y, workspace, workspace_bytes, st);
    }
    default:
      return TCS_ERR_UNSUPPORTED_DTYPE;
  }
}

}  // extern "C"
