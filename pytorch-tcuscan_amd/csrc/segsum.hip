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
//
// SpMV over a vector larger than about half an L2 splits the gather by
// column range across the XCDs.  MI355X dispatches workgroup b to XCD
// b mod 8 and every XCD has its own 4 MB L2, so with two workgroups per
// tile, the even one summing the products whose column lies in the lower
// half of x and the odd one the upper half, each L2 caches only its half
// of x (2^22 fp16 columns: random gathers 0.61 -> 0.41 ms).  The two
// halves are two independent segmented reductions, each with its own
// look-back chain, writing exact double partial sums per row; a combine
// pass adds the two partials of every row and rounds once to float, so the
// result is bit-identical to the single-chain sum.  Correctness never
// depends on where workgroups run, only the cache hit rate does.
#include <type_traits>

#include "tcs_common.hpp"

namespace tcs {

// Tiles of kSegNW waves x J groups x 64 lanes x G items; each value source
// picks J and G (16 bytes per group load).  kSegMinTile, the smallest tile
// any source uses, sizes the workspaces.
constexpr int kSegNW = 8;
constexpr int64_t kSegMinTile = 8192;
// Tiles reduced by one workgroup, all loaded before the first is reduced.
// Measured on MI355X (deferred carries, SpMV 2^22 x 2^22 with 6.7e7
// lognormal-row nonzeros / fp16 seg-sum of the same rows): 1 tile 0.529 /
// 0.100 ms, 2 tiles 0.548 / 0.122 ms, 4 tiles 0.761 / 0.124 ms; more tiles
// per workgroup cost occupancy (SpMV: 79 VGPRs at 1, 113 at 2) and
// serialise each workgroup's phases.
constexpr int kSegTPW = 1;

template <class IdxT>
__device__ __forceinline__ int64_t row_end(const IdxT* __restrict__ ends, int64_t r, int64_t n_ends,
                                           int64_t nnz) {
  if (r < 0) return 0;
  const int64_t e = r < n_ends ? (int64_t)ends[r] : nnz;
  return e < nnz ? e : nnz;
}

// Row pass: zero empty rows, record each tile's first row.  tile_row[b] = R
// marks a tile past the last row.
// With `part` set (the split SpMV) empty rows zero both partial sums
// part[r] and part[R + r] instead of y[r].
// Each thread takes kRowsPerThread consecutive rows, their ends loaded
// together (4M rows: 9.8 us, 10.8 with one row per thread).
constexpr int kRowsPerThread = 4;

template <int64_t TILE, class IdxT, class OutT>
__global__ __launch_bounds__(256) void seg_rows_kernel(const IdxT* __restrict__ ends, int64_t R,
                                                       int64_t n_ends, int64_t nnz,
                                                       OutT* __restrict__ y, double* __restrict__ part,
                                                       uint64_t* __restrict__ tile_row,
                                                       uint32_t num_tiles) {
  const int64_t r0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * kRowsPerThread;
  if (r0 >= R) return;
  int64_t e[kRowsPerThread + 1];  // e[k] = end of row r0 - 1 + k
#pragma unroll
  for (int k = 0; k <= kRowsPerThread; ++k) e[k] = r0 - 1 + k < R ? row_end(ends, r0 - 1 + k, n_ends, nnz) : 0;
#pragma unroll
  for (int k = 0; k < kRowsPerThread; ++k) {
    const int64_t r = r0 + k;
    if (r >= R) break;
    const int64_t s = e[k], t = e[k + 1];
    if (t <= s) {
      if (part) {
        part[r] = 0.0;
        part[R + r] = 0.0;
      } else {
        y[r] = OutT(0);
      }
    } else {
      for (int64_t b = (s + TILE - 1) / TILE; b * TILE < t && b < num_tiles; ++b)
        tile_row[b] = (uint64_t)r;
    }
    if (r == R - 1) {
      const int64_t last = t > s ? t : s;
      for (int64_t b = (last + TILE - 1) / TILE; b < num_tiles; ++b) tile_row[b] = (uint64_t)R;
    }
  }
}

// Value sources.  Plain values: 16-byte groups (16 int8 or 8 fp16 / bf16
// values; fp32 takes two loads per group of 8); 16K-element tiles for int8
// (one tile's chain of dependent loads then covers 16 KiB instead of 8),
// 8K for wider values.
template <class C>
struct SrcPlain {
  using Acc = typename C::Acc;
  using In = typename C::In;
  static constexpr int G = sizeof(In) == 1 ? 16 : 8;
  static constexpr int J = 2;
  static constexpr int64_t kTile = (int64_t)kSegNW * J * kWave * G;
  const In* __restrict__ x;
  struct Regs { Packed<In, G> v[J]; };
  template <bool FULL>
  __device__ __forceinline__ void load(Regs& r, int j, int64_t gb, int64_t t1) const {
    load_items<In, G, FULL>(x, gb, t1, r.v[j]);
  }
  template <bool FULL>
  __device__ __forceinline__ void gather(Regs&, int, int64_t, int64_t) const {}
  __device__ __forceinline__ Acc value(const Regs& r, int j, int i) const { return C::conv(r.v[j][i]); }
  __device__ __forceinline__ Acc value_at(int64_t p) const { return C::conv(x[p]); }
  __device__ __forceinline__ void barrier(Regs& r) const {
#pragma unroll
    for (int j = 0; j < J; ++j) reg_barrier(r.v[j]);
  }
  __device__ __forceinline__ void set_part(int, int) {}
  static constexpr bool kSplittable = false;
};

// SpMV products vals[p] * xv[cols[p]], formed exactly: an fp16 product has
// at most 22 significant bits, so float holds it exactly; fp32 products are
// formed in double.  Only columns in [c_lo, c_lo + c_n) contribute (all of
// [0, x_len) unless set_part narrows it to one part of a split), others
// and columns outside [0, x_len) contribute exactly 0.
template <class C>
struct SrcSpmv {
  using Acc = double;
  using In = typename C::In;
  const In* __restrict__ vals;
  const int32_t* __restrict__ cols;
  const In* __restrict__ xv;
  int64_t x_len;
  uint32_t c_lo = 0, c_n = 0;
  static constexpr bool kSplittable = true;
  // 8K-element tiles (measured: 4K tiles 0.639 vs 0.563 ms on the 2^22
  // lognormal matrix).
  static constexpr int G = 8, J = 2;
  static constexpr int64_t kTile = (int64_t)kSegNW * J * kWave * G;
  struct Regs {
    Packed<In, G> v[J];
    Packed<In, G> g[J];
    uint32_t c[J][G];  // column indices, from load() until gather()
    uint32_t ok[J];        // bit i: item i of group j contributes
  };
  // Part `part` of `parts` equal column ranges (the last one shorter).
  __device__ __forceinline__ void set_part(int part, int parts) {
    // Columns are int32: nothing at or past 2^31 is ever gathered.
    const uint64_t len = x_len < ((int64_t)1 << 31) ? (uint64_t)x_len : ((uint64_t)1 << 31);
    const uint64_t w = (len + (uint64_t)parts - 1) / (uint64_t)parts;
    const uint64_t lo = w * (uint64_t)part < len ? w * (uint64_t)part : len;
    const uint64_t hi = lo + w < len ? lo + w : len;
    c_lo = (uint32_t)lo;
    c_n = (uint32_t)(hi - lo);
  }
  __device__ __forceinline__ bool in_range(int32_t c) const { return (uint32_t)c - c_lo < c_n; }
  // Values and column indices stream through once: nontemporal loads on
  // whole groups, so that they do not displace the gathered vector x from
  // the caches.
  template <class T, int W>
  __device__ __forceinline__ static void load_nt(const T* __restrict__ p, int64_t base, uint32_t (&w)[W]) {
    static_assert(W * 4 == G * (int)sizeof(T) && W % 4 == 0, "whole 16-byte vectors");
    const u32x4* vp = reinterpret_cast<const u32x4*>(p + base);
#pragma unroll
    for (int k = 0; k < W / 4; ++k) {
      const u32x4 t = __builtin_nontemporal_load(vp + k);
      w[4 * k] = t[0];
      w[4 * k + 1] = t[1];
      w[4 * k + 2] = t[2];
      w[4 * k + 3] = t[3];
    }
  }
  // The streamed loads of group j; gather() then fetches x for them (the
  // caller issues other loads in between, so that the column indices'
  // latency is not waited out twice).
  template <bool FULL>
  __device__ __forceinline__ void load(Regs& r, int j, int64_t gb, int64_t t1) const {
    if constexpr (FULL) {
      load_nt<In>(vals, gb, r.v[j].w);
      load_nt<int32_t>(cols, gb, r.c[j]);
    } else {
      load_items<In, G, false>(vals, gb, t1, r.v[j]);
      int32_t c[G];
      load_items<int32_t, G, false>(cols, gb, t1, c);
      __builtin_memcpy(r.c[j], c, sizeof(c));
    }
  }
  template <bool FULL>
  __device__ __forceinline__ void gather(Regs& r, int j, int64_t gb, int64_t t1) const {
    In g[G];
    uint32_t ok = 0;
#pragma unroll
    for (int i = 0; i < G; ++i) {
      const bool k = in_range((int32_t)r.c[j][i]) && (FULL || gb + i < t1);
      // A 32-bit byte offset from a uniform base (x_len < 2^31): one VGPR
      // per address instead of a 64-bit pointer.
      const uint32_t off = r.c[j][i] * (uint32_t)sizeof(In);
      g[i] = k ? *reinterpret_cast<const In*>(reinterpret_cast<const char*>(xv) + off) : In(0);
      ok |= (k ? 1u : 0u) << i;
    }
    __builtin_memcpy(r.g[j].w, g, sizeof(g));
    r.ok[j] = ok;
  }
  __device__ __forceinline__ static double prod(In a, In b) {
    if constexpr (sizeof(In) == 2) {
      return (double)((float)C::conv(a) * (float)C::conv(b));
    } else {
      return (double)C::conv(a) * (double)C::conv(b);
    }
  }
  __device__ __forceinline__ Acc value(const Regs& r, int j, int i) const {
    return (r.ok[j] >> i & 1u) ? prod(r.v[j][i], r.g[j][i]) : 0.0;
  }
  __device__ __forceinline__ Acc value_at(int64_t p) const {
    const int32_t c = cols[p];
    if (!in_range(c)) return 0.0;
    return prod(vals[p], xv[c]);
  }
  __device__ __forceinline__ void barrier(Regs& r) const {
#pragma unroll
    for (int j = 0; j < J; ++j) {
      reg_barrier(r.v[j]);
      reg_barrier(r.g[j]);
    }
  }
};

// One tile's loaded state, per thread: the values (and for SpMV column
// indices and gathered x), the tile's span and rows, and the row ends this
// thread marks.
template <class Src>
struct SegTile {
  typename Src::Regs regs;
  int64_t t0, t1, r_lo, r_hi, cnt;
  bool full, k16;
  uint32_t e_early[2], ep_early[2];
};

// SPLIT workgroups per tile group: workgroup b handles part b % SPLIT of the
// values (see the split SpMV above) with that part's own descriptors and
// writes its row sums as doubles to part[p * R + row]; SPLIT = 1 writes y.
// Each workgroup reduces TPW consecutive tiles (b / SPLIT) * TPW + m, and
// every load of all of them (values, column indices, row ends, gathers of
// x) is issued before the first is reduced.
//
// No tile waits for its look-back: a tile's rows that start inside it are
// complete without the carry from earlier tiles, so they are written as soon
// as the tile is reduced, and the tile publishes its aggregate at once.  At
// most one row end per tile needs the carry (the first end, when no row
// starts before it: the row came in from an earlier tile); its sum without
// the carry waits in LDS.  After the last tile one walk finds the carry into
// the first tile, the later tiles' carries follow locally from the earlier
// tiles' aggregates, and the open ends and missing inclusive descriptors are
// written then.
template <class Src, class IdxT, class OutT, int SPLIT, int TPW>
__global__ __launch_bounds__(kSegNW * kWave) void seg_reduce_kernel(
    Src src, const IdxT* __restrict__ ends, int64_t R, int64_t n_ends, int64_t nnz,
    OutT* __restrict__ y, double* __restrict__ part_out, const uint64_t* __restrict__ tile_row,
    uint32_t num_tiles, void* ws) {
  using Acc = typename Src::Acc;
  constexpr int G = Src::G, J = Src::J, NW = kSegNW;
  constexpr int64_t TILE = Src::kTile;
  constexpr int64_t GSPAN = (int64_t)kWave * G, WSEG = GSPAN * J;
  static_assert(32 % G == 0 || G == 16, "a group's bits lie in one 32-bit word");
  __shared__ uint32_t s_head[TILE / 32];  // bit q: a row starts at t0 + q
  __shared__ uint32_t s_endb[TILE / 32];  // bit q: a row ends at t0 + q (inclusive)
  // At an end position q: k = r - (r_lo - 1) of the row ending there.
  // Tiles above 8K keep k in 16 bits (half the LDS) when the tile's rows
  // number at most 65536 (k16); otherwise each end's row is found by a
  // binary search over the row ends.
  constexpr bool kRowK16 = TILE > 8192;
  using RowK = typename std::conditional<kRowK16, uint16_t, uint32_t>::type;
  __shared__ RowK s_rowk[TILE];
  __shared__ Acc s_wval[NW];
  __shared__ int s_wflag[NW];
  __shared__ Acc s_open_val[TPW];
  __shared__ int64_t s_open_row[TPW];
  __shared__ int s_open_has[TPW];
  __shared__ LaunchTag s_tag;

  // Every workgroup arrives, including those that return early below.
  if (threadIdx.x == 0) s_tag = launch_arrive(ws);
  const int lane = __lane_id();
  const int wid = threadIdx.x / kWave;
  const uint32_t tile_base = blockIdx.x / SPLIT * (uint32_t)TPW;
  const int part = SPLIT > 1 ? (int)(blockIdx.x % SPLIT) : 0;
  src.set_part(part, SPLIT);
  const int64_t end_all = row_end(ends, R - 1, n_ends, nnz);
  const int64_t lane_off = (int64_t)wid * WSEG + (int64_t)lane * G;
  constexpr int kEarly = 2;

  // Loads that do not depend on one another go out together: the last row
  // end, each tile's first rows and its values (bounded by nnz; values past
  // the last row end belong to no row and are never stored).
  SegTile<Src> st[TPW];
#pragma unroll
  for (int m = 0; m < TPW; ++m) {
    const uint32_t tile = tile_base + m;
    SegTile<Src>& a = st[m];
    a.t0 = (int64_t)tile * TILE;
    a.t1 = a.t0 + TILE < nnz ? a.t0 + TILE : nnz;
    a.full = a.t0 + TILE <= nnz;
    if (tile >= num_tiles) continue;
    a.r_lo = (int64_t)tile_row[tile];
    a.r_hi = tile + 1 < num_tiles ? (int64_t)tile_row[tile + 1] : R - 1;
    if (a.full) {
#pragma unroll
      for (int j = 0; j < J; ++j) src.template load<true>(a.regs, j, a.t0 + lane_off + j * GSPAN, a.t1);
    } else {
#pragma unroll
      for (int j = 0; j < J; ++j) src.template load<false>(a.regs, j, a.t0 + lane_off + j * GSPAN, a.t1);
    }
  }
  if (st[0].t0 >= end_all) return;  // no row reaches these tiles (nor any later one)

  // Rows touching a tile: r_lo (holds t0) .. r_hi; their ends are
  // end(r_lo - 1 + k) for k in [0, cnt).  Each thread's first kEarly ends
  // (k = thread, thread + 512) are loaded before any gather is issued.
  // Row ends are below nnz < 2^32 (checked on the host): 32-bit registers.
#pragma unroll
  for (int m = 0; m < TPW; ++m) {
    SegTile<Src>& a = st[m];
    if (tile_base + m >= num_tiles) continue;
    if (a.r_lo > R - 1) a.r_lo = R - 1;
    if (a.r_hi > R - 1) a.r_hi = R - 1;
    if (a.r_hi < a.r_lo) a.r_hi = a.r_lo;
    a.cnt = a.r_hi - a.r_lo + 2;
    a.k16 = !kRowK16 || a.cnt <= 65536;
#pragma unroll
    for (int q = 0; q < kEarly; ++q) {
      const int64_t k = threadIdx.x + (int64_t)q * NW * kWave;
      a.e_early[q] = k < a.cnt ? (uint32_t)row_end(ends, a.r_lo - 1 + k, n_ends, nnz) : 0u;
      a.ep_early[q] = (k > 0 && k < a.cnt) ? (uint32_t)row_end(ends, a.r_lo - 2 + k, n_ends, nnz) : 0u;
    }
  }
#pragma unroll
  for (int m = 0; m < TPW; ++m) {
    SegTile<Src>& a = st[m];
    if (tile_base + m >= num_tiles) continue;
    if (a.full) {
#pragma unroll
      for (int j = 0; j < J; ++j) src.template gather<true>(a.regs, j, a.t0 + lane_off + j * GSPAN, a.t1);
    } else {
#pragma unroll
      for (int j = 0; j < J; ++j) src.template gather<false>(a.regs, j, a.t0 + lane_off + j * GSPAN, a.t1);
    }
  }

  uint64_t* desc = ws_desc(ws) + (size_t)part * num_tiles * (kDescBytes / 8);
  Acc agg[TPW];
  int aggf[TPW];
  int nv = 0;  // tiles of this group that hold row ends (block-uniform)
#pragma unroll
  for (int m = 0; m < TPW; ++m) {
    SegTile<Src>& a = st[m];
    const uint32_t tile = tile_base + m;
    agg[m] = Acc(0);
    aggf[m] = 0;
    // Block-uniform: later tiles of the group may lie past the last tile or
    // the last row.
    if (tile >= num_tiles || a.t0 >= end_all) break;
    nv = m + 1;
    const int64_t t0 = a.t0, t1 = a.t1;
    const int64_t wbase = t0 + lane_off;

    // The previous tile's outputs have read s_rowk and s_wval.
    if (m > 0) __syncthreads();
    // Mark the boundaries.  Among equal ends (empty rows) only the first,
    // the row that holds the element before it, marks them, so every
    // position is written by at most one thread.
    for (int k = threadIdx.x; k < TILE / 32; k += NW * kWave) {
      s_head[k] = 0;
      s_endb[k] = 0;
    }
    if (threadIdx.x == 0) s_open_has[m] = 0;
    __syncthreads();
    auto mark = [&](int64_t k, int64_t e, int64_t ep) {
      if (k > 0 && ep == e) return;  // an empty row: the same boundary as before
      if (e >= t0 && e < t1) {
        const int q = (int)(e - t0);
        atomicOr(&s_head[q >> 5], 1u << (q & 31));
      }
      if (e > t0 && e <= t1) {
        const int q = (int)(e - 1 - t0);
        atomicOr(&s_endb[q >> 5], 1u << (q & 31));
        if (a.k16) s_rowk[q] = (RowK)k;
      }
    };
#pragma unroll
    for (int q = 0; q < kEarly; ++q) {
      const int64_t k = threadIdx.x + (int64_t)q * NW * kWave;
      if (k < a.cnt) mark(k, (int64_t)a.e_early[q], k > 0 ? (int64_t)a.ep_early[q] : -1);
    }
    for (int64_t k = threadIdx.x + (int64_t)kEarly * NW * kWave; k < a.cnt; k += NW * kWave)
      mark(k, row_end(ends, a.r_lo - 1 + k, n_ends, nnz), row_end(ends, a.r_lo - 2 + k, n_ends, nnz));
    __syncthreads();

    // Heads and ends of each lane group (G elements at a G-aligned offset).
    uint32_t hmask[J], emask[J];
#pragma unroll
    for (int j = 0; j < J; ++j) {
      const int q = (int)(wbase + j * GSPAN - t0);
      constexpr uint32_t kGroupBits = G == 32 ? ~0u : ((1u << G) - 1u);
      hmask[j] = (s_head[q >> 5] >> (q & 31)) & kGroupBits;
      emask[j] = (s_endb[q >> 5] >> (q & 31)) & kGroupBits;
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
        const Acc v = src.value(a.regs, j, i);
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

    // Every wave combines the wave totals before it (its exclusive prefix
    // within the tile) and all of them (the tile's aggregate).
    Acc wx = Acc(0), ta = Acc(0);
    int wxf = 0, taf = 0;
#pragma unroll
    for (int k = 0; k < NW; ++k) {
      const Acc v = s_wval[k];
      const int f = s_wflag[k];
      if (k == wid) {
        wx = ta;
        wxf = taf;
      }
      ta = f ? v : ta + v;
      taf |= f;
    }
    agg[m] = ta;
    aggf[m] = taf;
    if (threadIdx.x == 0) {
      if (tile == 0) {
        if (num_tiles > 1) desc_publish<Acc>(desc, 0, ta, s_tag.tag, kFlagIncl);
      } else {
        desc_publish<Acc>(desc, tile, ta, s_tag.tag, taf ? kFlagIncl : kFlagAgg);
      }
    }

    // Row sums.  A run with a row start before it in the tile is complete;
    // the one run without (at most one end per tile) waits for the carry.
    src.barrier(a.regs);
#pragma unroll
    for (int j = 0; j < J; ++j) {
      if (!emask[j]) continue;
      Acc run = gexf[j] ? gex[j] : wx + gex[j];
      int rf = gexf[j] | wxf;
      const int q0 = (int)(wbase + j * GSPAN - t0);
#pragma unroll
      for (int i = 0; i < G; ++i) {
        const Acc v = src.value(a.regs, j, i);
        if (hmask[j] >> i & 1u) {
          run = v;
          rf = 1;
        } else {
          run = run + v;
        }
        if (emask[j] >> i & 1u) {
          int64_t row;
          if (!kRowK16 || a.k16) {
            row = a.r_lo - 1 + (int64_t)s_rowk[q0 + i];
          } else {
            // The row holding position t0 + q0 + i: the first whose end is past it.
            int64_t lo = a.r_lo - 1 < 0 ? 0 : a.r_lo - 1, hi = a.r_hi;
            while (lo < hi) {
              const int64_t mid = (lo + hi) >> 1;
              if (row_end(ends, mid, n_ends, nnz) <= t0 + q0 + i) lo = mid + 1; else hi = mid;
            }
            row = lo;
          }
          if (row < 0) row = 0;
          if (row > R - 1) row = R - 1;
          if (rf || tile == 0) {
            if constexpr (SPLIT > 1) {
              part_out[(int64_t)part * R + row] = (double)run;
            } else {
              y[row] = (OutT)run;
            }
          } else {
            s_open_val[m] = run;
            s_open_row[m] = row;
            s_open_has[m] = 1;
          }
        }
      }
    }
  }

  // The carries: one walk for the first tile when any tile needs it, then
  // tile by tile from the aggregates.
  __syncthreads();
  if (wid != 0) return;
  bool need = false, closed = false;
#pragma unroll
  for (int m = 0; m < TPW; ++m) {
    if (m < nv && !closed && (s_open_has[m] || !aggf[m])) need = true;
    if (m < nv) closed = closed || aggf[m];
  }
  Acc carry = Acc(0);
  if (need && tile_base != 0) {
    // Fallback: {row head inside tile u, sum since its last head}.
    auto fb = [&](int64_t u) {
      const int64_t u0 = u * TILE;
      const int64_t u1 = u0 + TILE < end_all ? u0 + TILE : end_all;
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
    carry = lookback<Acc>(desc, tile_base, s_tag.tag, s_tag.polls, fb);
  }
#pragma unroll
  for (int m = 0; m < TPW; ++m) {
    if (m >= nv) break;
    const uint32_t tile = tile_base + m;
    if (lane == 0) {
      if (tile != 0 && !aggf[m]) desc_publish<Acc>(desc, tile, carry + agg[m], s_tag.tag, kFlagIncl);
      if (s_open_has[m]) {
        const Acc v = carry + s_open_val[m];
        if constexpr (SPLIT > 1) {
          part_out[(int64_t)part * R + s_open_row[m]] = (double)v;
        } else {
          y[s_open_row[m]] = (OutT)v;
        }
      }
    }
    carry = aggf[m] ? agg[m] : carry + agg[m];
  }
}

// y[r] = float(part[r] + part[R + r]) for every row, and the partials
// zeroed again: the workspace carries no doubles from one call to the next
// (a later call may read these bytes as look-back descriptors).
__global__ __launch_bounds__(256) void spmv_combine_kernel(double* __restrict__ part, int64_t R,
                                                           float* __restrict__ y) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= R) return;
  const double a = part[r], b = part[R + r];
  y[r] = (float)(a + b);
  part[r] = 0.0;
  part[R + r] = 0.0;
}

// Workspace of one call: the look-back region (header, `parts` descriptor
// chains of num_tiles tiles), the first row of every tile, and with
// parts > 1 the two partial sums of every row.
struct SegSumPlan {
  uint32_t num_tiles;
  int parts;
  size_t ws_bytes;
  size_t tile_row_off;
  size_t part_off;
};

// tile: the value source's tile (kSegMinTile, the smallest, sizes the
// workspaces a caller allocates without knowing the source).
inline SegSumPlan seg_plan(int64_t nnz, int parts = 1, int64_t rows = 0, int64_t tile = kSegMinTile) {
  SegSumPlan p;
  p.num_tiles = (uint32_t)((nnz + tile - 1) / tile);
  p.parts = parts;
  const size_t lb = lookback_ws_bytes((uint64_t)p.num_tiles * (uint64_t)parts);
  p.tile_row_off = lb;
  p.part_off = lb + (((size_t)p.num_tiles * 8 + 255) & ~(size_t)255);
  p.ws_bytes = p.part_off + (parts > 1 ? (((size_t)rows * 16 + 255) & ~(size_t)255) : 0);
  return p;
}

// Vectors of more than this many bytes take the split SpMV: each L2 (4 MB)
// then caches half of x.
constexpr size_t kSpmvSplitMinXBytes = (size_t)2 << 20;

inline int spmv_parts(int64_t x_len, int vals_dtype) {
  const size_t eb = vals_dtype == TCS_F32 ? 4 : 2;
  return (size_t)x_len * eb > kSpmvSplitMinXBytes ? 2 : 1;
}

template <class Src, class IdxT, class OutT>
static int launch_seg(const Src& src, const void* ends_v, int64_t R, int64_t n_ends, int64_t nnz,
                      void* y_v, void* ws, size_t ws_bytes, hipStream_t st, int parts = 1) {
  static_assert(Src::kTile % kSegMinTile == 0, "tiles are multiples of the smallest tile");
  SegSumPlan plan = seg_plan(nnz, parts, R, Src::kTile);
  // One chain when there is nothing to gather or no room for the split.
  if (parts > 1 && (nnz == 0 || !ws || ws_bytes < plan.ws_bytes)) plan = seg_plan(nnz, 1, 0, Src::kTile);
  if (!ws || ws_bytes < plan.ws_bytes) return TCS_ERR_WORKSPACE;
  const auto* ends = static_cast<const IdxT*>(ends_v);
  auto* y = static_cast<OutT*>(y_v);
  auto* tile_row = reinterpret_cast<uint64_t*>(static_cast<char*>(ws) + plan.tile_row_off);
  double* part = plan.parts > 1 ? reinterpret_cast<double*>(static_cast<char*>(ws) + plan.part_off) : nullptr;
  const int64_t row_blocks = (R + 255) / 256;
  const int64_t rows_blocks = (R + 256 * kRowsPerThread - 1) / (256 * kRowsPerThread);
  hipLaunchKernelGGL((seg_rows_kernel<Src::kTile, IdxT, OutT>), dim3((uint32_t)rows_blocks), dim3(256), 0, st,
                     ends, R, n_ends, nnz, y, part, tile_row, plan.num_tiles);
  TCS_CHECK_LAUNCH();
  if (plan.num_tiles == 0) return TCS_OK;
  const uint32_t groups = (plan.num_tiles + kSegTPW - 1) / kSegTPW;
  if (plan.parts > 1) {
    if constexpr (Src::kSplittable && std::is_same<OutT, float>::value) {
      hipLaunchKernelGGL((seg_reduce_kernel<Src, IdxT, OutT, 2, kSegTPW>), dim3(groups * 2u),
                         dim3(kSegNW * kWave), 0, st, src, ends, R, n_ends, nnz, y, part, tile_row,
                         plan.num_tiles, ws);
      TCS_CHECK_LAUNCH();
      hipLaunchKernelGGL(spmv_combine_kernel, dim3((uint32_t)row_blocks), dim3(256), 0, st, part, R, y);
      TCS_CHECK_LAUNCH();
      return TCS_OK;
    } else {
      return TCS_ERR_UNSUPPORTED_DTYPE;
    }
  }
  hipLaunchKernelGGL((seg_reduce_kernel<Src, IdxT, OutT, 1, kSegTPW>), dim3(groups), dim3(kSegNW * kWave),
                     0, st, src, ends, R, n_ends, nnz, y, nullptr, tile_row, plan.num_tiles, ws);
  TCS_CHECK_LAUNCH();
  return TCS_OK;
}

template <class Src, class OutT>
static int launch_seg_idx(const Src& src, const void* indptr, int idx_dtype, int64_t R,
                          int64_t n_ends, int64_t nnz, void* y, void* ws, size_t ws_bytes,
                          hipStream_t st, int parts = 1) {
  if (idx_dtype == TCS_I32)
    return launch_seg<Src, int32_t, OutT>(src, indptr, R, n_ends, nnz, y, ws, ws_bytes, st, parts);
  if (idx_dtype == TCS_U32)
    return launch_seg<Src, uint32_t, OutT>(src, indptr, R, n_ends, nnz, y, ws, ws_bytes, st, parts);
  if (idx_dtype == TCS_I64)
    return launch_seg<Src, int64_t, OutT>(src, indptr, R, n_ends, nnz, y, ws, ws_bytes, st, parts);
  return TCS_ERR_UNSUPPORTED_DTYPE;
}

}  // namespace tcs

extern "C" {

size_t tcs_seg_sum_workspace_bytes(int64_t nnz) {
  return tcs::seg_plan(nnz < 0 ? 0 : nnz).ws_bytes;
}

size_t tcs_spmv_workspace_bytes(int64_t nnz, int64_t num_rows, int64_t x_len, int vals_dtype) {
  if (nnz < 0) nnz = 0;
  if (num_rows < 0) num_rows = 0;
  if (x_len < 0) x_len = 0;
  return tcs::seg_plan(nnz, tcs::spmv_parts(x_len, vals_dtype), num_rows).ws_bytes;
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
  const int parts = spmv_parts(x_len, vals_dtype);
  switch (vals_dtype) {
    case TCS_F16: {
      SrcSpmv<CodecF16> src{static_cast<const uint16_t*>(vals), cols, static_cast<const uint16_t*>(x), x_len};
      return launch_seg_idx<SrcSpmv<CodecF16>, float>(src, indptr_ends, ends_dtype, num_rows, num_rows, nnz,
                                                      y, workspace, workspace_bytes, st, parts);
    }
    case TCS_F32: {
      SrcSpmv<CodecF32> src{static_cast<const float*>(vals), cols, static_cast<const float*>(x), x_len};
      return launch_seg_idx<SrcSpmv<CodecF32>, float>(src, indptr_ends, ends_dtype, num_rows, num_rows, nnz,
                                                      y, workspace, workspace_bytes, st, parts);
    }
    default:
      return TCS_ERR_UNSUPPORTED_DTYPE;
  }
}

}  // extern "C"
