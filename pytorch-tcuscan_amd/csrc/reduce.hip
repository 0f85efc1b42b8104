// Reduction and down-sweep building blocks of the reference's multi-core
// scan, with the reference's block geometry:
//
//  reduce_tiles(x, tile_len, num_blocks): sum of block b =
//    [b*BL, min((b+1)*BL, n)), BL = tile_len * ceil(ceil(n/tile_len) /
//    num_blocks) (KernelReduceTiles, src/kernels/kernel_reduce_tiles.h:
//    79-135, work split by scalar::GetWorkDistribution,
//    src/kernels/tcuscan_utils.h:1389-1408).
//  complete_blocks(x, sums, tile_len): x[i] + sum(sums[:b]) for block b of
//    n / len(sums) elements (src/kernels/kernel_complete_blocks.h:109-179,
//    src/complete_blocks.cpp:36-40).
//  complete_rows(x, sums, w, h): blocks as reduce_tiles with tiles of
//    w * min(h, ceil(8192 / w)) elements; inside block b the rows of w
//    elements get a running carry that starts at sum(sums[:b]) and becomes
//    each row's last output (src/kernels/kernel_complete_rows.h:99-161).
//
// Floating sums accumulate in double and round once; integer sums are
// exact (int32 results wrap).  reduce_tiles splits every block over many
// workgroups and adds the partials in a fixed order, so results are
// deterministic.  complete_rows evaluates the row carries as a segmented
// scan of the row-end values (one launch of the scan kernel) rather than
// the reference's serial carry.
//
// The workspace scratch of these ops holds doubles and signed integers,
// which a later look-back launch on the same workspace could read as
// descriptors; each op zeroes its scratch before it returns (the last
// reader of the partials clears them; complete_blocks / complete_rows end
// with a fill).
#include "tcs_common.hpp"

namespace tcs {

constexpr int kRedBlock = 256;
constexpr int64_t kRedChunk = 32768;

template <class C>
__global__ __launch_bounds__(kRedBlock) void reduce_partials_kernel(
    const typename C::In* __restrict__ x, int64_t n, int64_t block_len, int64_t cpb,
    typename C::Acc* __restrict__ partial) {
  using Acc = typename C::Acc;
  const int64_t b = blockIdx.x / cpb, c = blockIdx.x % cpb;
  const int64_t bs = b * block_len;
  int64_t be = bs + block_len;
  if (be > n) be = n;
  int64_t s = bs + c * kRedChunk, e = s + kRedChunk;
  if (e > be) e = be;
  Acc acc = Acc(0);
  for (int64_t p = s + threadIdx.x; p < e; p += kRedBlock) acc = acc + C::conv(x[p]);
  __shared__ Acc sh[kRedBlock / kWave];
  const Acc w = wave_sum(acc);
  if (__lane_id() == 0) sh[threadIdx.x / kWave] = w;
  __syncthreads();
  if (threadIdx.x == 0) {
    Acc t = Acc(0);
    for (int k = 0; k < kRedBlock / kWave; ++k) t = t + sh[k];
    partial[blockIdx.x] = t;
  }
}

template <class Acc, class Out>
__global__ void reduce_final_kernel(Acc* __restrict__ partial, int64_t nb, int64_t cpb,
                                    Out* __restrict__ out) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  Acc t = Acc(0);
  for (int64_t c = 0; c < cpb; ++c) {
    t = t + partial[b * cpb + c];
    partial[b * cpb + c] = Acc(0);
  }
  out[b] = (Out)t;
}

// Exclusive prefix of `sums` (one workgroup walks the array in chunks).
template <class T, class Acc>
__global__ __launch_bounds__(1024) void excl_prefix_kernel(const T* __restrict__ sums, int64_t nb,
                                                           Acc* __restrict__ prefix) {
  __shared__ Acc s_w[16];
  __shared__ Acc s_carry;
  const int lane = __lane_id(), wid = threadIdx.x / kWave;
  if (threadIdx.x == 0) s_carry = Acc(0);
  __syncthreads();
  for (int64_t base = 0; base < nb; base += 1024) {
    const int64_t i = base + threadIdx.x;
    const Acc v = i < nb ? (Acc)sums[i] : Acc(0);
    const Acc incl = wave_incl_scan(v);
    if (lane == kWave - 1) s_w[wid] = incl;
    __syncthreads();
    Acc before = s_carry;
    for (int k = 0; k < wid; ++k) before = before + s_w[k];
    if (i < nb) prefix[i] = before + incl - v;
    __syncthreads();
    if (threadIdx.x == 0) {
      Acc t = s_carry;
      for (int k = 0; k < 16; ++k) t = t + s_w[k];
      s_carry = t;
    }
    __syncthreads();
  }
}

// complete_blocks: out[i] = x[i] + prefix[i / block_len].
template <class T, class Acc>
__global__ __launch_bounds__(256) void add_block_prefix_kernel(const T* __restrict__ x,
                                                               const Acc* __restrict__ prefix,
                                                               T* __restrict__ y, int64_t n,
                                                               int64_t block_len) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  y[i] = (T)((Acc)x[i] + prefix[i / block_len]);
}

// complete_rows helpers: gather the last element of every row, then add
// each row's carry prefix[b] + (first row of the block ? 0 : I[k-1]) where
// I is the per-block inclusive scan of the row ends.
template <class T>
__global__ __launch_bounds__(256) void gather_row_ends_kernel(const T* __restrict__ x, int64_t n,
                                                              int64_t w, T* __restrict__ e,
                                                              int64_t rows) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (k >= rows) return;
  int64_t p = k * w + w - 1;
  if (p >= n) p = n - 1;
  e[k] = x[p];
}

template <class T, class Acc, class ScanT>
__global__ __launch_bounds__(256) void add_row_carry_kernel(const T* __restrict__ x,
                                                            const Acc* __restrict__ prefix,
                                                            const ScanT* __restrict__ incl,
                                                            T* __restrict__ y, int64_t n, int64_t w,
                                                            int64_t block_len) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int64_t b = i / block_len;
  const int64_t k = i / w;
  Acc c = prefix[b];
  if (k * w != b * block_len) c = c + (Acc)incl[k - 1];
  y[i] = (T)((Acc)x[i] + c);
}

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

}  // namespace tcs

extern "C" {

// Scratch of every op starts after the reserved header (kHeaderBytes): the
// first bytes of a workspace hold the persistent scan's ticket word, which
// must stay zero between calls.
size_t tcs_reduce_tiles_workspace_bytes(int64_t n, int64_t tile_len, int64_t num_blocks) {
  if (n <= 0 || tile_len <= 0 || num_blocks <= 0) return tcs::kHeaderBytes + 256;
  const int64_t bl = tile_len * tcs::ceil_div(tcs::ceil_div(n, tile_len), num_blocks);
  const int64_t cpb = tcs::ceil_div(bl, tcs::kRedChunk);
  return tcs::kHeaderBytes + (size_t)(num_blocks * cpb) * 8;
}

int tcs_reduce_tiles(const void* x, int x_dtype, void* out, int64_t n, int64_t tile_len,
                     int64_t num_blocks, void* workspace, size_t workspace_bytes, void* stream) {
  using namespace tcs;
  if (n < 0 || tile_len <= 0 || num_blocks <= 0 || !out || (n > 0 && !x)) return TCS_ERR_INVALID_ARG;
  if (x_dtype != TCS_F16 && x_dtype != TCS_BF16 && x_dtype != TCS_F32 && x_dtype != TCS_I8 &&
      x_dtype != TCS_I32)
    return TCS_ERR_UNSUPPORTED_DTYPE;
  const int64_t bl = n > 0 ? tile_len * ceil_div(ceil_div(n, tile_len), num_blocks) : 1;
  const int64_t cpb = n > 0 ? ceil_div(bl, kRedChunk) : 1;
  if (!workspace || workspace_bytes < kHeaderBytes + (size_t)(num_blocks * cpb) * 8) return TCS_ERR_WORKSPACE;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const uint32_t grid = (uint32_t)(num_blocks * cpb);
  const uint32_t fin = (uint32_t)ceil_div(num_blocks, 256);
#define TCS_RED(CODEC, OUT)                                                                           \
  do {                                                                                                \
    using Acc = typename CODEC::Acc;                                                                  \
    auto* part = reinterpret_cast<Acc*>(static_cast<char*>(workspace) + kHeaderBytes);               \
    hipLaunchKernelGGL(reduce_partials_kernel<CODEC>, dim3(grid), dim3(kRedBlock), 0, st,            \
                       static_cast<const typename CODEC::In*>(x), n, bl, cpb, part);                  \
    TCS_CHECK_LAUNCH();                                                                               \
    hipLaunchKernelGGL((reduce_final_kernel<Acc, OUT>), dim3(fin), dim3(256), 0, st, part, num_blocks, \
                       cpb, static_cast<OUT*>(out));                                                  \
    TCS_CHECK_LAUNCH();                                                                               \
  } while (0)
  switch (x_dtype) {
    case TCS_F16: TCS_RED(CodecF16, float); break;
    case TCS_BF16: TCS_RED(CodecBF16, float); break;
    case TCS_F32: TCS_RED(CodecF32, float); break;
    case TCS_I8: TCS_RED(CodecI8, int); break;
    default: TCS_RED(CodecI32, int); break;
  }
#undef TCS_RED
  return TCS_OK;
}

size_t tcs_complete_blocks_workspace_bytes(int64_t num_blocks) {
  return tcs::kHeaderBytes + (size_t)(num_blocks > 0 ? num_blocks : 1) * 8;
}

int tcs_complete_blocks(const void* x, int dtype, const void* sums, int64_t num_blocks, void* y,
                        int64_t n, int64_t tile_len, void* workspace, size_t workspace_bytes,
                        void* stream) {
  using namespace tcs;
  if (n < 0 || num_blocks <= 0 || tile_len <= 0 || !sums || !y || (n > 0 && !x)) return TCS_ERR_INVALID_ARG;
  if (n % num_blocks != 0) return TCS_ERR_INVALID_ARG;
  if (dtype != TCS_F32 && dtype != TCS_I32) return TCS_ERR_UNSUPPORTED_DTYPE;
  if (!workspace || workspace_bytes < kHeaderBytes + (size_t)num_blocks * 8) return TCS_ERR_WORKSPACE;
  if ((reinterpret_cast<uintptr_t>(workspace) & 7u) != 0) return TCS_ERR_WORKSPACE;
  if (n == 0) return TCS_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int64_t bl = n / num_blocks;
  char* scratch = static_cast<char*>(workspace) + kHeaderBytes;
  const uint32_t grid = (uint32_t)ceil_div(n, 256);
  if (dtype == TCS_F32) {
    auto* pre = reinterpret_cast<double*>(scratch);
    hipLaunchKernelGGL((excl_prefix_kernel<float, double>), dim3(1), dim3(1024), 0, st,
                       static_cast<const float*>(sums), num_blocks, pre);
    hipLaunchKernelGGL((add_block_prefix_kernel<float, double>), dim3(grid), dim3(256), 0, st,
                       static_cast<const float*>(x), pre, static_cast<float*>(y), n, bl);
  } else {
    auto* pre = reinterpret_cast<int64_t*>(scratch);
    hipLaunchKernelGGL((excl_prefix_kernel<int32_t, int64_t>), dim3(1), dim3(1024), 0, st,
                       static_cast<const int32_t*>(sums), num_blocks, pre);
    hipLaunchKernelGGL((add_block_prefix_kernel<int32_t, int64_t>), dim3(grid), dim3(256), 0, st,
                       static_cast<const int32_t*>(x), pre, static_cast<int32_t*>(y), n, bl);
  }
  TCS_CHECK_LAUNCH();
  return fill_words(scratch, (size_t)num_blocks * 8, 0, st);
}

// complete_rows workspace: the nested scan's look-back state first (its
// ticket word is the workspace header), then block prefixes, row ends and
// their scan.
struct RowsPlan {
  int64_t height, tile, block_len, rows, rows_per_block;
  size_t pre_off, e_off, i_off, scan_off, scan_bytes, total;
};

static RowsPlan rows_plan(int64_t n, int64_t num_sums, int64_t w, int64_t h) {
  using tcs::ceil_div;
  RowsPlan p{};
  p.height = h < ceil_div(8192, w) ? h : ceil_div(8192, w);
  p.tile = w * p.height;
  p.block_len = n > 0 ? p.tile * ceil_div(ceil_div(n, p.tile), num_sums) : p.tile;
  p.rows = ceil_div(n, w);
  p.rows_per_block = p.block_len / w;
  auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
  p.scan_off = 0;
  p.scan_bytes = up(tcs_scan_workspace_bytes(p.rows, 128, 0));
  p.pre_off = p.scan_bytes;
  p.e_off = p.pre_off + up((size_t)num_sums * 8);
  p.i_off = p.e_off + up((size_t)p.rows * 4);
  p.total = p.i_off + up((size_t)p.rows * 4);
  return p;
}

size_t tcs_complete_rows_workspace_bytes(int64_t n, int64_t num_sums, int64_t tile_width,
                                         int64_t tile_height) {
  if (n <= 0 || num_sums <= 0 || tile_width <= 0 || tile_height <= 0) return tcs::kHeaderBytes;
  return rows_plan(n, num_sums, tile_width, tile_height).total;
}

int tcs_complete_rows(const void* x, int dtype, const void* sums, int64_t num_sums, void* y,
                      int64_t n, int64_t tile_width, int64_t tile_height, void* workspace,
                      size_t workspace_bytes, void* stream) {
  using namespace tcs;
  if (n < 0 || num_sums <= 0 || tile_width <= 0 || tile_height <= 0 || !sums || !y || (n > 0 && !x))
    return TCS_ERR_INVALID_ARG;
  if (dtype != TCS_F32 && dtype != TCS_I32) return TCS_ERR_UNSUPPORTED_DTYPE;
  if (n == 0) return TCS_OK;
  const RowsPlan p = rows_plan(n, num_sums, tile_width, tile_height);
  if (!workspace || workspace_bytes < p.total) return TCS_ERR_WORKSPACE;
  if ((reinterpret_cast<uintptr_t>(workspace) & 7u) != 0) return TCS_ERR_WORKSPACE;
  hipStream_t st = static_cast<hipStream_t>(stream);
  char* ws = static_cast<char*>(workspace);
  const uint32_t rgrid = (uint32_t)ceil_div(p.rows, 256);
  const uint32_t grid = (uint32_t)ceil_div(n, 256);
  const int code = dtype == TCS_F32 ? TCS_F32 : TCS_I32;
  if (dtype == TCS_F32) {
    auto* pre = reinterpret_cast<double*>(ws + p.pre_off);
    auto* e = reinterpret_cast<float*>(ws + p.e_off);
    auto* inc = reinterpret_cast<float*>(ws + p.i_off);
    hipLaunchKernelGGL((excl_prefix_kernel<float, double>), dim3(1), dim3(1024), 0, st,
                       static_cast<const float*>(sums), num_sums, pre);
    hipLaunchKernelGGL((gather_row_ends_kernel<float>), dim3(rgrid), dim3(256), 0, st,
                       static_cast<const float*>(x), n, tile_width, e, p.rows);
    TCS_CHECK_LAUNCH();
    const int rc = tcs_scan_rows(e, code, inc, p.rows, p.rows_per_block, 128, 0, ws + p.scan_off,
                                 p.scan_bytes, stream);
    if (rc != TCS_OK) return rc;
    hipLaunchKernelGGL((add_row_carry_kernel<float, double, float>), dim3(grid), dim3(256), 0, st,
                       static_cast<const float*>(x), pre, inc, static_cast<float*>(y), n, tile_width,
                       p.block_len);
  } else {
    auto* pre = reinterpret_cast<int64_t*>(ws + p.pre_off);
    auto* e = reinterpret_cast<int32_t*>(ws + p.e_off);
    auto* inc = reinterpret_cast<int32_t*>(ws + p.i_off);
    hipLaunchKernelGGL((excl_prefix_kernel<int32_t, int64_t>), dim3(1), dim3(1024), 0, st,
                       static_cast<const int32_t*>(sums), num_sums, pre);
    hipLaunchKernelGGL((gather_row_ends_kernel<int32_t>), dim3(rgrid), dim3(256), 0, st,
                       static_cast<const int32_t*>(x), n, tile_width, e, p.rows);
    TCS_CHECK_LAUNCH();
    const int rc = tcs_scan_rows(e, code, inc, p.rows, p.rows_per_block, 128, 0, ws + p.scan_off,
                                 p.scan_bytes, stream);
    if (rc != TCS_OK) return rc;
    hipLaunchKernelGGL((add_row_carry_kernel<int32_t, int64_t, int32_t>), dim3(grid), dim3(256), 0, st,
                       static_cast<const int32_t*>(x), pre, inc, static_cast<int32_t*>(y), n,
                       tile_width, p.block_len);
  }
  TCS_CHECK_LAUNCH();
  return fill_words(ws + p.pre_off, p.total - p.pre_off, 0, st);
}

}  // extern "C"
