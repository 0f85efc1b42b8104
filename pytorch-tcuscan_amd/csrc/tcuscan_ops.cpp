// tcuscan_ops: the PyTorch-facing module of tcuscan on MI355X.
//
// Mirrors the reference's pybind11 module (src/pybind11.cpp:36-142): the
// same function names, positional order, defaults and output dtypes, with
// every compute call forwarded to the C ABI of libtcuscan_hip.so
// (include/tcuscan.h).  Tensors are borrowed, outputs are new tensors,
// scratch space comes from a per-(device, stream) workspace that is zeroed
// once and then reused (the look-back kernels leave it ready for the next
// call).  Calls are stream-ordered on the current HIP stream and do not
// synchronise except where an output size must reach the host (compaction,
// flag segment sums), as in the reference.
#include <torch/extension.h>

#include <ATen/Parallel.h>
#include <ATen/hip/impl/HIPGuardImplMasqueradingAsCUDA.h>
#include <ATen/hip/impl/HIPStreamMasqueradingAsCUDA.h>

#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>
#include <utility>

#include "../../include/tcuscan.h"

namespace tcuscan {
namespace {

// ---------------------------------------------------------------------------
// Plumbing
// ---------------------------------------------------------------------------
int dtype_code(at::ScalarType t) {
  switch (t) {
    case at::kHalf: return TCS_F16;
    case at::kBFloat16: return TCS_BF16;
    case at::kFloat: return TCS_F32;
    case at::kChar: return TCS_I8;
    case at::kShort: return TCS_I16;
    case at::kInt: return TCS_I32;
    case at::kUInt32: return TCS_U32;
    case at::kLong: return TCS_I64;
    default: return -1;
  }
}

void check(int status, const char* op) {
  TORCH_CHECK(status == TCS_OK, op, ": ", tcs_status_string(status));
}

void* stream_of(const at::Tensor& t) {
  return static_cast<void*>(at::hip::getCurrentHIPStreamMasqueradingAsCUDA(t.device().index()).stream());
}

// Contiguous, 16-byte aligned view of `t` (a copy only when `t` is a
// strided or offset view).
at::Tensor prep(const at::Tensor& t) {
  at::Tensor c = t.contiguous();
  if ((reinterpret_cast<uintptr_t>(c.data_ptr()) & 15u) != 0) c = c.clone();
  return c;
}

void require_device(const at::Tensor& t, const char* op, const char* name) {
  TORCH_CHECK(t.is_cuda(), op, ": ", name, " must be a GPU tensor");
}

// Per-(device, stream) workspace, zeroed on allocation and again every
// TCS_WORKSPACE_MAX_CALLS calls (launch tags repeat after 2^30 - 1 look-back
// launches on one workspace).  `polls` is the look-back poll limit set by
// _set_lookback_poll_limit (-1: the default), re-applied whenever the
// workspace is replaced or re-zeroed.
struct Workspace {
  at::Tensor t;
  uint64_t calls = 0;
  int polls = -1;
};
std::mutex g_ws_mu;
std::map<std::pair<int, void*>, Workspace> g_ws;

// keep = true fetches the workspace for a call that reads what the previous
// call on it left there (a counted compaction after its count pass): it is
// neither replaced nor re-zeroed, and the call is not counted.
Workspace& workspace_entry(const at::Tensor& like, size_t bytes, bool keep = false) {
  const int dev = like.device().index();
  void* s = stream_of(like);
  auto& w = g_ws[{dev, s}];
  if (keep) {
    TORCH_CHECK(w.t.defined() && (size_t)w.t.numel() >= bytes, "workspace: the count pass left no workspace");
    return w;
  }
  bool fresh = false;
  if (!w.t.defined() || (size_t)w.t.numel() < bytes) {
    size_t want = bytes < (1u << 20) ? (1u << 20) : bytes;
    if (w.t.defined() && want < 2 * (size_t)w.t.numel()) want = 2 * (size_t)w.t.numel();
    w.t = at::zeros({(int64_t)want}, at::TensorOptions().dtype(at::kByte).device(like.device()));
    w.calls = 0;
    fresh = true;
  } else if (++w.calls >= TCS_WORKSPACE_MAX_CALLS) {
    const at::hip::HIPGuardMasqueradingAsCUDA guard(like.device());
    check(tcs_workspace_reset(w.t.data_ptr(), (size_t)w.t.numel(), s), "workspace reset");
    w.calls = 0;
    fresh = true;
  }
  if (fresh && w.polls >= 0) {
    const at::hip::HIPGuardMasqueradingAsCUDA guard(like.device());
    check(tcs_workspace_set_poll_limit(w.t.data_ptr(), (size_t)w.t.numel(), w.polls, s), "workspace poll limit");
  }
  return w;
}

std::pair<void*, size_t> workspace(const at::Tensor& like, size_t bytes, bool keep = false) {
  std::lock_guard<std::mutex> lock(g_ws_mu);
  Workspace& w = workspace_entry(like, bytes, keep);
  return {w.t.data_ptr(), (size_t)w.t.numel()};
}

at::ScalarType scan_out_dtype(at::ScalarType t) {
  return (t == at::kHalf || t == at::kBFloat16 || t == at::kFloat) ? at::kFloat : at::kInt;
}

// Inclusive scan of n elements of x (already prepped) into y.
void scan_into(const at::Tensor& x, at::Tensor& y, int64_t n, int S, double start, const char* op) {
  TORCH_CHECK(S > 0, op, ": S must be positive");
  const int code = dtype_code(x.scalar_type());
  auto ws = workspace(x, tcs_scan_workspace_bytes(n, S, 0));
  const at::hip::HIPGuardMasqueradingAsCUDA guard(x.device());
  check(tcs_scan(x.data_ptr(), code, y.data_ptr(), n, S, 0, start, ws.first, ws.second, stream_of(x)), op);
}

void scan_rows_into(const at::Tensor& x, at::Tensor& y, int64_t n, int64_t row_len, int S, const char* op) {
  TORCH_CHECK(S > 0, op, ": S must be positive");
  TORCH_CHECK(row_len > 0, op, ": row length must be positive");
  const int code = dtype_code(x.scalar_type());
  auto ws = workspace(x, tcs_scan_workspace_bytes(n, S, 0));
  const at::hip::HIPGuardMasqueradingAsCUDA guard(x.device());
  check(tcs_scan_rows(x.data_ptr(), code, y.data_ptr(), n, row_len, S, 0, ws.first, ws.second,
                      stream_of(x)),
        op);
}

}  // namespace

// ---------------------------------------------------------------------------
// Scan family (reference: src/torch/torch_scan.h, torch_scan_cpu.h)
// ---------------------------------------------------------------------------

// src/torch/torch_scan.h:47-110
at::Tensor run_scan_single_core(const at::Tensor& x_in, int S, double starting_sum) {
  const char* op = "run_scan_single_core";
  const auto t = x_in.scalar_type();
  TORCH_CHECK(t == at::kHalf || t == at::kFloat || t == at::kChar || t == at::kBFloat16, op,
              ": x must be float16, bfloat16, float32 or int8, got ", t);
  require_device(x_in, op, "x");
  at::Tensor x = prep(x_in);
  at::Tensor z = at::empty({x.numel()}, x.options().dtype(scan_out_dtype(t)));
  scan_into(x, z, x.numel(), S, starting_sum, op);
  return z;
}

at::Tensor multi_core_scan(const at::Tensor& x_in, int S, const char* op) {
  const auto t = x_in.scalar_type();
  TORCH_CHECK(t == at::kHalf || t == at::kChar || t == at::kBFloat16, op,
              ": x must be float16, bfloat16 or int8, got ", t);
  require_device(x_in, op, "x");
  at::Tensor x = prep(x_in);
  at::Tensor z = at::empty({x.numel()}, x.options().dtype(scan_out_dtype(t)));
  scan_into(x, z, x.numel(), S, 0.0, op);
  return z;
}

// src/torch/torch_scan.h:184-241
at::Tensor run_scan_multi_core(const at::Tensor& x, int S) {
  return multi_core_scan(x, S, "run_scan_multi_core");
}

// src/torch/torch_scan.h:251-302 (the L2-chunking distinction of the
// reference has no counterpart on a single-pass scan; same kernel).
at::Tensor run_scan_multi_core_no_l2(const at::Tensor& x, int S) {
  return multi_core_scan(x, S, "run_scan_multi_core_no_l2");
}

// Not in the reference: the plain scan with an explicit tile size
// (tiles_per_wg S x S tiles, 0 = default) and local scan method ("auto",
// "dpp" or "mfma"); every choice gives bit-identical results.
at::Tensor run_scan_ex(const at::Tensor& x_in, int S, int64_t tiles_per_wg, const std::string& method,
                       double starting_sum, int64_t lookback_window) {
  const char* op = "run_scan_ex";
  TORCH_CHECK(S > 0, op, ": S must be positive");
  TORCH_CHECK(tiles_per_wg >= 0, op, ": tiles_per_wg must be non-negative");
  TORCH_CHECK(lookback_window >= 0 && lookback_window <= 64, op, ": lookback_window must be in [0, 64]");
  int m = TCS_METHOD_AUTO;
  if (method == "dpp") {
    m = TCS_METHOD_DPP;
  } else if (method == "mfma") {
    m = TCS_METHOD_MFMA;
  } else {
    TORCH_CHECK(method == "auto", op, ": method must be 'auto', 'dpp' or 'mfma', got '", method, "'");
  }
  const auto t = x_in.scalar_type();
  TORCH_CHECK(t == at::kHalf || t == at::kFloat || t == at::kChar || t == at::kBFloat16 || t == at::kInt,
              op, ": x must be float16, bfloat16, float32, int8 or int32, got ", t);
  require_device(x_in, op, "x");
  at::Tensor x = prep(x_in);
  at::Tensor z = at::empty({x.numel()}, x.options().dtype(scan_out_dtype(t)));
  const int64_t n = x.numel();
  const int tpw = (int)tiles_per_wg;
  auto ws = workspace(x, tcs_scan_workspace_bytes(n, S, tpw));
  const at::hip::HIPGuardMasqueradingAsCUDA guard(x.device());
  check(tcs_scan_method(x.data_ptr(), dtype_code(t), z.data_ptr(), n, S, tpw, starting_sum, m,
                        (int)lookback_window, ws.first, ws.second, stream_of(x)),
        op);
  return z;
}

// src/torch/torch_scan.h:416-466: full scan; S is upper.size(0).
at::Tensor run_scan_multi_cube(const at::Tensor& x_in, const at::Tensor& upper,
                               const at::Tensor& lower_strict) {
  const char* op = "run_scan_multi_cube";
  TORCH_CHECK(upper.dim() == 2 && upper.size(0) == upper.size(1), op, ": upper must be S x S");
  TORCH_CHECK(lower_strict.sizes() == upper.sizes(), op, ": lower_strict must be S x S");
  const auto t = x_in.scalar_type();
  TORCH_CHECK(t == at::kHalf || t == at::kBFloat16, op, ": x must be float16 or bfloat16, got ", t);
  require_device(x_in, op, "x");
  at::Tensor x = prep(x_in);
  at::Tensor z = at::empty({x.numel()}, x.options().dtype(at::kFloat));
  scan_into(x, z, x.numel(), (int)upper.size(0), 0.0, op);
  return z;
}

// src/torch/torch_scan.h:119-175: row-wise scan of [B, L]
at::Tensor run_scan_batch(const at::Tensor& x_in, int S) {
  const char* op = "run_scan_batch";
  const auto t = x_in.scalar_type();
  TORCH_CHECK(t == at::kHalf || t == at::kFloat || t == at::kBFloat16 || t == at::kChar, op,
              ": x must be float16, bfloat16, float32 or int8, got ", t);
  TORCH_CHECK(x_in.dim() >= 1, op, ": x must have at least one dimension");
  require_device(x_in, op, "x");
  at::Tensor x = prep(x_in);
  const int64_t L = x.size(-1);
  TORCH_CHECK(L > 0, op, ": rows must be non-empty");
  const int64_t B = x.numel() / L;
  at::Tensor z = at::empty({B, L}, x.options().dtype(scan_out_dtype(t)));
  scan_rows_into(x, z, x.numel(), L, S, op);
  return z;
}

// src/torch/torch_scan.h:312-351: scan of each S-row of x[M, S]
at::Tensor run_row_scan(const at::Tensor& x_in, int S) {
  const char* op = "run_row_scan";
  const auto t = x_in.scalar_type();
  TORCH_CHECK(t == at::kHalf || t == at::kBFloat16 || t == at::kFloat || t == at::kChar, op,
              ": x must be float16, bfloat16, float32 or int8, got ", t);
  TORCH_CHECK(S > 0, op, ": S must be positive");
  require_device(x_in, op, "x");
  at::Tensor x = prep(x_in);
  const int64_t M = x.dim() == 0 ? 1 : x.size(0);
  const int64_t n = M * (int64_t)S;
  TORCH_CHECK(x.numel() >= n, op, ": x holds ", x.numel(), " elements, fewer than M*S = ", n);
  at::Tensor z = at::empty({n}, x.options().dtype(scan_out_dtype(t)));
  scan_rows_into(x, z, n, S, S, op);
  return z;
}

// src/torch/torch_scan.h:362-405: scan of each S*S block, S = upper.size(0)
at::Tensor run_block_scan(const at::Tensor& x_in, const at::Tensor& upper,
                          const at::Tensor& lower_strict) {
  const char* op = "run_block_scan";
  TORCH_CHECK(upper.dim() == 2 && upper.size(0) == upper.size(1), op, ": upper must be S x S");
  TORCH_CHECK(lower_strict.sizes() == upper.sizes(), op, ": lower_strict must be S x S");
  const auto t = x_in.scalar_type();
  TORCH_CHECK(t == at::kHalf || t == at::kBFloat16 || t == at::kFloat || t == at::kChar, op,
              ": x must be float16, bfloat16, float32 or int8, got ", t);
  require_device(x_in, op, "x");
  at::Tensor x = prep(x_in);
  const int64_t s = upper.size(0);
  at::Tensor z = at::empty({x.numel()}, x.options().dtype(scan_out_dtype(t)));
  scan_rows_into(x, z, x.numel(), s * s, (int)s, op);
  return z;
}

// src/torch/torch_seg_ops.h:164-193 (and run_seg_scan_vec :46-73):
// f[i] = 1 starts a segment at i; output float32.
at::Tensor seg_scan_flags(const at::Tensor& x_in, const at::Tensor& f_in, int S, const char* op) {
  const auto t = x_in.scalar_type();
  TORCH_CHECK(t == at::kHalf || t == at::kBFloat16 || t == at::kFloat || t == at::kChar, op,
              ": x must be float16, bfloat16, float32 or int8, got ", t);
  TORCH_CHECK(f_in.scalar_type() == at::kChar || f_in.scalar_type() == at::kBool ||
                  f_in.scalar_type() == at::kByte,
              op, ": f must be int8, uint8 or bool");
  TORCH_CHECK(f_in.numel() == x_in.numel(), op, ": x and f must have the same length");
  TORCH_CHECK(S > 0, op, ": S must be positive");
  require_device(x_in, op, "x");
  require_device(f_in, op, "f");
  at::Tensor x = prep(x_in);
  at::Tensor f = prep(f_in.scalar_type() == at::kChar ? f_in : f_in.to(at::kChar));
  const int64_t n = x.numel();
  at::Tensor z = at::empty({n}, x.options().dtype(scan_out_dtype(t)));
  auto ws = workspace(x, tcs_scan_workspace_bytes(n, S, 0));
  const at::hip::HIPGuardMasqueradingAsCUDA guard(x.device());
  check(tcs_seg_scan_flags(x.data_ptr(), dtype_code(t), static_cast<const int8_t*>(f.data_ptr()),
                           z.data_ptr(), n, S, 0, ws.first, ws.second, stream_of(x)),
        op);
  return t == at::kChar ? z : z;
}

at::Tensor run_seg_scan(const at::Tensor& x, const at::Tensor& f, int S) {
  return seg_scan_flags(x, f, S, "run_seg_scan");
}
at::Tensor run_seg_scan_vec(const at::Tensor& x, const at::Tensor& f, int S) {
  return seg_scan_flags(x, f, S, "run_seg_scan_vec");
}

// src/torch/torch_scan_cpu.h:23-107: three-phase blocked scan on the host.
template <typename T>
static void blocked_inclusive_scan(T* a, int64_t n) {
  if (n <= 0) return;
  const int64_t threads = std::max<int64_t>(1, at::get_num_threads());
  const int64_t blocks = std::min<int64_t>(threads, n);
  const int64_t bs = (n + blocks - 1) / blocks;
  std::vector<T> totals(blocks, T(0)), offset(blocks, T(0));
  at::parallel_for(0, blocks, 1, [&](int64_t b0, int64_t b1) {
    for (int64_t b = b0; b < b1; ++b) {
      const int64_t s = b * bs, e = std::min(s + bs, n);
      if (s >= e) continue;
      T run = a[s];
      for (int64_t i = s + 1; i < e; ++i) {
        run += a[i];
        a[i] = run;
      }
      totals[b] = run;
    }
  });
  T running = T(0);
  for (int64_t b = 0; b < blocks; ++b) {
    offset[b] = running;
    running += totals[b];
  }
  at::parallel_for(0, blocks, 1, [&](int64_t b0, int64_t b1) {
    for (int64_t b = b0; b < b1; ++b) {
      const T off = offset[b];
      if (off == T(0)) continue;
      const int64_t s = b * bs, e = std::min(s + bs, n);
      for (int64_t i = s; i < e; ++i) a[i] += off;
    }
  });
}

at::Tensor run_scan_cpu(const at::Tensor& x) {
  const char* op = "run_scan_cpu";
  TORCH_CHECK(x.device().is_cpu(), op, ": x must be a CPU tensor");
  const auto t = x.scalar_type();
  TORCH_CHECK(t == at::kHalf || t == at::kFloat || t == at::kChar, op,
              ": x must be float16, float32 or int8, got ", t);
  at::Tensor z = x.contiguous().reshape({-1}).to(scan_out_dtype(t)).clone();
  if (t == at::kChar) {
    blocked_inclusive_scan<int32_t>(z.data_ptr<int32_t>(), z.numel());
  } else {
    blocked_inclusive_scan<float>(z.data_ptr<float>(), z.numel());
  }
  return z;
}

// ---------------------------------------------------------------------------
// Reduction / down-sweep building blocks (reference: src/torch/torch_reduce.h)
// ---------------------------------------------------------------------------

// src/torch/torch_reduce.h:42-75
at::Tensor run_reduce_tiles(const at::Tensor& x_in, int64_t tile_len, int64_t num_blocks) {
  const char* op = "run_reduce_tiles";
  const auto t = x_in.scalar_type();
  TORCH_CHECK(t == at::kHalf || t == at::kChar || t == at::kBFloat16 || t == at::kFloat || t == at::kInt,
              op, ": x must be float16, bfloat16, float32, int8 or int32, got ", t);
  TORCH_CHECK(tile_len > 0 && num_blocks > 0, op, ": tile_len and num_blocks must be positive");
  require_device(x_in, op, "x");
  at::Tensor x = prep(x_in);
  at::Tensor z = at::empty({num_blocks}, x.options().dtype(scan_out_dtype(t)));
  auto ws = workspace(x, tcs_reduce_tiles_workspace_bytes(x.numel(), tile_len, num_blocks));
  const at::hip::HIPGuardMasqueradingAsCUDA guard(x.device());
  check(tcs_reduce_tiles(x.data_ptr(), dtype_code(t), z.data_ptr(), x.numel(), tile_len, num_blocks,
                         ws.first, ws.second, stream_of(x)),
        op);
  return z;
}

// src/torch/torch_reduce.h:143-182
at::Tensor run_complete_rows(const at::Tensor& x_in, const at::Tensor& sums_in, int tile_width,
                             int tile_height) {
  const char* op = "run_complete_rows";
  // The reference accepts float32 and int64 (int64 runs its int32 kernel
  // and the output is int32, torch_reduce.h:145-147,165-170); int32 is
  // accepted too.  Integer inputs and sums are taken as int32 values (an
  // int64 outside the int32 range wraps, as .to(torch.int32) does).
  const auto t = x_in.scalar_type();
  const auto st = sums_in.scalar_type();
  TORCH_CHECK(t == at::kFloat || t == at::kInt || t == at::kLong, op,
              ": x must be float32, int32 or int64, got ", t);
  TORCH_CHECK(t == at::kFloat ? st == at::kFloat : (st == at::kInt || st == at::kLong), op,
              ": sums must be float32 for float32 x and int32 or int64 for integer x, got ", st);
  TORCH_CHECK(tile_width > 0 && tile_height > 0, op, ": tile sizes must be positive");
  TORCH_CHECK(sums_in.numel() > 0, op, ": sums must not be empty");
  require_device(x_in, op, "x");
  require_device(sums_in, op, "sums");
  at::Tensor x = prep(t == at::kLong ? x_in.to(at::kInt) : x_in);
  at::Tensor sums = prep(st == at::kLong ? sums_in.to(at::kInt) : sums_in);
  at::Tensor z = at::empty({x.numel()}, x.options());
  auto ws = workspace(x, tcs_complete_rows_workspace_bytes(x.numel(), sums.numel(), tile_width, tile_height));
  const at::hip::HIPGuardMasqueradingAsCUDA guard(x.device());
  check(tcs_complete_rows(x.data_ptr(), dtype_code(x.scalar_type()), sums.data_ptr(), sums.numel(), z.data_ptr(),
                          x.numel(), tile_width, tile_height, ws.first, ws.second, stream_of(x)),
        op);
  return z;
}

// src/torch/torch_reduce.h:197-231
at::Tensor run_complete_blocks(const at::Tensor& x_in, const at::Tensor& sums_in, int tile_length) {
  const char* op = "run_complete_blocks";
  const auto t = x_in.scalar_type();
  TORCH_CHECK(t == at::kFloat || t == at::kInt, op, ": x must be float32 or int32, got ", t);
  TORCH_CHECK(sums_in.scalar_type() == t, op, ": sums must have the dtype of x");
  TORCH_CHECK(tile_length > 0, op, ": tile_length must be positive");
  TORCH_CHECK(sums_in.numel() > 0 && x_in.numel() % sums_in.numel() == 0, op,
              ": len(x) must be a multiple of len(sums)");
  require_device(x_in, op, "x");
  require_device(sums_in, op, "sums");
  at::Tensor x = prep(x_in), sums = prep(sums_in);
  at::Tensor z = at::empty_like(x);
  auto ws = workspace(x, tcs_complete_blocks_workspace_bytes(sums.numel()));
  const at::hip::HIPGuardMasqueradingAsCUDA guard(x.device());
  check(tcs_complete_blocks(x.data_ptr(), dtype_code(t), sums.data_ptr(), sums.numel(), z.data_ptr(),
                            x.numel(), tile_length, ws.first, ws.second, stream_of(x)),
        op);
  return z;
}

// ---------------------------------------------------------------------------
// Segmented sums and SpMV (reference: src/torch/torch_seg_ops.h, torch_spmv.h)
// ---------------------------------------------------------------------------
namespace {

int index_code(const at::Tensor& t) {
  switch (t.scalar_type()) {
    case at::kInt: return TCS_I32;
    case at::kUInt32: return TCS_U32;
    case at::kLong: return TCS_I64;
    default: return -1;
  }
}

// Segment sums with row ends = ends_base[1 : 1 + num_ends] (the element
// after the first), num_segments outputs.
at::Tensor seg_sum_impl(const at::Tensor& x_in, const at::Tensor& idx_in, int64_t num_segments,
                        int64_t num_ends, const char* op) {
  const auto t = x_in.scalar_type();
  TORCH_CHECK(t == at::kHalf || t == at::kChar || t == at::kBFloat16 || t == at::kFloat, op,
              ": x must be float16, bfloat16, float32 or int8, got ", t);
  TORCH_CHECK(index_code(idx_in) >= 0, op, ": indptr must be int32, uint32 or int64, got ",
              idx_in.scalar_type());
  TORCH_CHECK(num_segments >= 0, op, ": indptr is too short");
  require_device(x_in, op, "x");
  require_device(idx_in, op, "indptr");
  at::Tensor x = prep(x_in);
  at::Tensor idx = idx_in.contiguous();
  at::Tensor z = at::empty({num_segments}, x.options().dtype(scan_out_dtype(t)));
  if (num_segments == 0) return z;
  auto ws = workspace(x, tcs_seg_sum_workspace_bytes(x.numel()));
  const void* ends = static_cast<const char*>(idx.data_ptr()) + idx.element_size();
  const at::hip::HIPGuardMasqueradingAsCUDA guard(x.device());
  check(tcs_seg_sum_indptr(x.numel() ? x.data_ptr() : nullptr, dtype_code(t), ends, index_code(idx),
                           num_segments, num_ends, z.data_ptr(), x.numel(), ws.first, ws.second,
                           stream_of(x)),
        op);
  return z;
}

void check_segm_offsets(const c10::optional<at::Tensor>& so, const char* op) {
  // The reference lets callers pass the per-core first segments of its
  // fixed nnz partition (torch_seg_ops.h:362-375).  Work here is split by
  // nonzeros inside the kernel, so the hint is accepted and not needed.
  if (so.has_value() && so->defined()) {
    TORCH_CHECK(so->dim() == 1, op, ": segm_offsets must be 1-D");
  }
}

}  // namespace

// src/torch/torch_seg_ops.h:211-266 (full CSR indptr, n + 1 entries)
at::Tensor run_seg_sum_single_core(const at::Tensor& x, const at::Tensor& indptr, int s) {
  TORCH_CHECK(s > 0, "run_seg_sum_single_core: s must be positive");
  const int64_t R = indptr.numel() - 1;
  return seg_sum_impl(x, indptr, R, R, "run_seg_sum_single_core");
}

// src/torch/torch_seg_ops.h:333-426
at::Tensor run_seg_sum_multi_core(const at::Tensor& x, const at::Tensor& indptr, int s,
                                  c10::optional<at::Tensor> segm_offsets) {
  TORCH_CHECK(s > 0, "run_seg_sum_multi_core: s must be positive");
  check_segm_offsets(segm_offsets, "run_seg_sum_multi_core");
  const int64_t R = indptr.numel() - 1;
  return seg_sum_impl(x, indptr, R, R, "run_seg_sum_multi_core");
}

// src/torch/torch_seg_ops.h:278-322: indptr holds the n segment starts
// (indptr[:-1] of a CSR matrix); the last segment ends at len(x).
at::Tensor run_seg_sum_single_cube(const at::Tensor& x, const at::Tensor& upper,
                                   const at::Tensor& lower_strict, const at::Tensor& indptr, int s) {
  const char* op = "run_seg_sum_single_cube";
  TORCH_CHECK(s > 0, op, ": s must be positive");
  TORCH_CHECK(upper.dim() == 2 && lower_strict.sizes() == upper.sizes(), op,
              ": upper / lower_strict must be S x S");
  const int64_t R = indptr.numel();
  return seg_sum_impl(x, indptr, R, R > 0 ? R - 1 : 0, op);
}

// src/torch/torch_seg_ops.h:443-522 (fp16 only, full CSR indptr)
at::Tensor run_seg_sum_multi_cube(const at::Tensor& x, const at::Tensor& upper,
                                  const at::Tensor& lower_strict, const at::Tensor& indptr,
                                  c10::optional<at::Tensor> segm_offsets) {
  const char* op = "run_seg_sum_multi_cube";
  TORCH_CHECK(x.scalar_type() == at::kHalf, op, ": x must be fp16, got ", x.scalar_type());
  TORCH_CHECK(upper.dim() == 2 && lower_strict.sizes() == upper.sizes(), op,
              ": upper / lower_strict must be S x S");
  check_segm_offsets(segm_offsets, op);
  const int64_t R = indptr.numel() - 1;
  return seg_sum_impl(x, indptr, R, R, op);
}

// src/torch/torch_spmv.h:112-209
at::Tensor run_spmv_v2(const at::Tensor& vals_in, const at::Tensor& indptr_in,
                       const at::Tensor& cols_in, const at::Tensor& x_in, int s,
                       c10::optional<at::Tensor> segm_offsets) {
  const char* op = "run_spmv_v2";
  const auto vt = vals_in.scalar_type(), xt = x_in.scalar_type();
  TORCH_CHECK((vt == at::kHalf && xt == at::kHalf) || (vt == at::kFloat && xt == at::kFloat), op,
              ": vals and x must both be fp16 or both be fp32, got vals=", vt, " x=", xt);
  TORCH_CHECK(indptr_in.scalar_type() == at::kInt || indptr_in.scalar_type() == at::kUInt32, op,
              ": indptr must be int32 or uint32, got ", indptr_in.scalar_type());
  TORCH_CHECK(cols_in.scalar_type() == at::kInt || cols_in.scalar_type() == at::kUInt32, op,
              ": cols must be int32 or uint32, got ", cols_in.scalar_type());
  TORCH_CHECK(cols_in.numel() == vals_in.numel(), op, ": cols and vals must have the same length");
  TORCH_CHECK(s > 0, op, ": s must be positive");
  check_segm_offsets(segm_offsets, op);
  require_device(vals_in, op, "vals");
  require_device(indptr_in, op, "indptr");
  require_device(cols_in, op, "cols");
  require_device(x_in, op, "x");
  at::Tensor vals = prep(vals_in), cols = prep(cols_in), x = x_in.contiguous();
  at::Tensor indptr = indptr_in.contiguous();
  const int64_t R = indptr.numel() - 1;
  TORCH_CHECK(R >= 0, op, ": indptr must not be empty");
  at::Tensor z = at::empty({R}, vals.options().dtype(at::kFloat));
  if (R == 0) return z;
  auto ws = workspace(vals, tcs_spmv_workspace_bytes(vals.numel(), R, x.numel(), dtype_code(vt)));
  const void* ends = static_cast<const char*>(indptr.data_ptr()) + indptr.element_size();
  const at::hip::HIPGuardMasqueradingAsCUDA guard(vals.device());
  check(tcs_spmv_csr(vals.numel() ? vals.data_ptr() : nullptr, dtype_code(vt), ends, index_code(indptr),
                     static_cast<const int32_t*>(cols.numel() ? cols.data_ptr() : nullptr),
                     x.numel() ? x.data_ptr() : nullptr, x.numel(), z.data_ptr<float>(), R,
                     vals.numel(), ws.first, ws.second, stream_of(vals)),
        op);
  return z;
}

// src/torch/torch_spmv.h:235-302: fp16 vals and x, int32 indptr; S is the
// side of `upper` (the tiling hint of the reference's block scan).
at::Tensor run_spmv_v2_multi_cube(const at::Tensor& vals, const at::Tensor& indptr, const at::Tensor& cols,
                                  const at::Tensor& x, const at::Tensor& upper, const at::Tensor& lower_strict) {
  const char* op = "run_spmv_v2_multi_cube";
  TORCH_CHECK(vals.scalar_type() == at::kHalf && x.scalar_type() == at::kHalf, op,
              ": vals and x must both be fp16, got vals=", vals.scalar_type(), " x=", x.scalar_type());
  TORCH_CHECK(indptr.scalar_type() == at::kInt, op, ": indptr must be int32, got ", indptr.scalar_type());
  TORCH_CHECK(upper.dim() == 2 && upper.size(0) == upper.size(1) && lower_strict.sizes() == upper.sizes(), op,
              ": upper / lower_strict must be S x S");
  return run_spmv_v2(vals, indptr, cols, x, (int)upper.size(0), c10::nullopt);
}

// src/torch/torch_seg_ops.h:108-153: z[i] = scan_x[i] - diff[scan_f[i] - 1]
// (scan_x in segment 0), float32.
at::Tensor run_seg_scan_mc_revert(const at::Tensor& scan_x_in, const at::Tensor& scan_f_in,
                                  const at::Tensor& diff_in) {
  const char* op = "run_seg_scan_mc_revert";
  TORCH_CHECK(scan_x_in.scalar_type() == at::kFloat, op, ": x must be float32, got ", scan_x_in.scalar_type());
  TORCH_CHECK(scan_f_in.scalar_type() == at::kInt, op, ": f must be int32, got ", scan_f_in.scalar_type());
  TORCH_CHECK(diff_in.scalar_type() == at::kFloat, op, ": diff must be float32, got ", diff_in.scalar_type());
  TORCH_CHECK(scan_f_in.numel() == scan_x_in.numel(), op, ": x and f must have the same length");
  require_device(scan_x_in, op, "x");
  require_device(scan_f_in, op, "f");
  require_device(diff_in, op, "diff");
  at::Tensor sx = prep(scan_x_in), sf = prep(scan_f_in), d = diff_in.contiguous();
  const int64_t n = sx.numel();
  const at::hip::HIPGuardMasqueradingAsCUDA guard(sx.device());
  at::Tensor z = at::empty({n}, sx.options());
  check(tcs_seg_scan_revert(sx.data_ptr<float>(), sf.data_ptr<int32_t>(), d.numel() ? d.data_ptr<float>() : nullptr,
                            d.numel(), z.data_ptr<float>(), n, stream_of(sx)),
        op);
  return z;
}

// Top-p mask step (scripts/profiling/profile_tcuscan_ops.py:352-353, torch
// code in the reference): kept = where(mass - x > p, 0, x) in float32, for
// the probabilities x sorted largest first and their inclusive mass.
at::Tensor run_top_p_keep(const at::Tensor& sorted_in, const at::Tensor& mass_in, double p) {
  const char* op = "run_top_p_keep";
  TORCH_CHECK(sorted_in.scalar_type() == at::kHalf, op, ": probs_sort must be float16, got ",
              sorted_in.scalar_type());
  TORCH_CHECK(mass_in.scalar_type() == at::kFloat, op, ": probs_sum must be float32, got ", mass_in.scalar_type());
  TORCH_CHECK(sorted_in.numel() == mass_in.numel(), op, ": probs_sort and probs_sum must have the same length");
  require_device(sorted_in, op, "probs_sort");
  require_device(mass_in, op, "probs_sum");
  at::Tensor x = prep(sorted_in), m = prep(mass_in);
  const at::hip::HIPGuardMasqueradingAsCUDA guard(x.device());
  at::Tensor kept = at::empty({x.numel()}, m.options());
  check(tcs_top_p_keep(x.data_ptr(), m.data_ptr<float>(), (float)p, kept.data_ptr<float>(), x.numel(), stream_of(x)),
        op);
  return kept;
}

// src/torch/torch_gather.h:92-172
at::Tensor run_csr_gather(const at::Tensor& values_in, const at::Tensor& cols_in, const at::Tensor& x_in) {
  const char* op = "run_csr_gather";
  TORCH_CHECK(values_in.dim() == 1, op, ": values must be 1D, got ", values_in.dim(), "D");
  TORCH_CHECK(cols_in.dim() == 1, op, ": cols must be 1D, got ", cols_in.dim(), "D");
  TORCH_CHECK(x_in.dim() == 1, op, ": x must be 1D, got ", x_in.dim(), "D");
  TORCH_CHECK(cols_in.numel() == values_in.numel(), op, ": cols length (", cols_in.numel(),
              ") must match values length (", values_in.numel(), ")");
  const auto t = values_in.scalar_type();
  TORCH_CHECK(t == at::kHalf || t == at::kShort || t == at::kFloat, op,
              ": values dtype must be fp16, int16, or fp32, got ", t);
  TORCH_CHECK(x_in.scalar_type() == t, op, ": x dtype (", x_in.scalar_type(),
              ") must match values dtype (", t, ")");
  TORCH_CHECK(cols_in.scalar_type() == at::kInt, op, ": cols dtype must be int32, got ",
              cols_in.scalar_type());
  TORCH_CHECK(values_in.device() == cols_in.device() && values_in.device() == x_in.device(), op,
              ": values, cols and x must be on the same device");
  require_device(values_in, op, "values");
  at::Tensor v = prep(values_in), c = prep(cols_in), x = x_in.contiguous();
  at::Tensor z = at::empty_like(v);
  const at::hip::HIPGuardMasqueradingAsCUDA guard(v.device());
  check(tcs_csr_gather(v.data_ptr(), dtype_code(t), static_cast<const int32_t*>(c.data_ptr()),
                       x.numel() ? x.data_ptr() : v.data_ptr(), x.numel(), z.data_ptr(), v.numel(),
                       stream_of(v)),
        op);
  return z;
}

// ---------------------------------------------------------------------------
// Compaction (reference: src/torch/torch_compress.h)
// ---------------------------------------------------------------------------
namespace {

at::Tensor mask_i8(const at::Tensor& m, const char* op) {
  TORCH_CHECK(m.scalar_type() == at::kChar || m.scalar_type() == at::kByte || m.scalar_type() == at::kBool,
              op, ": mask must be int8 (uint8 and bool are accepted), got ", m.scalar_type());
  return prep(m.scalar_type() == at::kChar ? m : m.to(at::kChar));
}

// Number of mask == 1 entries, brought to the host (the reference syncs
// the same way to size its outputs, torch_compress.h:62-64).
int64_t count_ones(const at::Tensor& mask, const char* op) {
  at::Tensor cnt = at::empty({1}, mask.options().dtype(at::kLong));
  auto ws = workspace(mask, tcs_compress_workspace_bytes(mask.numel()));
  const at::hip::HIPGuardMasqueradingAsCUDA guard(mask.device());
  check(tcs_count_mask(static_cast<const int8_t*>(mask.data_ptr()), mask.numel(), cnt.data_ptr<int64_t>(),
                       ws.first, ws.second, stream_of(mask)),
        op);
  return cnt.item<int64_t>();
}

// counted = true: count_ones(mask) ran last on this stream's workspace, so
// the compaction takes its chunk offsets from there (tcs_compress_counted);
// otherwise it finds them by look-back (tcs_compress).
void compress_into(const at::Tensor& x, const at::Tensor* idx_in, int idx_mode, const at::Tensor& mask,
                   at::Tensor& out_x, at::Tensor* out_idx, const char* op, bool counted) {
  const int64_t n = x.numel();
  if (n == 0) return;
  auto ws = workspace(x, tcs_compress_workspace_bytes(n), counted);
  const at::hip::HIPGuardMasqueradingAsCUDA guard(x.device());
  auto fn = counted ? tcs_compress_counted : tcs_compress;
  check(fn(x.data_ptr(), (int)x.element_size(), idx_in ? idx_in->data_ptr() : nullptr, idx_mode,
           idx_in ? (int)idx_in->element_size() : 4, static_cast<const int8_t*>(mask.data_ptr()),
           out_x.numel() ? out_x.data_ptr() : nullptr, (out_idx && out_idx->numel()) ? out_idx->data_ptr() : nullptr,
           n, out_x.numel(), ws.first, ws.second, stream_of(x)),
        op);
}

void compress_checks(const at::Tensor& x, const at::Tensor& mask, int S, const char* op) {
  const auto es = x.element_size();
  TORCH_CHECK(es == 1 || es == 2 || es == 4 || es == 8, op, ": unsupported element size");
  TORCH_CHECK(mask.scalar_type() == at::kChar || mask.scalar_type() == at::kByte || mask.scalar_type() == at::kBool,
              op, ": mask must be int8 (uint8 and bool are accepted), got ", mask.scalar_type());
  TORCH_CHECK(mask.numel() == x.numel(), op, ": x and mask must have the same length");
  TORCH_CHECK(S > 0, op, ": S must be positive");
  require_device(x, op, "x");
  require_device(mask, op, "mask");
}

}  // namespace

// src/torch/torch_compress.h:41-98
at::Tensor run_compress(const at::Tensor& x_in, const at::Tensor& mask_in, int S) {
  const char* op = "run_compress";
  compress_checks(x_in, mask_in, S, op);
  at::Tensor x = prep(x_in), mask = mask_i8(mask_in, op);
  const int64_t k = count_ones(mask, op);
  at::Tensor z = at::empty({k}, x.options());
  compress_into(x, nullptr, 0, mask, z, nullptr, op, true);
  return z;
}

// src/torch/torch_compress.h:109-161: the output length is given.
at::Tensor run_compress_pos(const at::Tensor& x_in, const at::Tensor& mask_in, int64_t output_len, int S) {
  const char* op = "run_compress_pos";
  compress_checks(x_in, mask_in, S, op);
  TORCH_CHECK(output_len >= 0, op, ": output_len must be non-negative");
  at::Tensor x = prep(x_in), mask = mask_i8(mask_in, op);
  at::Tensor z = at::empty({output_len}, x.options());
  compress_into(x, nullptr, 0, mask, z, nullptr, op, false);
  return z;
}

// src/torch/torch_compress.h:172-239
std::tuple<at::Tensor, at::Tensor> run_compress_ind(const at::Tensor& x_in, const at::Tensor& idx_in_t,
                                                    const at::Tensor& mask_in, int S) {
  const char* op = "run_compress_ind";
  compress_checks(x_in, mask_in, S, op);
  TORCH_CHECK(idx_in_t.numel() == x_in.numel(), op, ": indices_in and x must have the same length");
  TORCH_CHECK(idx_in_t.element_size() == 4 || idx_in_t.element_size() == 8, op,
              ": indices_in must have 4- or 8-byte elements");
  require_device(idx_in_t, op, "indices_in");
  at::Tensor x = prep(x_in), mask = mask_i8(mask_in, op), idx = prep(idx_in_t);
  const int64_t k = count_ones(mask, op);
  at::Tensor z = at::empty({k}, x.options());
  at::Tensor zi = at::empty({k}, idx.options());
  compress_into(x, &idx, 1, mask, z, &zi, op, true);
  return std::make_tuple(z, zi);
}

// src/torch/torch_compress.h:249-309
std::tuple<at::Tensor, at::Tensor> run_compress_ind_no_arange(const at::Tensor& x_in,
                                                              const at::Tensor& mask_in, int S) {
  const char* op = "run_compress_ind_no_arange";
  compress_checks(x_in, mask_in, S, op);
  at::Tensor x = prep(x_in), mask = mask_i8(mask_in, op);
  const int64_t k = count_ones(mask, op);
  at::Tensor z = at::empty({k}, x.options());
  at::Tensor zi = at::empty({k}, x.options().dtype(at::kInt));
  compress_into(x, nullptr, 2, mask, z, &zi, op, true);
  return std::make_tuple(z, zi);
}

// src/torch/torch_seg_ops.h:85-107: f[i] = 1 marks the last element of a
// segment; one float32 sum per flag.  Summed per segment directly (the
// reference differences a global prefix, which loses precision).
at::Tensor run_seg_sum(const at::Tensor& x_in, const at::Tensor& f_in, int S) {
  const char* op = "run_seg_sum";
  TORCH_CHECK(S > 0, op, ": S must be positive");
  TORCH_CHECK(f_in.numel() == x_in.numel(), op, ": x and f must have the same length");
  const auto t = x_in.scalar_type();
  TORCH_CHECK(t == at::kHalf || t == at::kBFloat16 || t == at::kFloat || t == at::kChar, op,
              ": x must be float16, bfloat16, float32 or int8, got ", t);
  require_device(x_in, op, "x");
  require_device(f_in, op, "f");
  at::Tensor x = prep(x_in), f = mask_i8(f_in, op);
  const int64_t k = count_ones(f, op);
  at::Tensor ends = at::empty({k + 1}, x.options().dtype(at::kInt));
  at::Tensor unused = at::empty({k}, f.options());
  if (k > 0) {
    at::Tensor ends_tail = ends.narrow(0, 1, k);
    // ends[1 + j] = (position of the j-th flag) + 1.
    auto ws = workspace(f, tcs_compress_workspace_bytes(f.numel()), true);
    const at::hip::HIPGuardMasqueradingAsCUDA guard(x.device());
    check(tcs_compress_counted(f.data_ptr(), 1, nullptr, 3, 4, static_cast<const int8_t*>(f.data_ptr()),
                       unused.data_ptr(), ends_tail.data_ptr(), f.numel(), k, ws.first, ws.second,
                       stream_of(x)),
          op);
  }
  at::Tensor sums = seg_sum_impl(x, ends, k, k, op);
  return sums.scalar_type() == at::kFloat ? sums : sums.to(at::kFloat);
}

int abi_version() { return tcs_abi_version(); }

// The batch tuner's cache: (kind, dtype code, log2n, S, tiles_per_wg,
// method, window, ms) per entry, for tcuscan.tuning to persist.
std::vector<std::tuple<int, int, int, int, int, int, int, double>> tuned_entries() {
  std::vector<std::tuple<int, int, int, int, int, int, int, double>> out;
  const int n = tcs_batch_tuned_count();
  for (int i = 0; i < n; ++i) {
    int k = 0, d = 0, l = 0, S = 0, t = 0, m = 0, w = 0;
    float ms = -1.0f;
    check(tcs_batch_tuned_entry(i, &k, &d, &l, &S, &t, &m, &w, &ms), "_tuned_entries");
    out.emplace_back(k, d, l, S, t, m, w, (double)ms);
  }
  return out;
}

void set_tuned(int kind, int dtype, int log2n, int S, int tiles_per_wg, int method, int window, double ms) {
  check(tcs_batch_tuned_set(kind, dtype, log2n, S, tiles_per_wg, method, window, (float)ms), "_set_tuned");
}

// Workspace of the current stream of like's device (created if needed), so
// that tests can inspect and prepare its launch state.  The tensor shares
// the workspace's memory.
at::Tensor workspace_tensor(const at::Tensor& like) {
  require_device(like, "_workspace_tensor", "like");
  std::lock_guard<std::mutex> lock(g_ws_mu);
  return workspace_entry(like, 0).t;
}

// Calls counted on that workspace since it was last zeroed; the next call
// that finds TCS_WORKSPACE_MAX_CALLS re-zeroes it first.
void workspace_set_calls(const at::Tensor& like, int64_t calls) {
  require_device(like, "_workspace_set_calls", "like");
  TORCH_CHECK(calls >= 0 && calls < (int64_t)TCS_WORKSPACE_MAX_CALLS, "_workspace_set_calls: out of range");
  std::lock_guard<std::mutex> lock(g_ws_mu);
  workspace_entry(like, 0).calls = (uint64_t)calls;
}

// Look-back poll limit of the workspace of the current stream of like's
// device (0: every walk falls back at once to computing a missing
// predecessor from the input; negative: the default).  Results do not
// depend on it; tests use it to exercise the forward-progress fallback.
void set_lookback_poll_limit(const at::Tensor& like, int polls) {
  require_device(like, "_set_lookback_poll_limit", "like");
  std::lock_guard<std::mutex> lock(g_ws_mu);
  Workspace& w = workspace_entry(like, 0);
  w.polls = polls < 0 ? -1 : polls;
  const at::hip::HIPGuardMasqueradingAsCUDA guard(like.device());
  check(tcs_workspace_set_poll_limit(w.t.data_ptr(), (size_t)w.t.numel(), w.polls, stream_of(like)),
        "_set_lookback_poll_limit");
}

// ---------------------------------------------------------------------------
// Batched work plans (no counterpart in the reference): the C ABI's
// tcs_batch with torch tensors.  Inputs are bound by address, so they must
// be contiguous, 16-byte aligned GPU tensors that the caller keeps (the plan
// holds references); outputs are allocated by the add_* calls.
// ---------------------------------------------------------------------------
class BatchPlan {
 public:
  BatchPlan() { check(tcs_batch_create(&b_), "BatchPlan"); }
  ~BatchPlan() { tcs_batch_destroy(b_); }
  BatchPlan(const BatchPlan&) = delete;
  BatchPlan& operator=(const BatchPlan&) = delete;

  at::Tensor add_scan(const at::Tensor& x) {
    const char* op = "BatchPlan.add_scan";
    bind(x, op, "x");
    const auto t = x.scalar_type();
    at::Tensor y = at::empty({x.numel()}, x.options().dtype(scan_out_dtype(t)));
    added(tcs_batch_add_scan(b_, x.data_ptr(), dtype_code(t), y.data_ptr(), x.numel()), op, {x, y});
    return y;
  }

  at::Tensor add_seg_scan(const at::Tensor& x, const at::Tensor& flags) {
    const char* op = "BatchPlan.add_seg_scan";
    bind(x, op, "x");
    bind(flags, op, "flags");
    TORCH_CHECK(flags.scalar_type() == at::kChar && flags.numel() == x.numel(), op,
                ": flags must be int8 with one flag per element");
    const auto t = x.scalar_type();
    at::Tensor y = at::empty({x.numel()}, x.options().dtype(scan_out_dtype(t)));
    added(tcs_batch_add_seg_scan(b_, x.data_ptr(), dtype_code(t),
                                 static_cast<const int8_t*>(flags.data_ptr()), y.data_ptr(), x.numel()),
          op, {x, flags, y});
    return y;
  }

  // Full CSR indptr (num_segments + 1 entries), as run_seg_sum_multi_core.
  at::Tensor add_seg_sum(const at::Tensor& x, const at::Tensor& indptr) {
    const char* op = "BatchPlan.add_seg_sum";
    bind(x, op, "x");
    bind(indptr, op, "indptr");
    TORCH_CHECK(index_code(indptr) >= 0 && indptr.numel() >= 2, op,
                ": indptr must be int32, uint32 or int64 with at least two entries");
    const int64_t R = indptr.numel() - 1;
    at::Tensor y = at::empty({R}, x.options().dtype(scan_out_dtype(x.scalar_type())));
    const void* ends = static_cast<const char*>(indptr.data_ptr()) + indptr.element_size();
    added(tcs_batch_add_seg_sum(b_, x.data_ptr(), dtype_code(x.scalar_type()), ends, index_code(indptr), R,
                                R, y.data_ptr(), x.numel()),
          op, {x, indptr, y});
    return y;
  }

  at::Tensor add_spmv(const at::Tensor& vals, const at::Tensor& indptr, const at::Tensor& cols,
                      const at::Tensor& x) {
    const char* op = "BatchPlan.add_spmv";
    bind(vals, op, "vals");
    bind(indptr, op, "indptr");
    bind(cols, op, "cols");
    bind(x, op, "x");
    TORCH_CHECK(cols.scalar_type() == at::kInt && cols.numel() == vals.numel(), op,
                ": cols must be int32, one per value");
    TORCH_CHECK(index_code(indptr) >= 0 && indptr.numel() >= 2, op, ": bad indptr");
    TORCH_CHECK(x.scalar_type() == vals.scalar_type(), op, ": x and vals must share a dtype");
    const int64_t R = indptr.numel() - 1;
    at::Tensor y = at::empty({R}, vals.options().dtype(at::kFloat));
    const void* ends = static_cast<const char*>(indptr.data_ptr()) + indptr.element_size();
    added(tcs_batch_add_spmv(b_, vals.data_ptr(), dtype_code(vals.scalar_type()), ends, index_code(indptr),
                             static_cast<const int32_t*>(cols.data_ptr()), x.data_ptr(), x.numel(),
                             static_cast<float*>(y.data_ptr()), R, vals.numel()),
          op, {vals, indptr, cols, x, y});
    return y;
  }

  int num_ops() const { return tcs_batch_num_ops(b_); }

  void tune(int reps, bool force) {
    ensure_ws();
    const at::hip::HIPGuardMasqueradingAsCUDA guard(ws_.device());
    check(tcs_batch_tune(b_, ws_.data_ptr(), (size_t)ws_.numel(), reps, force ? 1 : 0, stream_of(ws_)),
          "BatchPlan.tune");
  }

  void run_eager() {
    ensure_ws();
    const at::hip::HIPGuardMasqueradingAsCUDA guard(ws_.device());
    check(tcs_batch_run_eager(b_, ws_.data_ptr(), (size_t)ws_.numel(), stream_of(ws_)), "BatchPlan.run_eager");
  }

  void instantiate(int max_branches) {
    ensure_ws();
    const at::hip::HIPGuardMasqueradingAsCUDA guard(ws_.device());
    check(tcs_batch_instantiate(b_, max_branches, ws_.data_ptr(), (size_t)ws_.numel(), stream_of(ws_)),
          "BatchPlan.instantiate");
  }

  void launch() {
    TORCH_CHECK(ws_.defined(), "BatchPlan.launch: call instantiate() first");
    const at::hip::HIPGuardMasqueradingAsCUDA guard(ws_.device());
    check(tcs_batch_launch(b_, stream_of(ws_)), "BatchPlan.launch");
  }

  // Per op: S, tiles_per_wg, method, window and the tuner's best time.
  std::vector<std::tuple<int, int, std::string, int, double>> configs() const {
    std::vector<std::tuple<int, int, std::string, int, double>> out;
    for (int i = 0; i < tcs_batch_num_ops(b_); ++i) {
      int S = 0, tpw = 0, m = 0, w = 0;
      float ms = -1.0f;
      check(tcs_batch_op_config(b_, i, &S, &tpw, &m, &w, &ms), "BatchPlan.configs");
      const char* mn = m == TCS_METHOD_DPP ? "dpp" : m == TCS_METHOD_MFMA ? "mfma" : "auto";
      out.emplace_back(S, tpw, std::string(mn), w, (double)ms);
    }
    return out;
  }

 private:
  void bind(const at::Tensor& t, const char* op, const char* name) {
    require_device(t, op, name);
    TORCH_CHECK(t.is_contiguous() && (reinterpret_cast<uintptr_t>(t.data_ptr()) & 15u) == 0, op, ": ",
                name, " must be contiguous and 16-byte aligned (the plan binds its address)");
    TORCH_CHECK(t.numel() > 0, op, ": ", name, " must not be empty");
    if (device_ < 0) device_ = t.device().index();
    TORCH_CHECK(t.device().index() == device_, op, ": all tensors of a plan must be on one device");
  }
  void added(int rc, const char* op, std::initializer_list<at::Tensor> keep) {
    TORCH_CHECK(rc >= 0, op, ": ", tcs_status_string(-rc));
    for (const auto& t : keep) keep_.push_back(t);
  }
  void ensure_ws() {
    TORCH_CHECK(device_ >= 0, "BatchPlan: no ops added");
    const size_t need = tcs_batch_workspace_bytes(b_);
    if (!ws_.defined() || (size_t)ws_.numel() < need)
      ws_ = at::zeros({(int64_t)need}, at::TensorOptions().dtype(at::kByte).device(at::kCUDA, device_));
  }

  tcs_batch* b_ = nullptr;
  std::vector<at::Tensor> keep_;
  at::Tensor ws_;
  int device_ = -1;
};

// ---------------------------------------------------------------------------
// Filters, where, searchsorted (reference: src/torch/torch_compress.h:319-398,
// src/torch/torch_searchsorted.h:33-82)
// ---------------------------------------------------------------------------

// x[x >= pivot] / x[x <= pivot] (values) or the int32 positions of the kept
// elements; pivot compared in x's type.  One count pass sizes the output
// (the reference reads a count back the same way), then one fused
// compare-and-compact pass.
at::Tensor filter_impl(const at::Tensor& x_in, double pivot, int cmp, bool values, int S, const char* op) {
  TORCH_CHECK(S > 0, op, ": S must be positive");
  const auto t = x_in.scalar_type();
  TORCH_CHECK(t == at::kHalf || t == at::kBFloat16 || t == at::kFloat, op,
              ": x must be float16, bfloat16 or float32, got ", t);
  require_device(x_in, op, "x");
  at::Tensor x = prep(x_in);
  const int64_t n = x.numel();
  TORCH_CHECK(n < ((int64_t)1 << 31), op, ": at most 2^31 - 1 elements");
  auto ws = workspace(x, tcs_compress_workspace_bytes(n));
  const at::hip::HIPGuardMasqueradingAsCUDA guard(x.device());
  at::Tensor cnt = at::zeros({1}, x.options().dtype(at::kLong));
  if (n > 0)
    check(tcs_count_filter(x.data_ptr(), dtype_code(t), pivot, cmp, n, cnt.data_ptr<int64_t>(), ws.first,
                           ws.second, stream_of(x)),
          op);
  const int64_t k = cnt.item<int64_t>();
  at::Tensor out = values ? at::empty({k}, x.options()) : at::empty({k}, x.options().dtype(at::kInt));
  if (k > 0) {
    // The count pass left the chunk counts in this workspace.
    auto wk = workspace(x, tcs_compress_workspace_bytes(n), true);
    check(tcs_filter_counted(x.data_ptr(), dtype_code(t), pivot, cmp, values ? out.data_ptr() : nullptr,
                             values ? nullptr : out.data_ptr<int32_t>(), n, k, wk.first, wk.second,
                             stream_of(x)),
          op);
  }
  return out;
}

// src/torch/torch_compress.h:319-322
at::Tensor run_filter_greater_equal(const at::Tensor& x, float pivot, int S) {
  return filter_impl(x, pivot, TCS_CMP_GE, true, S, "run_filter_greater_eq");
}

// src/torch/torch_compress.h:332-335
at::Tensor run_filter_less_equal(const at::Tensor& x, float pivot, int S) {
  return filter_impl(x, pivot, TCS_CMP_LE, true, S, "run_filter_less_eq");
}

// src/torch/torch_compress.h:345-398: int32 indices where x >= pivot.
at::Tensor run_where(const at::Tensor& x, float pivot, int S) {
  return filter_impl(x, pivot, TCS_CMP_GE, false, S, "run_where");
}

// ---------------------------------------------------------------------------
// Split (reference: src/torch/torch_split.h:25-141)
// ---------------------------------------------------------------------------

std::pair<at::Tensor, at::Tensor> split_impl(const at::Tensor& x_in, const at::Tensor& mask_in,
                                             const at::Tensor* idx_in, int S, const char* op) {
  TORCH_CHECK(S > 0, op, ": S must be positive");
  TORCH_CHECK(mask_in.scalar_type() == at::kChar, op, ": mask must be int8, got ", mask_in.scalar_type());
  TORCH_CHECK(mask_in.numel() == x_in.numel(), op, ": mask and x must have the same length");
  const int eb = (int)x_in.element_size();
  TORCH_CHECK(eb == 1 || eb == 2 || eb == 4 || eb == 8, op, ": unsupported dtype ", x_in.scalar_type());
  require_device(x_in, op, "x");
  require_device(mask_in, op, "mask");
  at::Tensor x = prep(x_in), mask = prep(mask_in), idx;
  if (idx_in) {
    require_device(*idx_in, op, "indices_in");
    TORCH_CHECK(idx_in->scalar_type() == at::kInt, op, ": indices_in must be int32, got ", idx_in->scalar_type());
    TORCH_CHECK(idx_in->numel() == x_in.numel(), op, ": indices_in and x must have the same length");
    idx = prep(*idx_in);
  }
  const int64_t n = x.numel();
  TORCH_CHECK(n < ((int64_t)1 << 31), op, ": at most 2^31 - 1 elements");
  const at::hip::HIPGuardMasqueradingAsCUDA guard(x.device());
  at::Tensor z = at::empty({n}, x.options());
  at::Tensor zi = idx_in ? at::empty({n}, x.options().dtype(at::kInt)) : at::Tensor();
  if (n > 0) {
    auto ws = workspace(x, tcs_split_workspace_bytes(n));
    check(tcs_split(x.data_ptr(), eb, mask.data_ptr<int8_t>(), idx_in ? idx.data_ptr<int32_t>() : nullptr,
                    z.data_ptr(), idx_in ? zi.data_ptr<int32_t>() : nullptr, n, ws.first, ws.second,
                    stream_of(x)),
          op);
  }
  return {z, zi};
}

at::Tensor run_split(const at::Tensor& x, const at::Tensor& mask, int S) {
  return split_impl(x, mask, nullptr, S, "run_split").first;
}

std::tuple<at::Tensor, at::Tensor> run_split_ind(const at::Tensor& x, const at::Tensor& mask,
                                                 const at::Tensor& indices_in, int S) {
  auto r = split_impl(x, mask, &indices_in, S, "run_split_ind");
  return {r.first, r.second};
}

// ---------------------------------------------------------------------------
// Radix sort (reference: src/torch/torch_sort.h:21-81)
// ---------------------------------------------------------------------------

std::tuple<at::Tensor, at::Tensor> sort_impl(const at::Tensor& x_in, bool descending, const char* op) {
  const int dt = dtype_code(x_in.scalar_type());
  TORCH_CHECK(dt >= 0 && tcs_sort_workspace_bytes(1, dt) > 0, op,
              ": x must be float16, bfloat16, float32, int8, int16, int32 or uint32, got ", x_in.scalar_type());
  require_device(x_in, op, "x");
  at::Tensor x = prep(x_in);
  const int64_t n = x.numel();
  TORCH_CHECK(n < ((int64_t)1 << 31), op, ": at most 2^31 - 1 elements");
  const at::hip::HIPGuardMasqueradingAsCUDA guard(x.device());
  at::Tensor z = at::empty({n}, x.options());
  at::Tensor zi = at::empty({n}, x.options().dtype(at::kInt));
  if (n > 0) {
    auto ws = workspace(x, tcs_sort_workspace_bytes(n, dt));
    at::Tensor scratch = at::empty({(int64_t)tcs_sort_scratch_bytes(n, dt)}, x.options().dtype(at::kByte));
    check(tcs_sort(x.data_ptr(), dt, nullptr, z.data_ptr(), zi.data_ptr<int32_t>(), n, descending ? 1 : 0,
                   ws.first, ws.second, scratch.data_ptr(), (size_t)scratch.numel(), stream_of(x)),
          op);
  }
  return {z, zi};
}

// Ascending, as the reference's kernels sort (radix_sort.cpp: descending = false).
std::tuple<at::Tensor, at::Tensor> run_radix_sort(const at::Tensor& x, int S) {
  TORCH_CHECK(S > 0, "run_radix_sort: S must be positive");
  return sort_impl(x, false, "run_radix_sort");
}

std::tuple<at::Tensor, at::Tensor> run_radix_sort_ex(const at::Tensor& x, bool descending) {
  return sort_impl(x, descending, "run_radix_sort_ex");
}

// ---------------------------------------------------------------------------
// Top-k (reference: src/torch/torch_topk.h:23-207)
// ---------------------------------------------------------------------------

std::tuple<at::Tensor, at::Tensor> topk_impl(const at::Tensor& x_in, int64_t k, bool largest, const char* op) {
  const int dt = dtype_code(x_in.scalar_type());
  TORCH_CHECK(dt >= 0 && tcs_topk_workspace_bytes(1, 1, dt) > 0, op,
              ": x must be float16, bfloat16, float32, int8, int16, int32 or uint32, got ", x_in.scalar_type());
  require_device(x_in, op, "x");
  at::Tensor x = prep(x_in);
  const int64_t n = x.numel();
  TORCH_CHECK(n < ((int64_t)1 << 31), op, ": at most 2^31 - 1 elements");
  TORCH_CHECK(k >= 0 && k <= n, op, ": k must be in [0, ", n, "], got ", k);
  const at::hip::HIPGuardMasqueradingAsCUDA guard(x.device());
  at::Tensor v = at::empty({k}, x.options());
  at::Tensor i = at::empty({k}, x.options().dtype(at::kInt));
  if (k > 0) {
    auto ws = workspace(x, tcs_topk_workspace_bytes(n, k, dt));
    at::Tensor scratch = at::empty({(int64_t)tcs_topk_scratch_bytes(n, k, dt)}, x.options().dtype(at::kByte));
    check(tcs_topk(x.data_ptr(), dt, n, k, largest ? 1 : 0, v.data_ptr(), i.data_ptr<int32_t>(), ws.first,
                   ws.second, scratch.data_ptr(), (size_t)scratch.numel(), stream_of(x)),
          op);
  }
  return {v, i};
}

// src/torch/torch_topk.h:149-207; x_min / x_max bound the reference's value
// bisection and are not needed by a radix select.
std::tuple<at::Tensor, at::Tensor> run_topk_fp16(const at::Tensor& x, int64_t k, double x_min, double x_max, int S) {
  (void)x_min;
  (void)x_max;
  TORCH_CHECK(S > 0, "run_topk_fp16: S must be positive");
  TORCH_CHECK(x.scalar_type() == at::kHalf, "run_topk_fp16: x must be float16, got ", x.scalar_type());
  return topk_impl(x, k, true, "run_topk_fp16");
}

// src/torch/torch_topk.h:69-136
std::tuple<at::Tensor, at::Tensor> run_topk_int16(const at::Tensor& x, int64_t k, int64_t x_min, int64_t x_max,
                                                  int S) {
  (void)x_min;
  (void)x_max;
  TORCH_CHECK(S > 0, "run_topk_int16: S must be positive");
  TORCH_CHECK(x.scalar_type() == at::kShort, "run_topk_int16: x must be int16, got ", x.scalar_type());
  return topk_impl(x, k, true, "run_topk_int16");
}

std::tuple<at::Tensor, at::Tensor> run_topk(const at::Tensor& x, int64_t k, bool largest) {
  return topk_impl(x, k, largest, "run_topk");
}

// src/torch/torch_topk.h:23-60: k is accepted and unused, as there.
at::Tensor run_topk_pivot_fp16(const at::Tensor& x_in, int64_t k, int num_samples) {
  const char* op = "run_topk_pivot_fp16";
  (void)k;
  TORCH_CHECK(x_in.scalar_type() == at::kHalf, op, ": x must be float16, got ", x_in.scalar_type());
  TORCH_CHECK(num_samples >= 4 && num_samples <= 64, op, ": num_samples must be in [4, 64]");
  TORCH_CHECK(x_in.numel() > 0, op, ": x must not be empty");
  require_device(x_in, op, "x");
  at::Tensor x = prep(x_in);
  const int64_t blocks = tcs_topk_pivot_blocks(x.numel(), num_samples);
  const at::hip::HIPGuardMasqueradingAsCUDA guard(x.device());
  at::Tensor out = at::empty({blocks}, x.options());
  check(tcs_topk_pivot(x.data_ptr(), x.numel(), num_samples, out.data_ptr(), blocks, stream_of(x)), op);
  return out;
}

// src/torch/torch_compare.h:33-67: int32 [1] count of (x cmp pivot); the
// reference's tile_len is a tiling hint with no counterpart here.
at::Tensor run_count_if(const at::Tensor& x_in, float pivot, uint32_t tile_len, uint8_t compare_mode) {
  const char* op = "run_count_if";
  TORCH_CHECK(tile_len > 0, op, ": tile_len must be positive");
  TORCH_CHECK(compare_mode <= TCS_CMP_NE, op, ": compare_mode must be 0..5 (LT, GT, EQ, LE, GE, NE)");
  const auto t = x_in.scalar_type();
  TORCH_CHECK(t == at::kHalf || t == at::kBFloat16 || t == at::kFloat, op,
              ": x must be float16, bfloat16 or float32, got ", t);
  require_device(x_in, op, "x");
  at::Tensor x = prep(x_in);
  const int64_t n = x.numel();
  auto ws = workspace(x, tcs_compress_workspace_bytes(n));
  const at::hip::HIPGuardMasqueradingAsCUDA guard(x.device());
  at::Tensor cnt = at::zeros({1}, x.options().dtype(at::kLong));
  if (n > 0)
    check(tcs_count_filter(x.data_ptr(), dtype_code(t), pivot, compare_mode, n, cnt.data_ptr<int64_t>(),
                           ws.first, ws.second, stream_of(x)),
          op);
  return cnt.to(at::kInt);
}

// run_searchsorted takes the bucket index above these sizes.
constexpr int64_t kIndexedMinLen = (int64_t)1 << 16, kIndexedMinNeedles = (int64_t)1 << 16;

// src/torch/torch_searchsorted.h:33-82
at::Tensor run_searchsorted(const at::Tensor& sorted_in, const at::Tensor& values_in) {
  const char* op = "run_searchsorted";
  TORCH_CHECK(sorted_in.scalar_type() == at::kInt, op, ": sorted must be int32, got ", sorted_in.scalar_type());
  TORCH_CHECK(values_in.scalar_type() == at::kInt, op, ": values must be int32, got ", values_in.scalar_type());
  TORCH_CHECK(sorted_in.numel() < ((int64_t)1 << 31), op, ": at most 2^31 - 1 sorted elements");
  require_device(sorted_in, op, "sorted");
  require_device(values_in, op, "values");
  at::Tensor sorted = sorted_in.contiguous(), values = values_in.contiguous();
  at::Tensor out = at::empty({values.numel()}, values.options());  // flat, as the reference returns
  const at::hip::HIPGuardMasqueradingAsCUDA guard(values.device());
  // Haystacks beyond the LDS samples with many needles: the bucket index
  // (its build costs about one search per 16 haystack elements).
  if (sorted.numel() > kIndexedMinLen && values.numel() >= kIndexedMinNeedles) {
    // The bucket table is scratch of this call only: a buffer of its own, not
    // the stream's look-back workspace (whose header it would overwrite).
    const size_t tb = tcs_searchsorted_workspace_bytes(sorted.numel());
    at::Tensor table = at::empty({(int64_t)tb}, values.options().dtype(at::kByte));
    check(tcs_searchsorted_indexed(sorted.data_ptr<int32_t>(), sorted.numel(), values.data_ptr<int32_t>(),
                                   out.data_ptr<int32_t>(), values.numel(), table.data_ptr(), tb,
                                   stream_of(values)),
          op);
  } else {
    check(tcs_searchsorted(sorted.data_ptr<int32_t>(), sorted.numel(), values.data_ptr<int32_t>(),
                           out.data_ptr<int32_t>(), values.numel(), stream_of(values)),
          op);
  }
  return out;
}

void register_server(py::module& m);  // serve.cpp

}  // namespace tcuscan

PYBIND11_MODULE(tcuscan_ops, m) {
  tcuscan::register_server(m);
  py::class_<tcuscan::BatchPlan>(m, "BatchPlan",
                                 "Batched scans / seg-scans / seg-sums / SpMVs replayed as one HIP graph")
      .def(py::init<>())
      .def("add_scan", &tcuscan::BatchPlan::add_scan, py::arg("x"))
      .def("add_seg_scan", &tcuscan::BatchPlan::add_seg_scan, py::arg("x"), py::arg("flags"))
      .def("add_seg_sum", &tcuscan::BatchPlan::add_seg_sum, py::arg("x"), py::arg("indptr"))
      .def("add_spmv", &tcuscan::BatchPlan::add_spmv, py::arg("vals"), py::arg("indptr"), py::arg("cols"),
           py::arg("x"))
      .def("num_ops", &tcuscan::BatchPlan::num_ops)
      .def("tune", &tcuscan::BatchPlan::tune, py::arg("reps") = 3, py::arg("force") = false)
      .def("run_eager", &tcuscan::BatchPlan::run_eager)
      .def("instantiate", &tcuscan::BatchPlan::instantiate, py::arg("max_branches") = 4)
      .def("launch", &tcuscan::BatchPlan::launch)
      .def("configs", &tcuscan::BatchPlan::configs);
  namespace py = pybind11;
  m.doc() = "tcuscan operators for AMD Instinct MI355X (HIP, gfx950)";
  m.def("abi_version", &tcuscan::abi_version, "C ABI version of libtcuscan_hip.so");
  m.def("_workspace_tensor", &tcuscan::workspace_tensor, py::arg("like"),
        "The per-stream workspace of like's device (shares its memory)");
  m.def("_workspace_set_calls", &tcuscan::workspace_set_calls, py::arg("like"), py::arg("calls"),
        "Sets the call count of the per-stream workspace");
  m.def("_set_lookback_poll_limit", &tcuscan::set_lookback_poll_limit, py::arg("like"), py::arg("polls"),
        "Look-back poll limit of the per-stream workspace (0: always fall back; -1: default)");
  m.def("run_scan_multi_core", &tcuscan::run_scan_multi_core, "Multi-core Scan");
  m.def("run_scan_multi_core_no_l2", &tcuscan::run_scan_multi_core_no_l2,
        "Multi-core Scan (no L2 cache optimization)");
  m.def("run_scan_single_core", &tcuscan::run_scan_single_core, py::arg("x"), py::arg("S"),
        py::arg("starting_sum") = 0, "Scan Single Core");
  m.def("run_scan_batch", &tcuscan::run_scan_batch, "Scan Batch");
  m.def("run_row_scan", &tcuscan::run_row_scan, "Row scan of each S-row");
  m.def("run_block_scan", &tcuscan::run_block_scan, "Block scan on blocks of length S^2");
  m.def("run_scan_multi_cube", &tcuscan::run_scan_multi_cube, "Multi-cube scan");
  m.def("run_filter_greater_eq", &tcuscan::run_filter_greater_equal, "Filter by great-equal pivot.");
  m.def("run_filter_less_eq", &tcuscan::run_filter_less_equal, "Filter by less-equal pivot.");
  m.def("run_where", &tcuscan::run_where, "Indices where x >= pivot (int32).");
  m.def("run_count_if", &tcuscan::run_count_if, "Count if");
  m.def("run_split", &tcuscan::run_split, "Split (stable partition by mask == 1)");
  m.def("run_topk_fp16", &tcuscan::run_topk_fp16, "TopK (fp16): k largest values and their indices");
  m.def("run_topk_int16", &tcuscan::run_topk_int16, "TopK (int16): k largest values and their indices");
  m.def("run_topk", &tcuscan::run_topk, py::arg("x"), py::arg("k"), py::arg("largest") = true,
        "TopK of any sortable dtype, largest or smallest first");
  m.def("run_topk_pivot_fp16", &tcuscan::run_topk_pivot_fp16, py::arg("x"), py::arg("k"),
        py::arg("num_samples") = 32, "K-largest value estimator (fp16)");
  m.def("run_radix_sort", &tcuscan::run_radix_sort, "Radix sort with indices (ascending, stable)");
  m.def("run_radix_sort_ex", &tcuscan::run_radix_sort_ex, py::arg("x"), py::arg("descending") = false,
        "Radix sort with indices, ascending or descending (stable)");
  m.def("run_split_ind", &tcuscan::run_split_ind, "Split with indices");
  m.def("run_searchsorted", &tcuscan::run_searchsorted, py::arg("sorted"), py::arg("values"),
        "Binary search (lower_bound) into a sorted int32 array");
  m.def("run_scan_ex", &tcuscan::run_scan_ex, py::arg("x"), py::arg("S"), py::arg("tiles_per_wg") = 0,
        py::arg("method") = "auto", py::arg("starting_sum") = 0.0, py::arg("lookback_window") = 0,
        "Scan with an explicit tile size, local scan method ('auto', 'dpp', 'mfma') and look-back window");
  m.def("run_seg_scan", &tcuscan::run_seg_scan, "Segmented Scan");
  m.def("run_seg_scan_vec", &tcuscan::run_seg_scan_vec, "Segmented Scan (vector-only)");
  m.def("run_scan_cpu", &tcuscan::run_scan_cpu, "Scan on CPUs");
  m.def("run_reduce_tiles", &tcuscan::run_reduce_tiles, "Sum-reduce over tiles");
  m.def("run_complete_rows", &tcuscan::run_complete_rows, "Down-sweep (second) phase of MCSCAN");
  m.def("run_complete_blocks", &tcuscan::run_complete_blocks,
        "Block-wise down-sweep (second) phase of block scan");
  m.def("run_seg_sum", &tcuscan::run_seg_sum, "Segmented Sum");
  m.def("run_seg_sum_single_core", &tcuscan::run_seg_sum_single_core, "Segmented Sum (single-core)");
  m.def("run_seg_sum_multi_core", &tcuscan::run_seg_sum_multi_core, py::arg("x"), py::arg("indptr"),
        py::arg("s"), py::arg("segm_offsets") = py::none(), "Segmented Sum (multi-core)");
  m.def("run_seg_sum_single_cube", &tcuscan::run_seg_sum_single_cube, "Segmented Sum (single-cube)");
  m.def("run_seg_sum_multi_cube", &tcuscan::run_seg_sum_multi_cube, py::arg("x"), py::arg("upper"),
        py::arg("lower_strict"), py::arg("indptr"), py::arg("segm_offsets") = py::none(),
        "Segmented Sum (multi-cube)");
  m.def("_tuned_entries", &tcuscan::tuned_entries, "The batch tuner's cache of this process");
  m.def("_set_tuned", &tcuscan::set_tuned, py::arg("kind"), py::arg("dtype"), py::arg("log2n"), py::arg("S"),
        py::arg("tiles_per_wg"), py::arg("method"), py::arg("window"), py::arg("ms") = -1.0,
        "Adds or replaces one entry of the batch tuner's cache");
  m.def("run_spmv_v2", &tcuscan::run_spmv_v2, py::arg("vals"), py::arg("indptr"), py::arg("cols"),
        py::arg("x"), py::arg("s"), py::arg("segm_offsets") = py::none(),
        "Sparse Matrix-Vector Multiplication Using Segmented Sum");
  m.def("run_csr_gather", &tcuscan::run_csr_gather, "CSR gather");
  m.def("run_spmv_v2_multi_cube", &tcuscan::run_spmv_v2_multi_cube, py::arg("vals"), py::arg("indptr"),
        py::arg("cols"), py::arg("x"), py::arg("upper"), py::arg("lower_strict"),
        "SpMV with the multi-cube block scan (fp16)");
  m.def("run_seg_scan_mc_revert", &tcuscan::run_seg_scan_mc_revert, "Vector Revert for MC Segmented Scan");
  m.def("run_top_p_keep", &tcuscan::run_top_p_keep, "Top-p kept weights (mask step of nucleus sampling)",
        py::arg("probs_sort"), py::arg("probs_sum"), py::arg("p"));
  m.def("run_compress", &tcuscan::run_compress, "Compaction/compress");
  m.def("run_compress_pos", &tcuscan::run_compress_pos,
        "Compaction/compress with pre-computed output positions");
  m.def("run_compress_ind", &tcuscan::run_compress_ind, "Compaction that returns indices");
  m.def("run_compress_ind_no_arange", &tcuscan::run_compress_ind_no_arange,
        "Compaction that returns indices (no input indices)");
}
