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
#include <tuple>
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

// Per-(device, stream) workspace, zeroed on allocation.
std::mutex g_ws_mu;
std::map<std::pair<int, void*>, at::Tensor> g_ws;

std::pair<void*, size_t> workspace(const at::Tensor& like, size_t bytes) {
  const int dev = like.device().index();
  void* s = stream_of(like);
  std::lock_guard<std::mutex> lock(g_ws_mu);
  auto& w = g_ws[{dev, s}];
  if (!w.defined() || (size_t)w.numel() < bytes) {
    size_t want = bytes < (1u << 20) ? (1u << 20) : bytes;
    if (w.defined() && want < 2 * (size_t)w.numel()) want = 2 * (size_t)w.numel();
    w = at::zeros({(int64_t)want}, at::TensorOptions().dtype(at::kByte).device(like.device()));
  }
  return {w.data_ptr(), (size_t)w.numel()};
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
  require_device(x_in, op, "x");
  const auto t = x_in.scalar_type();
  TORCH_CHECK(t == at::kHalf || t == at::kFloat || t == at::kChar || t == at::kBFloat16, op,
              ": x must be float16, bfloat16, float32 or int8, got ", t);
  at::Tensor x = prep(x_in);
  at::Tensor z = at::empty({x.numel()}, x.options().dtype(scan_out_dtype(t)));
  scan_into(x, z, x.numel(), S, starting_sum, op);
  return z;
}

at::Tensor multi_core_scan(const at::Tensor& x_in, int S, const char* op) {
  require_device(x_in, op, "x");
  const auto t = x_in.scalar_type();
  TORCH_CHECK(t == at::kHalf || t == at::kChar || t == at::kBFloat16, op,
              ": x must be float16, bfloat16 or int8, got ", t);
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

// src/torch/torch_scan.h:416-466: full scan; S is upper.size(0).
at::Tensor run_scan_multi_cube(const at::Tensor& x_in, const at::Tensor& upper,
                               const at::Tensor& lower_strict) {
  const char* op = "run_scan_multi_cube";
  require_device(x_in, op, "x");
  TORCH_CHECK(upper.dim() == 2 && upper.size(0) == upper.size(1), op, ": upper must be S x S");
  TORCH_CHECK(lower_strict.sizes() == upper.sizes(), op, ": lower_strict must be S x S");
  const auto t = x_in.scalar_type();
  TORCH_CHECK(t == at::kHalf || t == at::kBFloat16, op, ": x must be float16 or bfloat16, got ", t);
  at::Tensor x = prep(x_in);
  at::Tensor z = at::empty({x.numel()}, x.options().dtype(at::kFloat));
  scan_into(x, z, x.numel(), (int)upper.size(0), 0.0, op);
  return z;
}

// src/torch/torch_scan.h:119-175: row-wise scan of [B, L]
at::Tensor run_scan_batch(const at::Tensor& x_in, int S) {
  const char* op = "run_scan_batch";
  require_device(x_in, op, "x");
  const auto t = x_in.scalar_type();
  TORCH_CHECK(t == at::kHalf || t == at::kFloat || t == at::kBFloat16 || t == at::kChar, op,
              ": x must be float16, bfloat16, float32 or int8, got ", t);
  TORCH_CHECK(x_in.dim() >= 1, op, ": x must have at least one dimension");
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
  require_device(x_in, op, "x");
  const auto t = x_in.scalar_type();
  TORCH_CHECK(t == at::kHalf || t == at::kBFloat16 || t == at::kFloat || t == at::kChar, op,
              ": x must be float16, bfloat16, float32 or int8, got ", t);
  TORCH_CHECK(S > 0, op, ": S must be positive");
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
  require_device(x_in, op, "x");
  TORCH_CHECK(upper.dim() == 2 && upper.size(0) == upper.size(1), op, ": upper must be S x S");
  TORCH_CHECK(lower_strict.sizes() == upper.sizes(), op, ": lower_strict must be S x S");
  const auto t = x_in.scalar_type();
  TORCH_CHECK(t == at::kHalf || t == at::kBFloat16 || t == at::kFloat || t == at::kChar, op,
              ": x must be float16, bfloat16, float32 or int8, got ", t);
  at::Tensor x = prep(x_in);
  const int64_t s = upper.size(0);
  at::Tensor z = at::empty({x.numel()}, x.options().dtype(scan_out_dtype(t)));
  scan_rows_into(x, z, x.numel(), s * s, (int)s, op);
  return z;
}

// src/torch/torch_seg_ops.h:164-193 (and run_seg_scan_vec :46-73):
// f[i] = 1 starts a segment at i; output float32.
at::Tensor seg_scan_flags(const at::Tensor& x_in, const at::Tensor& f_in, int S, const char* op) {
  require_device(x_in, op, "x");
  require_device(f_in, op, "f");
  const auto t = x_in.scalar_type();
  TORCH_CHECK(t == at::kHalf || t == at::kBFloat16 || t == at::kFloat || t == at::kChar, op,
              ": x must be float16, bfloat16, float32 or int8, got ", t);
  TORCH_CHECK(f_in.scalar_type() == at::kChar || f_in.scalar_type() == at::kBool ||
                  f_in.scalar_type() == at::kByte,
              op, ": f must be int8, uint8 or bool");
  TORCH_CHECK(f_in.numel() == x_in.numel(), op, ": x and f must have the same length");
  TORCH_CHECK(S > 0, op, ": S must be positive");
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

int abi_version() { return tcs_abi_version(); }

}  // namespace tcuscan

PYBIND11_MODULE(tcuscan_ops, m) {
  namespace py = pybind11;
  m.doc() = "tcuscan operators for AMD Instinct MI355X (HIP, gfx950)";
  m.def("abi_version", &tcuscan::abi_version, "C ABI version of libtcuscan_hip.so");
  m.def("run_scan_multi_core", &tcuscan::run_scan_multi_core, "Multi-core Scan");
  m.def("run_scan_multi_core_no_l2", &tcuscan::run_scan_multi_core_no_l2,
        "Multi-core Scan (no L2 cache optimization)");
  m.def("run_scan_single_core", &tcuscan::run_scan_single_core, py::arg("x"), py::arg("S"),
        py::arg("starting_sum") = 0, "Scan Single Core");
  m.def("run_scan_batch", &tcuscan::run_scan_batch, "Scan Batch");
  m.def("run_row_scan", &tcuscan::run_row_scan, "Row scan of each S-row");
  m.def("run_block_scan", &tcuscan::run_block_scan, "Block scan on blocks of length S^2");
  m.def("run_scan_multi_cube", &tcuscan::run_scan_multi_cube, "Multi-cube scan");
  m.def("run_seg_scan", &tcuscan::run_seg_scan, "Segmented Scan");
  m.def("run_seg_scan_vec", &tcuscan::run_seg_scan_vec, "Segmented Scan (vector-only)");
  m.def("run_scan_cpu", &tcuscan::run_scan_cpu, "Scan on CPUs");
}
