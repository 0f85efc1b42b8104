// Batched work plans: many scans, flag-segmented scans, segmented sums and
// CSR SpMVs over fixed device buffers, tuned once and replayed as one HIP
// graph launch.
//
// The reference launches and synchronises per call (src/torch/*.h: every
// wrapper ends in aclrtSynchronizeStream) and picks its tiling by fixed
// host rules (src/tiling/*.h).  Here a plan records its ops, an autotuner
// times every tiling candidate of each scan op on the op's own buffers
// (tiles of S*S*tiles_per_wg elements for S in {32, 64, 128}, the local scan
// method and the look-back window) and keeps the fastest, and the whole
// plan is captured into one graph whose independent ops run on parallel
// branches.  Each op owns a region of the caller's workspace, so branches
// never share look-back state.  Every look-back kernel takes its launch tag
// from the epoch in its workspace region, which is what makes replays of
// one captured graph correct.
#include <algorithm>
#include <cmath>
#include <iterator>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "tcs_common.hpp"

namespace {

enum OpKind { kOpScan = 0, kOpSegScan = 1, kOpSegSum = 2, kOpSpmv = 3 };

struct ScanCfg {
  int S = 128;
  int tiles_per_wg = 0;
  int method = TCS_METHOD_AUTO;
  int window = 0;
  float ms = -1.0f;  // best time measured by the tuner (-1: not tuned)
};

struct Op {
  OpKind kind;
  const void* x;
  int x_dtype;
  const void* aux;  // flags (seg scan), ends (seg sum, SpMV)
  int aux_dtype;
  const int32_t* cols;
  const void* xv;  // SpMV vector
  int64_t xv_len;
  void* y;
  int64_t n;  // elements (scans) or nonzeros (seg sum, SpMV)
  int64_t rows;
  int64_t num_ends;
  ScanCfg cfg;
  size_t ws_off = 0, ws_bytes = 0;
};

// Tuned configurations, shared by every plan of the process: key = (kind,
// dtype, ceil(log2(n))).
std::mutex g_tune_mu;
std::map<std::tuple<int, int, int>, ScanCfg> g_tuned;

int ceil_log2(int64_t n) {
  int l = 0;
  while (l < 62 && ((int64_t)1 << l) < n) ++l;
  return l;
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

bool is_scan(OpKind k) { return k == kOpScan || k == kOpSegScan; }

int run_op(const Op& op, const ScanCfg& c, char* ws, hipStream_t st) {
  void* w = ws + op.ws_off;
  switch (op.kind) {
    case kOpScan:
      return tcs_scan_method(op.x, op.x_dtype, op.y, op.n, c.S, c.tiles_per_wg, 0.0, c.method, c.window,
                             w, op.ws_bytes, st);
    case kOpSegScan:
      return tcs::seg_scan_flags_cfg(op.x, op.x_dtype, static_cast<const int8_t*>(op.aux), op.y, op.n, c.S,
                                     c.tiles_per_wg, c.window, w, op.ws_bytes, st);
    case kOpSegSum:
      return tcs_seg_sum_indptr(op.x, op.x_dtype, op.aux, op.aux_dtype, op.rows, op.num_ends, op.y, op.n,
                                w, op.ws_bytes, st);
    case kOpSpmv:
      return tcs_spmv_csr(op.x, op.x_dtype, op.aux, op.aux_dtype, op.cols, op.xv, op.xv_len,
                          static_cast<float*>(op.y), op.rows, op.n, w, op.ws_bytes, st);
  }
  return TCS_ERR_INVALID_ARG;
}

// Candidates of one scan op: distinct effective tiles over S in {32, 64,
// 128} and tiles_per_wg in {1, 2, ..., 32}, times methods and windows.
std::vector<ScanCfg> candidates(const Op& op) {
  std::vector<ScanCfg> out;
  std::vector<int64_t> seen;
  const int g = (op.x_dtype == TCS_F32 || op.x_dtype == TCS_I32) ? 4 : 8;
  const int64_t max_tile = 64 * 64 * g;
  std::vector<int> methods = {TCS_METHOD_DPP};
  if (op.kind == kOpScan && (op.x_dtype == TCS_F16 || op.x_dtype == TCS_BF16 || op.x_dtype == TCS_I8))
    methods.push_back(TCS_METHOD_MFMA);
  for (int S : {32, 64, 128}) {
    for (int tpw = 1; tpw <= 32; tpw *= 2) {
      int64_t tile = tcs::tile_elems_for(S, tpw);
      if (tile > max_tile) tile = max_tile;
      if (std::find(seen.begin(), seen.end(), tile) != seen.end()) continue;
      seen.push_back(tile);
      for (int m : methods)
        for (int win : {6, 8, 16, 32}) {
          ScanCfg c;
          c.S = S;
          c.tiles_per_wg = tpw;
          c.method = m;
          c.window = win;
          out.push_back(c);
        }
    }
  }
  return out;
}

}  // namespace

struct tcs_batch {
  std::vector<Op> ops;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t origin = nullptr;  // capture stream (the legacy default stream cannot be captured)
  std::vector<hipStream_t> branches;
  std::vector<hipEvent_t> events;
  void* ws = nullptr;
  size_t ws_bytes = 0;
  // Look-back launches on the workspace since it was last zeroed (each op
  // run or graph launch launches at most one per op region); the plan
  // re-zeroes it before TCS_WORKSPACE_MAX_CALLS.
  uint64_t launches = 0;

  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
  ~tcs_batch() {
    drop_graph();
    if (origin) (void)hipStreamDestroy(origin);
    for (hipStream_t s : branches) (void)hipStreamDestroy(s);
    for (hipEvent_t e : events) (void)hipEventDestroy(e);
  }
};

namespace {

int add_op(tcs_batch* b, Op op) {
  if (!b) return -TCS_ERR_INVALID_ARG;
  if (!b->ops.empty() && (b->exec || b->graph)) b->drop_graph();
  size_t need = 0;
  switch (op.kind) {
    case kOpScan:
    case kOpSegScan:
      // Sized for the smallest tile any candidate uses (1024 elements).
      need = tcs_scan_workspace_bytes(op.n, 32, 1);
      break;
    case kOpSegSum:
      need = tcs_seg_sum_workspace_bytes(op.n);
      break;
    case kOpSpmv:
      need = tcs_spmv_workspace_bytes(op.n, op.rows, op.xv_len, op.x_dtype);
      break;
  }
  size_t off = 0;
  if (!b->ops.empty()) off = align256(b->ops.back().ws_off + b->ops.back().ws_bytes);
  op.ws_off = off;
  op.ws_bytes = align256(need);
  const int key_log = ceil_log2(op.n > 0 ? op.n : 1);
  {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    auto it = g_tuned.find({op.kind, op.x_dtype, key_log});
    if (it != g_tuned.end()) op.cfg = it->second;
  }
  b->ops.push_back(op);
  return (int)b->ops.size() - 1;
}

int check_ws(const tcs_batch* b, void* ws, size_t ws_bytes) {
  if (!ws) return TCS_ERR_WORKSPACE;
  if (ws_bytes < tcs_batch_workspace_bytes(b)) return TCS_ERR_WORKSPACE;
  return TCS_OK;
}

// Counts `k` upcoming launches on ws; re-zeroes the plan's part of it on
// the stream first when the count would reach the limit.
int count_launches(tcs_batch* b, void* ws, uint64_t k, hipStream_t st) {
  if (b->launches + k >= TCS_WORKSPACE_MAX_CALLS) {
    if (int e = tcs_workspace_reset(ws, tcs_batch_workspace_bytes(b), st)) return e;
    b->launches = 0;
  }
  b->launches += k;
  return TCS_OK;
}

}  // namespace

extern "C" {

int tcs_batch_create(tcs_batch** out) {
  if (!out) return TCS_ERR_INVALID_ARG;
  *out = new tcs_batch();
  return TCS_OK;
}

void tcs_batch_destroy(tcs_batch* b) { delete b; }

int tcs_batch_num_ops(const tcs_batch* b) { return b ? (int)b->ops.size() : 0; }

int tcs_batch_add_scan(tcs_batch* b, const void* x, int x_dtype, void* y, int64_t n) {
  if (!x || !y || n <= 0) return -TCS_ERR_INVALID_ARG;
  Op op{};
  op.kind = kOpScan;
  op.x = x;
  op.x_dtype = x_dtype;
  op.y = y;
  op.n = n;
  return add_op(b, op);
}

int tcs_batch_add_seg_scan(tcs_batch* b, const void* x, int x_dtype, const int8_t* flags, void* y,
                           int64_t n) {
  if (!x || !y || !flags || n <= 0) return -TCS_ERR_INVALID_ARG;
  Op op{};
  op.kind = kOpSegScan;
  op.x = x;
  op.x_dtype = x_dtype;
  op.aux = flags;
  op.y = y;
  op.n = n;
  return add_op(b, op);
}

int tcs_batch_add_seg_sum(tcs_batch* b, const void* x, int x_dtype, const void* ends, int ends_dtype,
                          int64_t num_segments, int64_t num_ends, void* y, int64_t nnz) {
  if (!x || !y || !ends || num_segments <= 0 || num_ends < 0 || num_ends > num_segments || nnz < 0)
    return -TCS_ERR_INVALID_ARG;
  Op op{};
  op.kind = kOpSegSum;
  op.x = x;
  op.x_dtype = x_dtype;
  op.aux = ends;
  op.aux_dtype = ends_dtype;
  op.rows = num_segments;
  op.num_ends = num_ends;
  op.y = y;
  op.n = nnz;
  return add_op(b, op);
}

int tcs_batch_add_spmv(tcs_batch* b, const void* vals, int vals_dtype, const void* indptr_ends,
                       int ends_dtype, const int32_t* cols, const void* x, int64_t x_len, float* y,
                       int64_t num_rows, int64_t nnz) {
  if (!vals || !indptr_ends || !cols || !x || !y || num_rows <= 0 || nnz < 0)
    return -TCS_ERR_INVALID_ARG;
  Op op{};
  op.kind = kOpSpmv;
  op.x = vals;
  op.x_dtype = vals_dtype;
  op.aux = indptr_ends;
  op.aux_dtype = ends_dtype;
  op.cols = cols;
  op.xv = x;
  op.xv_len = x_len;
  op.y = y;
  op.rows = num_rows;
  op.num_ends = num_rows;
  op.n = nnz;
  return add_op(b, op);
}

size_t tcs_batch_workspace_bytes(const tcs_batch* b) {
  if (!b || b->ops.empty()) return tcs::kHeaderBytes;
  return align256(b->ops.back().ws_off + b->ops.back().ws_bytes);
}

int tcs_batch_run_eager(tcs_batch* b, void* ws, size_t ws_bytes, void* stream) {
  if (!b) return TCS_ERR_INVALID_ARG;
  if (int e = check_ws(b, ws, ws_bytes)) return e;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (int e = count_launches(b, ws, 1, st)) return e;
  for (const Op& op : b->ops)
    if (int e = run_op(op, op.cfg, static_cast<char*>(ws), st)) return e;
  return TCS_OK;
}

int tcs_batch_tune(tcs_batch* b, void* ws, size_t ws_bytes, int reps, int force, void* stream) {
  if (!b || reps < 1) return TCS_ERR_INVALID_ARG;
  if (int e = check_ws(b, ws, ws_bytes)) return e;
  hipStream_t st = static_cast<hipStream_t>(stream);
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) return TCS_ERR_LAUNCH;
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return TCS_ERR_LAUNCH;
  }
  int rc = TCS_OK;
  for (Op& op : b->ops) {
    if (!is_scan(op.kind)) continue;
    const auto key = std::make_tuple((int)op.kind, op.x_dtype, ceil_log2(op.n));
    if (!force) {
      std::lock_guard<std::mutex> lock(g_tune_mu);
      auto it = g_tuned.find(key);
      if (it != g_tuned.end()) {
        op.cfg = it->second;
        continue;
      }
    }
    ScanCfg best;
    for (ScanCfg c : candidates(op)) {
      if ((rc = count_launches(b, ws, (uint64_t)reps + 1, st))) break;
      if ((rc = run_op(op, c, static_cast<char*>(ws), st))) break;  // warm-up
      if (hipEventRecord(e0, st) != hipSuccess) { rc = TCS_ERR_LAUNCH; break; }
      for (int r = 0; r < reps && !rc; ++r) rc = run_op(op, c, static_cast<char*>(ws), st);
      if (rc) break;
      if (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
        rc = TCS_ERR_LAUNCH;
        break;
      }
      float ms = 0.0f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      c.ms = ms / reps;
      if (best.ms < 0.0f || c.ms < best.ms) best = c;
    }
    if (rc) break;
    op.cfg = best;
    std::lock_guard<std::mutex> lock(g_tune_mu);
    g_tuned[key] = best;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (!rc) b->drop_graph();
  return rc;
}

int tcs_batch_tuned_count(void) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  return (int)g_tuned.size();
}

int tcs_batch_tuned_entry(int i, int* kind, int* dtype, int* log2n, int* S, int* tiles_per_wg, int* method,
                          int* window, float* ms) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  if (i < 0 || i >= (int)g_tuned.size()) return TCS_ERR_INVALID_ARG;
  auto it = g_tuned.begin();
  std::advance(it, i);
  if (kind) *kind = std::get<0>(it->first);
  if (dtype) *dtype = std::get<1>(it->first);
  if (log2n) *log2n = std::get<2>(it->first);
  if (S) *S = it->second.S;
  if (tiles_per_wg) *tiles_per_wg = it->second.tiles_per_wg;
  if (method) *method = it->second.method;
  if (window) *window = it->second.window;
  if (ms) *ms = it->second.ms;
  return TCS_OK;
}

int tcs_batch_tuned_set(int kind, int dtype, int log2n, int S, int tiles_per_wg, int method, int window,
                        float ms) {
  if ((kind != kOpScan && kind != kOpSegScan) || log2n < 0 || log2n > 62) return TCS_ERR_INVALID_ARG;
  if (dtype != TCS_F16 && dtype != TCS_BF16 && dtype != TCS_F32 && dtype != TCS_I8 && dtype != TCS_I32)
    return TCS_ERR_UNSUPPORTED_DTYPE;
  if ((S != 32 && S != 64 && S != 128) || tiles_per_wg < 1 || tiles_per_wg > 32 ||
      (tiles_per_wg & (tiles_per_wg - 1)) != 0 || window < 1 || window > tcs::kWave)
    return TCS_ERR_INVALID_ARG;
  if (method != TCS_METHOD_DPP && method != TCS_METHOD_MFMA) return TCS_ERR_INVALID_ARG;
  if (method == TCS_METHOD_MFMA && (kind != kOpScan || (dtype != TCS_F16 && dtype != TCS_BF16 && dtype != TCS_I8)))
    return TCS_ERR_UNSUPPORTED_DTYPE;
  ScanCfg c;
  c.S = S;
  c.tiles_per_wg = tiles_per_wg;
  c.method = method;
  c.window = window;
  c.ms = ms;
  std::lock_guard<std::mutex> lock(g_tune_mu);
  g_tuned[std::make_tuple(kind, dtype, log2n)] = c;
  return TCS_OK;
}

int tcs_batch_op_config(const tcs_batch* b, int op, int* S, int* tiles_per_wg, int* method, int* window,
                        float* ms) {
  if (!b || op < 0 || op >= (int)b->ops.size()) return TCS_ERR_INVALID_ARG;
  const ScanCfg& c = b->ops[op].cfg;
  if (S) *S = c.S;
  if (tiles_per_wg) *tiles_per_wg = c.tiles_per_wg;
  if (method) *method = c.method;
  if (window) *window = c.window;
  if (ms) *ms = c.ms;
  return TCS_OK;
}

int tcs_batch_instantiate(tcs_batch* b, int max_branches, void* ws, size_t ws_bytes, void* stream) {
  if (!b || b->ops.empty() || max_branches < 1) return TCS_ERR_INVALID_ARG;
  if (int e = check_ws(b, ws, ws_bytes)) return e;
  b->drop_graph();
  // Work already queued on the caller's stream (e.g. zeroing the
  // workspace) must finish before the graph is first launched; launches
  // are ordered on the caller's stream, so a sync here is enough.
  if (hipStreamSynchronize(static_cast<hipStream_t>(stream)) != hipSuccess) return TCS_ERR_LAUNCH;
  if (!b->origin && hipStreamCreateWithFlags(&b->origin, hipStreamNonBlocking) != hipSuccess)
    return TCS_ERR_LAUNCH;
  hipStream_t st = b->origin;
  const int nb = std::min<int>(max_branches, (int)b->ops.size());
  while ((int)b->branches.size() < nb) {
    hipStream_t s;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return TCS_ERR_LAUNCH;
    b->branches.push_back(s);
  }
  while ((int)b->events.size() < nb + 1) {
    hipEvent_t e;
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return TCS_ERR_LAUNCH;
    b->events.push_back(e);
  }
  if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) return TCS_ERR_LAUNCH;
  int rc = TCS_OK;
  // Fork: every branch waits on the origin, ops are dealt to branches
  // round-robin, and the origin joins every branch.
  if (hipEventRecord(b->events[nb], st) != hipSuccess) rc = TCS_ERR_LAUNCH;
  for (int k = 0; k < nb && !rc; ++k)
    if (hipStreamWaitEvent(b->branches[k], b->events[nb], 0) != hipSuccess) rc = TCS_ERR_LAUNCH;
  for (size_t i = 0; i < b->ops.size() && !rc; ++i)
    rc = run_op(b->ops[i], b->ops[i].cfg, static_cast<char*>(ws), b->branches[i % nb]);
  for (int k = 0; k < nb && !rc; ++k) {
    if (hipEventRecord(b->events[k], b->branches[k]) != hipSuccess ||
        hipStreamWaitEvent(st, b->events[k], 0) != hipSuccess)
      rc = TCS_ERR_LAUNCH;
  }
  hipGraph_t g = nullptr;
  const hipError_t ce = hipStreamEndCapture(st, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (ce != hipSuccess || !g) return TCS_ERR_LAUNCH;
  hipGraphExec_t ex = nullptr;
  if (hipGraphInstantiate(&ex, g, nullptr, nullptr, 0) != hipSuccess) {
    (void)hipGraphDestroy(g);
    return TCS_ERR_LAUNCH;
  }
  b->graph = g;
  b->exec = ex;
  b->ws = ws;
  b->ws_bytes = ws_bytes;
  return TCS_OK;
}

int tcs_batch_launch(tcs_batch* b, void* stream) {
  if (!b || !b->exec) return TCS_ERR_INVALID_ARG;
  if (int e = count_launches(b, b->ws, 1, static_cast<hipStream_t>(stream))) return e;
  if (hipGraphLaunch(b->exec, static_cast<hipStream_t>(stream)) != hipSuccess) return TCS_ERR_LAUNCH;
  return TCS_OK;
}

}  // extern "C"
