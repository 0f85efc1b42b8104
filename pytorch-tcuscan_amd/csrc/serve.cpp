// Serving loop: continuous batching of mixed scan, seg-sum, SpMV and
// compaction requests over device-resident inputs, with admission under an
// HBM budget.  No counterpart in the reference (its wrappers block per call).
//
// Inputs are registered once as pool entries (device tensors the server
// keeps).  run() replays an open-loop schedule of (arrival time, entry)
// requests: a dispatcher loop on the calling thread releases each request
// at its arrival time, admits queued requests in FIFO order while their
// outputs fit in the server's HBM arena (first fit; a request that does not
// fit waits for completions to free space), and launches every request
// admitted in one pass as one batch on the least-busy of its HIP streams,
// so requests that arrive while others run are admitted as memory frees up
// and their work overlaps on the device.  A batch's completion event frees
// its outputs and stamps its requests' completion times; latency is
// completion minus arrival.  Each stream owns a zeroed look-back workspace
// (ops on one stream run in order, so they share it).
#include <torch/extension.h>

#include <ATen/hip/impl/HIPGuardImplMasqueradingAsCUDA.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <deque>
#include <map>
#include <thread>
#include <vector>

#include "../../include/tcuscan.h"

namespace tcuscan {
namespace {

int dcode(at::ScalarType t) {
  switch (t) {
    case at::kHalf: return TCS_F16;
    case at::kBFloat16: return TCS_BF16;
    case at::kFloat: return TCS_F32;
    case at::kChar: return TCS_I8;
    case at::kInt: return TCS_I32;
    case at::kLong: return TCS_I64;
    default: return -1;
  }
}

size_t up256(size_t b) { return (b + 255) & ~(size_t)255; }

void hip_ok(hipError_t e, const char* what) {
  TORCH_CHECK(e == hipSuccess, "Server: ", what, ": ", hipGetErrorString(e));
}

// First-fit allocator over [0, size) in 256-byte units, coalescing frees.
class Arena {
 public:
  explicit Arena(size_t size) { free_[0] = size; }
  int64_t alloc(size_t bytes) {
    bytes = up256(bytes);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= bytes) {
        const size_t off = it->first, rest = it->second - bytes;
        free_.erase(it);
        if (rest) free_[off + bytes] = rest;
        used_ += bytes;
        return (int64_t)off;
      }
    }
    return -1;
  }
  void release(size_t off, size_t bytes) {
    bytes = up256(bytes);
    used_ -= bytes;
    auto it = free_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  size_t used() const { return used_; }

 private:
  std::map<size_t, size_t> free_;
  size_t used_ = 0;
};

enum Kind { kScan = 0, kSegSum = 1, kSpmv = 2, kCompress = 3 };

struct Entry {
  Kind kind;
  at::Tensor a, b, c, d;  // scan: x | seg sum: x, indptr | SpMV: vals, indptr, cols, x | compress: x, mask
  int64_t n = 0, rows = 0;
  at::ScalarType out_dtype = at::kFloat;
  int64_t out_elems = 0;
  size_t out_bytes = 0;   // arena bytes of one request (outputs, + count for compress)
  size_t ws_bytes = 0;
  double bytes_moved = 0; // algorithmic HBM bytes of one request
};

}  // namespace

class Server {
 public:
  Server(int64_t budget_bytes, int num_streams, int max_batch, int device)
      : budget_((size_t)budget_bytes), nstreams_(num_streams), max_batch_(max_batch), device_(device) {
    TORCH_CHECK(budget_bytes > 0 && num_streams >= 1 && num_streams <= 16 && max_batch >= 1,
                "Server: bad budget, stream count or batch size");
    const at::hip::HIPGuardMasqueradingAsCUDA guard(at::Device(at::kCUDA, device_));
    arena_t_ = at::empty({budget_bytes}, at::TensorOptions().dtype(at::kByte).device(at::kCUDA, device_));
    for (int k = 0; k < nstreams_; ++k) {
      hipStream_t s;
      hip_ok(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "stream");
      streams_.push_back(s);
    }
  }
  ~Server() {
    for (hipStream_t s : streams_) (void)hipStreamDestroy(s);
    for (hipEvent_t e : events_) (void)hipEventDestroy(e);
    if (pinned_) (void)hipHostFree(pinned_);
  }

  int add_scan(const at::Tensor& x) {
    Entry e;
    e.kind = kScan;
    e.a = bind(x, "x");
    const auto t = x.scalar_type();
    TORCH_CHECK(t == at::kHalf || t == at::kChar || t == at::kBFloat16 || t == at::kFloat, "Server.add_scan: dtype");
    e.n = x.numel();
    e.out_dtype = (t == at::kChar) ? at::kInt : at::kFloat;
    e.out_elems = e.n;
    e.out_bytes = (size_t)e.n * 4;
    e.ws_bytes = tcs_scan_workspace_bytes(e.n, 128, 0);
    e.bytes_moved = (double)e.n * (x.element_size() + 4);
    return add(std::move(e));
  }

  int add_seg_sum(const at::Tensor& x, const at::Tensor& indptr) {
    Entry e;
    e.kind = kSegSum;
    e.a = bind(x, "x");
    e.b = bind(indptr, "indptr");
    TORCH_CHECK(indptr.scalar_type() == at::kInt && indptr.numel() >= 2, "Server.add_seg_sum: int32 indptr");
    e.n = x.numel();
    e.rows = indptr.numel() - 1;
    e.out_dtype = x.scalar_type() == at::kChar ? at::kInt : at::kFloat;
    e.out_elems = e.rows;
    e.out_bytes = (size_t)e.rows * 4;
    e.ws_bytes = tcs_seg_sum_workspace_bytes(e.n);
    e.bytes_moved = (double)e.n * x.element_size() + 8.0 * e.rows;
    return add(std::move(e));
  }

  int add_spmv(const at::Tensor& vals, const at::Tensor& indptr, const at::Tensor& cols, const at::Tensor& xv) {
    Entry e;
    e.kind = kSpmv;
    e.a = bind(vals, "vals");
    e.b = bind(indptr, "indptr");
    e.c = bind(cols, "cols");
    e.d = bind(xv, "x");
    TORCH_CHECK(indptr.scalar_type() == at::kInt && cols.scalar_type() == at::kInt, "Server.add_spmv: int32 index");
    e.n = vals.numel();
    e.rows = indptr.numel() - 1;
    e.out_elems = e.rows;
    e.out_bytes = (size_t)e.rows * 4;
    e.ws_bytes = tcs_spmv_workspace_bytes(e.n, e.rows, xv.numel(), dcode(vals.scalar_type()));
    e.bytes_moved = (double)e.n * (vals.element_size() + 4) + 8.0 * e.rows + (double)xv.numel() * xv.element_size();
    return add(std::move(e));
  }

  int add_compress(const at::Tensor& x, const at::Tensor& mask) {
    Entry e;
    e.kind = kCompress;
    e.a = bind(x, "x");
    e.b = bind(mask, "mask");
    TORCH_CHECK(mask.scalar_type() == at::kChar && mask.numel() == x.numel(), "Server.add_compress: int8 mask");
    e.n = x.numel();
    e.out_dtype = x.scalar_type();
    e.out_elems = e.n;  // capacity; the kept count comes back with the result
    e.out_bytes = up256((size_t)e.n * x.element_size()) + 256;
    e.ws_bytes = tcs_compress_workspace_bytes(e.n);
    e.bytes_moved = (double)e.n * (x.element_size() + 2);  // x + mask (twice) read, about half written
    return add(std::move(e));
  }

  // Runs an open-loop schedule.  arrivals_s must be non-decreasing; keep
  // lists request indices whose outputs are kept (copied out) for checks.
  py::dict run(const std::vector<double>& arrivals_s, const std::vector<int64_t>& entries,
               const std::vector<int64_t>& keep) {
    const int64_t N = (int64_t)arrivals_s.size();
    TORCH_CHECK((int64_t)entries.size() == N, "Server.run: one entry per arrival");
    for (int64_t i = 0; i < N; ++i) {
      TORCH_CHECK(entries[i] >= 0 && entries[i] < (int64_t)pool_.size(), "Server.run: bad entry");
      TORCH_CHECK(i == 0 || arrivals_s[i] >= arrivals_s[i - 1], "Server.run: arrivals must be sorted");
    }
    const at::hip::HIPGuardMasqueradingAsCUDA guard(at::Device(at::kCUDA, device_));
    prepare(N);
    kept_.clear();
    std::vector<char> keep_mask(N, 0);
    for (int64_t r : keep) {
      TORCH_CHECK(r >= 0 && r < N, "Server.run: bad keep index");
      keep_mask[r] = 1;
      const Entry& e = pool_[entries[r]];
      kept_[r] = at::empty({e.out_elems}, e.a.options().dtype(e.out_dtype));
    }
    std::vector<double> admit(N, 0.0), done(N, 0.0);
    std::vector<int64_t> off(N, -1);
    struct Flight {
      int stream;
      hipEvent_t ev;
      std::vector<int64_t> reqs;
    };
    std::vector<Flight> flights;
    std::vector<int> busy(nstreams_, 0);
    std::deque<int64_t> queue;
    int64_t next = 0, finished = 0, batches = 0, host_req_n = 0;
    std::vector<double> host_batch_us;
    host_batch_us.reserve((size_t)N);
    size_t peak_used = 0;
    double bytes_total = 0;
    Arena arena(budget_);
    char* base = static_cast<char*>(arena_t_.data_ptr());
    {
      py::gil_scoped_release nogil;
      const auto t0 = std::chrono::steady_clock::now();
      auto now = [&]() { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); };
      while (finished < N) {
        bool progress = false;
        double t = now();
        for (size_t f = 0; f < flights.size();) {
          const hipError_t q = hipEventQuery(flights[f].ev);
          if (q == hipSuccess) {
            for (int64_t r : flights[f].reqs) {
              done[r] = t;
              arena.release((size_t)off[r], pool_[entries[r]].out_bytes);
              ++finished;
            }
            --busy[flights[f].stream];
            free_ev_.push_back(flights[f].ev);
            flights.erase(flights.begin() + f);
            progress = true;
          } else {
            TORCH_CHECK(q == hipErrorNotReady, "Server.run: ", hipGetErrorString(q));
            ++f;
          }
        }
        while (next < N && arrivals_s[next] <= t) queue.push_back(next++);
        std::vector<int64_t> batch;
        while (!queue.empty() && (int)batch.size() < max_batch_) {
          const int64_t r = queue.front();
          const int64_t o = arena.alloc(pool_[entries[r]].out_bytes);
          if (o < 0) break;
          off[r] = o;
          batch.push_back(r);
          queue.pop_front();
        }
        if (!batch.empty()) {
          const double h0 = t;  // the dispatcher's own cost: admission, launches, event record
          peak_used = std::max(peak_used, arena.used());
          int k = 0;
          for (int j = 1; j < nstreams_; ++j)
            if (busy[j] < busy[k]) k = j;
          for (int64_t r : batch) {
            admit[r] = t;
            const Entry& e = pool_[entries[r]];
            launch(e, base + off[r], k, keep_mask[r] ? &kept_[r] : nullptr, r);
            bytes_total += e.bytes_moved;
          }
          hipEvent_t ev = take_event();
          hip_ok(hipEventRecord(ev, streams_[k]), "event record");
          const size_t nb = batch.size();
          flights.push_back({k, ev, std::move(batch)});
          ++busy[k];
          ++batches;
          host_batch_us.push_back((now() - h0) * 1e6);
          host_req_n += (int64_t)nb;
          progress = true;
        }
        if (!progress) {
          if (flights.empty() && queue.empty() && next < N) {
            const double wait = arrivals_s[next] - now();
            if (wait > 200e-6) std::this_thread::sleep_for(std::chrono::duration<double>(wait - 100e-6));
          } else {
            std::this_thread::yield();
          }
        }
      }
    }
    // Kept compaction results carry their kept counts.
    for (auto& kv : kept_) {
      const Entry& e = pool_[entries[kv.first]];
      if (e.kind == kCompress) kv.second = kv.second.narrow(0, 0, (int64_t)pinned_[kv.first]);
    }
    std::vector<double> lat(N);
    for (int64_t i = 0; i < N; ++i) lat[i] = done[i] - arrivals_s[i];
    std::vector<double> sorted = lat;
    std::sort(sorted.begin(), sorted.end());
    auto pct = [&](double p) { return sorted.empty() ? 0.0 : sorted[std::min<int64_t>(N - 1, (int64_t)(p * N))]; };
    double mean = 0, wait_mean = 0;
    for (int64_t i = 0; i < N; ++i) {
      mean += lat[i];
      wait_mean += admit[i] - arrivals_s[i];
    }
    const double span = N ? (*std::max_element(done.begin(), done.end()) - arrivals_s[0]) : 0.0;
    py::dict out;
    out["requests"] = N;
    out["batches"] = batches;
    out["p50_ms"] = pct(0.50) * 1e3;
    out["p90_ms"] = pct(0.90) * 1e3;
    out["p99_ms"] = pct(0.99) * 1e3;
    out["max_ms"] = N ? sorted.back() * 1e3 : 0.0;
    out["mean_ms"] = N ? mean / N * 1e3 : 0.0;
    out["mean_queue_ms"] = N ? wait_mean / N * 1e3 : 0.0;
    out["span_s"] = span;
    out["throughput_rps"] = span > 0 ? N / span : 0.0;
    out["gbs"] = span > 0 ? bytes_total / span / 1e9 : 0.0;
    out["peak_arena_bytes"] = (int64_t)peak_used;
    // Host time of the dispatcher per launched batch (admission, launches,
    // event record), which every request of the batch waits for.
    std::vector<double> hb = host_batch_us;
    std::sort(hb.begin(), hb.end());
    double hsum = 0;
    for (double v : hb) hsum += v;
    out["host_us_per_batch_mean"] = hb.empty() ? 0.0 : hsum / (double)hb.size();
    out["host_us_per_batch_p50"] = hb.empty() ? 0.0 : hb[hb.size() / 2];
    out["host_us_per_batch_p99"] = hb.empty() ? 0.0 : hb[std::min(hb.size() - 1, (size_t)(0.99 * hb.size()))];
    out["host_us_per_request"] = host_req_n ? hsum / (double)host_req_n : 0.0;
    out["host_busy_frac"] = span > 0 ? hsum * 1e-6 / span : 0.0;
    out["latency_s"] = lat;
    return out;
  }

  at::Tensor kept(int64_t r) const {
    auto it = kept_.find(r);
    TORCH_CHECK(it != kept_.end(), "Server.kept: request ", r, " was not kept");
    return it->second;
  }

  // The result of one entry processed alone on stream 0 (synchronises).
  at::Tensor run_alone(int64_t entry) {
    TORCH_CHECK(entry >= 0 && entry < (int64_t)pool_.size(), "Server.run_alone: bad entry");
    const at::hip::HIPGuardMasqueradingAsCUDA guard(at::Device(at::kCUDA, device_));
    prepare(1);
    const Entry& e = pool_[entry];
    at::Tensor out = at::empty({e.out_elems}, e.a.options().dtype(e.out_dtype));
    Arena arena(budget_);
    const int64_t o = arena.alloc(e.out_bytes);
    launch(e, static_cast<char*>(arena_t_.data_ptr()) + o, 0, &out, 0);
    hip_ok(hipStreamSynchronize(streams_[0]), "sync");
    if (e.kind == kCompress) out = out.narrow(0, 0, (int64_t)pinned_[0]);
    return out;
  }

  int64_t pool_size() const { return (int64_t)pool_.size(); }
  double entry_bytes(int64_t i) const { return pool_.at(i).bytes_moved; }

 private:
  at::Tensor bind(const at::Tensor& t, const char* name) {
    TORCH_CHECK(t.is_cuda() && t.device().index() == device_, "Server: ", name, " must be on the server's GPU");
    TORCH_CHECK(t.is_contiguous() && (reinterpret_cast<uintptr_t>(t.data_ptr()) & 15u) == 0, "Server: ", name,
                " must be contiguous and 16-byte aligned");
    return t;
  }
  int add(Entry&& e) {
    TORCH_CHECK(e.out_bytes <= budget_, "Server: a request's outputs exceed the HBM budget");
    pool_.push_back(std::move(e));
    return (int)pool_.size() - 1;
  }
  void prepare(int64_t n_req) {
    size_t need = tcs_scan_workspace_bytes(1, 128, 0);  // at least the workspace header
    for (const Entry& e : pool_) need = std::max(need, up256(e.ws_bytes));
    if (ws_.size() != streams_.size() || ws_bytes_ < need) {
      ws_.clear();
      for (size_t k = 0; k < streams_.size(); ++k)
        ws_.push_back(at::zeros({(int64_t)need}, at::TensorOptions().dtype(at::kByte).device(at::kCUDA, device_)));
      ws_bytes_ = need;
      ws_calls_.assign(streams_.size(), 0);
      hip_ok(hipDeviceSynchronize(), "workspace");
    }
    if (pinned_n_ < n_req) {
      if (pinned_) (void)hipHostFree(pinned_);
      pinned_ = nullptr;
      hip_ok(hipHostMalloc(reinterpret_cast<void**>(&pinned_), (size_t)n_req * sizeof(int64_t), 0), "pinned");
      pinned_n_ = n_req;
    }
  }
  hipEvent_t take_event() {
    if (!free_ev_.empty()) {
      hipEvent_t e = free_ev_.back();
      free_ev_.pop_back();
      return e;
    }
    hipEvent_t e;
    hip_ok(hipEventCreateWithFlags(&e, hipEventDisableTiming), "event");
    events_.push_back(e);
    return e;
  }
  void launch(const Entry& e, char* out, int k, at::Tensor* keep, int64_t slot) {
    hipStream_t st = streams_[k];
    void* ws = ws_[k].data_ptr();
    const size_t wb = ws_bytes_;
    int rc = TCS_OK;
    // Launch tags repeat after 2^30 - 1 look-back launches on a workspace:
    // re-zero it (stream-ordered) every TCS_WORKSPACE_MAX_CALLS requests.
    if (++ws_calls_[k] >= TCS_WORKSPACE_MAX_CALLS) {
      rc = tcs_workspace_reset(ws, wb, st);
      TORCH_CHECK(rc == TCS_OK, "Server: workspace reset failed: ", tcs_status_string(rc));
      ws_calls_[k] = 0;
    }
    switch (e.kind) {
      case kScan:
        rc = tcs_scan(e.a.data_ptr(), dcode(e.a.scalar_type()), out, e.n, 128, 0, 0.0, ws, wb, st);
        break;
      case kSegSum:
        rc = tcs_seg_sum_indptr(e.a.data_ptr(), dcode(e.a.scalar_type()),
                                static_cast<const int32_t*>(e.b.data_ptr()) + 1, TCS_I32, e.rows, e.rows, out,
                                e.n, ws, wb, st);
        break;
      case kSpmv:
        rc = tcs_spmv_csr(e.a.data_ptr(), dcode(e.a.scalar_type()), static_cast<const int32_t*>(e.b.data_ptr()) + 1,
                          TCS_I32, static_cast<const int32_t*>(e.c.data_ptr()), e.d.data_ptr(), e.d.numel(),
                          reinterpret_cast<float*>(out), e.rows, e.n, ws, wb, st);
        break;
      case kCompress: {
        int64_t* cnt = reinterpret_cast<int64_t*>(out + up256((size_t)e.n * e.a.element_size()));
        rc = tcs_count_mask(static_cast<const int8_t*>(e.b.data_ptr()), e.n, cnt, ws, wb, st);
        if (rc == TCS_OK)
          rc = tcs_compress_counted(e.a.data_ptr(), (int)e.a.element_size(), nullptr, 0, 4,
                                    static_cast<const int8_t*>(e.b.data_ptr()), out, nullptr, e.n, e.n, ws, wb, st);
        if (rc == TCS_OK)
          hip_ok(hipMemcpyAsync(pinned_ + slot, cnt, sizeof(int64_t), hipMemcpyDeviceToHost, st), "count copy");
        break;
      }
    }
    TORCH_CHECK(rc == TCS_OK, "Server: launch failed: ", tcs_status_string(rc));
    if (keep)
      hip_ok(hipMemcpyAsync(keep->data_ptr(), out, (size_t)e.out_elems * keep->element_size(),
                            hipMemcpyDeviceToDevice, st),
             "keep copy");
  }

  size_t budget_;
  int nstreams_, max_batch_, device_;
  at::Tensor arena_t_;
  std::vector<hipStream_t> streams_;
  std::vector<hipEvent_t> events_, free_ev_;
  std::vector<at::Tensor> ws_;
  std::vector<uint64_t> ws_calls_;
  size_t ws_bytes_ = 0;
  std::vector<Entry> pool_;
  std::map<int64_t, at::Tensor> kept_;
  int64_t* pinned_ = nullptr;
  int64_t pinned_n_ = 0;
};

void register_server(py::module& m) {
  py::class_<Server>(m, "Server",
                     "Continuous-batching request loop over device-resident inputs (HBM-budget admission)")
      .def(py::init<int64_t, int, int, int>(), py::arg("hbm_budget_bytes"), py::arg("num_streams") = 4,
           py::arg("max_batch") = 32, py::arg("device") = 0)
      .def("add_scan", &Server::add_scan, py::arg("x"))
      .def("add_seg_sum", &Server::add_seg_sum, py::arg("x"), py::arg("indptr"))
      .def("add_spmv", &Server::add_spmv, py::arg("vals"), py::arg("indptr"), py::arg("cols"), py::arg("x"))
      .def("add_compress", &Server::add_compress, py::arg("x"), py::arg("mask"))
      .def("run", &Server::run, py::arg("arrivals_s"), py::arg("entries"), py::arg("keep") = std::vector<int64_t>{})
      .def("kept", &Server::kept, py::arg("request"))
      .def("run_alone", &Server::run_alone, py::arg("entry"))
      .def("pool_size", &Server::pool_size)
      .def("entry_bytes", &Server::entry_bytes, py::arg("entry"));
}

}  // namespace tcuscan
