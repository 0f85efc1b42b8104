"""Benchmark of the tcuscan hot path on one MI355X.

Default workload: the large 1-D scan, N = 2^28 fp16 values uniform in
[-1, 1] (fixed seed) -> float32 inclusive prefix sums through
tcuscan.run_scan_multi_core(x, S=128).  A step is one call over the whole
input with the input already resident in HBM.  The metric is GB/s of
algorithmic traffic, N * (2 + 4) bytes per step (the reference's formula).

Timing.  `value` is the wall-clock rate of `--steps` back-to-back calls
after `--warmup` untimed ones, bracketed by device synchronisations, with
nothing else enqueued between the calls.  Separately, at least 100 calls
are timed one by one with HIP events on the stream the ops launch on (warm
caches), and 100 more after zeroing a 512 MiB buffer before each call
(cold caches: the reference's method); mean, median and p99 of both are
reported.

Printed: one JSON line with the driver's fields plus
  roofline      the scan kernel's average duration over the event-timed
                calls, as achieved GB/s against 8 TB/s, with its median,
                p99 and the cache-flushed distribution;
  cpu_baseline  the oracle's C port of the reference's run_scan_cpu
                (three-phase OpenMP blocked scan), timed on this host over
                whole copies of the same workload; beside it the host (core
                count, thread settings, lscpu model) and host libraries on
                bounded samples: torch.cumsum, torch.masked_select (values
                and positions), np.add.reduceat seg-sums and scipy's CSR
                product on the SpMV matrix;
  extras        the other workloads measured the same way: the N = 2^20
                fp16 scan's per-call latency (run_scan_single_core), the int8 and
                bf16 scans, run_compress_ind on N = 2^28 fp32 values with a
                Bernoulli(0.3) mask, run_filter_greater_eq on N = 2^28 fp16
                normals (pivot 0.5), run_radix_sort and run_topk_fp16
                (k = 2048) on N = 2^28 fp16 normals with torch.sort /
                torch.topk timed beside them, tcuscan.sample_top_p on 2^24
                fp16 probabilities (p = 0.1) beside the same recipe on
                torch.sort and torch.cumsum, run_searchsorted of 2^24 int32 needles
                into 2^24 sorted values, and run_spmv_v2 on the 2^22 x 2^22 CSR
                matrix with lognormal row lengths (mean 16, sigma 1.2,
                clipped to [0, 20000]; about 6.7e7 nonzeros).
`--only NAME` runs one workload for a fixed number of steps without the
JSON line (for rocprofv3); `--gpus N` with N > 1 exits: this build runs on
one GPU.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(REPO, "pytorch-tcuscan_amd")
sys.path.insert(0, PKG)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E peak


FLUSH_BYTES = 512 << 20  # larger than the 256 MB Infinity Cache and the L2s together
_flush_buf = None


def _flush():
    """Zeroes a 512 MiB device buffer: evicts the op's inputs from the L2s
    and the Infinity Cache before a cold-cache timed call (the reference's
    method, scripts/profiling/profile_tcuscan_ops.py:140-150)."""
    import torch

    global _flush_buf
    if _flush_buf is None:
        _flush_buf = torch.empty(FLUSH_BYTES // 4, dtype=torch.float32, device="cuda:0")
    _flush_buf.zero_()


def _spin(fn, spin_s: float):
    """Runs fn for at least `spin_s` seconds so the GPU is at its working
    clocks when timing starts (a few short kernels after idle measure the
    clock ramp)."""
    import torch

    t_end = time.perf_counter() + spin_s
    while time.perf_counter() < t_end:
        for _ in range(4):
            fn()
        torch.cuda.synchronize()


def _timed_steps(fn, steps: int, warmup: int, spin_s: float = 0.25):
    """Wall seconds of `steps` back-to-back calls of fn after `warmup`
    untimed ones, bracketed by device synchronisations; nothing else is
    enqueued between the calls."""
    import torch

    _spin(fn, spin_s)
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def _event_samples(fn, iters: int = 100, flush: bool = False):
    """Per-call durations (ms) of `iters` calls, each bracketed by its own
    HIP events on the current stream (where the ops launch their kernels).
    flush=True zeroes the 512 MiB flush buffer before each call, outside
    the events (cold caches, as the reference times its ops)."""
    import torch

    starts = [torch.cuda.Event(enable_timing=True) for _ in range(iters)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(iters)]
    for i in range(iters):
        if flush:
            _flush()
        starts[i].record()
        fn()
        ends[i].record()
    torch.cuda.synchronize()
    return [s.elapsed_time(e) for s, e in zip(starts, ends)]


def _dist(ms):
    """mean / median / p99 (nearest rank) of a list of durations."""
    v = sorted(ms)
    return {"mean_ms": sum(v) / len(v), "p50_ms": v[len(v) // 2],
            "p99_ms": v[min(len(v) - 1, int(round(0.99 * len(v))) - 1)], "n": len(v)}


def _measure(fn, steps: int, warmup: int, iters: int = 100, flush: bool = True, spin_s: float = 0.25):
    """The timing every workload gets: wall seconds of `steps` back-to-back
    calls, then `iters` event-timed calls with warm caches and, with flush,
    `iters` more after a cache flush each."""
    wall = _timed_steps(fn, steps, warmup, spin_s)
    out = {"wall_s": wall, "steps": steps, "hot": _dist(_event_samples(fn, iters))}
    if flush:
        out["flushed"] = _dist(_event_samples(fn, iters, flush=True))
    return out


def _progress(msg: str):
    """One line on stderr per workload (the JSON result goes to stdout)."""
    print(f"bench.py: {msg}", file=sys.stderr, flush=True)


def _rnd(d, k=4):
    return {a: (round(b, k) if isinstance(b, float) else b) for a, b in d.items()}


def _scan_workload(dtype_name: str, n: int, S: int, steps: int, warmup: int, method: str | None = None):
    """method None: run_scan_multi_core (the shipped local scan); "dpp" or
    "mfma": run_scan_ex with that local scan method."""
    import torch

    import tcuscan

    dev = torch.device("cuda:0")
    g = torch.Generator(device="cpu").manual_seed(0)
    if dtype_name == "int8":
        x = torch.randint(-4, 5, (n,), generator=g, dtype=torch.int8).to(dev)
        in_b, out_b = 1, 4
    else:
        dt = torch.float16 if dtype_name == "fp16" else torch.bfloat16
        x = (torch.rand(n, generator=g) * 2 - 1).to(dt).to(dev)
        in_b, out_b = 2, 4
    torch.cuda.synchronize()
    if method is None:
        fn = lambda: tcuscan.run_scan_multi_core(x, S)  # noqa: E731
    else:
        fn = lambda: tcuscan.run_scan_ex(x, S, 0, method)  # noqa: E731
    m = _measure(fn, steps, warmup, iters=max(100, steps))
    bytes_per_step = n * (in_b + out_b)
    del x
    torch.cuda.empty_cache()
    hot = m["hot"]
    return {
        "bytes_per_step": bytes_per_step,
        "wall_s": m["wall_s"],
        "event_avg_ms": hot["mean_ms"],
        "event_p50_ms": hot["p50_ms"],
        "event_p99_ms": hot["p99_ms"],
        "event_n": hot["n"],
        "flushed": m["flushed"],
        "gbs_wall": bytes_per_step * steps / m["wall_s"] / 1e9,
        "gbs_kernel": bytes_per_step / (hot["mean_ms"] * 1e-3) / 1e9,
    }


def _compress_workload(n: int, S: int, steps: int, warmup: int):
    """run_compress_ind on N fp32 U[0,1) values, int8 Bernoulli(0.3) mask,
    int32 arange indices.  Bytes: 4N + N + 4N read, 8K written."""
    import torch

    import tcuscan

    dev = torch.device("cuda:0")
    g = torch.Generator(device="cpu").manual_seed(0)
    x = torch.rand(n, generator=g).to(dev)
    mask = (torch.rand(n, generator=g) < 0.3).to(torch.int8).to(dev)
    idx = torch.arange(n, dtype=torch.int32, device=dev)
    k = int(mask.sum().item())
    torch.cuda.synchronize()
    m = _measure(lambda: tcuscan.run_compress_ind(x, idx, mask, S), steps, warmup)
    bytes_per_step = 9 * n + 8 * k
    del x, mask, idx
    torch.cuda.empty_cache()
    return {"bytes_per_step": bytes_per_step, "wall_s": m["wall_s"], "event_avg_ms": m["hot"]["mean_ms"],
            "gbs_wall": bytes_per_step * steps / m["wall_s"] / 1e9, "kept": k, "hot": m["hot"],
            "flushed": m["flushed"]}


def _latency_workload(log2n: int = 20, calls: int = 2000):
    """run_scan_single_core(x, 128) on N = 2^20 fp16 U[-1, 1] (seed 0): the
    per-call latency from the host's view (launch, kernel, and the caller
    waiting for the result), p50 / p99 over `calls` synchronised calls, and
    the kernel-only rate from back-to-back calls."""
    import torch

    import tcuscan

    dev = torch.device("cuda:0")
    n = 1 << log2n
    g = torch.Generator(device="cpu").manual_seed(0)
    x = (torch.rand(n, generator=g) * 2 - 1).half().to(dev)
    for _ in range(200):
        tcuscan.run_scan_single_core(x, 128)
    torch.cuda.synchronize()
    lat = []
    for _ in range(calls):
        t0 = time.perf_counter()
        tcuscan.run_scan_single_core(x, 128)
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - t0)
    lat.sort()
    wall = _timed_steps(lambda: tcuscan.run_scan_single_core(x, 128), 500, 50, spin_s=0.1)
    cold = _dist(_event_samples(lambda: tcuscan.run_scan_single_core(x, 128), 200, flush=True))
    hot = _dist(_event_samples(lambda: tcuscan.run_scan_single_core(x, 128), 200))
    return {"p50_us": lat[len(lat) // 2] * 1e6, "p99_us": lat[int(0.99 * len(lat))] * 1e6,
            "back_to_back_us": wall / 500 * 1e6, "gbs_back_to_back": 6 * n * 500 / wall / 1e9,
            "event_hot": hot, "event_flushed": cold}


def _filter_workload(n: int, S: int, steps: int, warmup: int, pivot: float = 0.5):
    """run_filter_greater_eq on N fp16 standard normals, pivot 0.5 (about
    31% kept): a counting pass (2N bytes read, plus one 8-byte read back that
    sizes the output) and a fused compare-and-compact pass (2N read, 2K
    written).  Bytes per step: 4N + 2K."""
    import torch

    import tcuscan

    dev = torch.device("cuda:0")
    g = torch.Generator(device="cpu").manual_seed(0)
    x = torch.randn(n, generator=g).half().to(dev)
    k = int((x >= pivot).sum().item())
    torch.cuda.synchronize()
    m = _measure(lambda: tcuscan.run_filter_greater_eq(x, pivot, S), steps, warmup, flush=False)
    bytes_per_step = 4 * n + 2 * k
    del x
    torch.cuda.empty_cache()
    return {"bytes_per_step": bytes_per_step, "wall_s": m["wall_s"], "event_avg_ms": m["hot"]["mean_ms"],
            "gbs_wall": bytes_per_step * steps / m["wall_s"] / 1e9, "kept": k}


def _split_workload(n: int, S: int, steps: int, warmup: int, p: float = 0.3):
    """run_split of N fp16 standard normals by an int8 mask with P(mask == 1)
    = 0.3: a count pass (N mask bytes read) and a placing pass (N mask bytes
    and 2N bytes of x read, 2N written).  Bytes per step: 6N."""
    import torch

    import tcuscan

    dev = torch.device("cuda:0")
    g = torch.Generator(device="cpu").manual_seed(0)
    x = torch.randn(n, generator=g).half().to(dev)
    m = (torch.rand(n, generator=g) < p).to(torch.int8).to(dev)
    torch.cuda.synchronize()
    r = _measure(lambda: tcuscan.run_split(x, m, S), steps, warmup, flush=False)
    bytes_per_step = 6 * n
    del x, m
    torch.cuda.empty_cache()
    return {"bytes_per_step": bytes_per_step, "wall_s": r["wall_s"], "event_avg_ms": r["hot"]["mean_ms"],
            "gbs_wall": bytes_per_step * steps / r["wall_s"] / 1e9}


def _searchsorted_workload(len_log2: int, needles_log2: int, steps: int, warmup: int):
    """run_searchsorted of 2^needles_log2 random int32 needles into 2^len_log2
    sorted random int32 values; reported as needles per second."""
    import torch

    import tcuscan

    dev = torch.device("cuda:0")
    g = torch.Generator(device="cpu").manual_seed(0)
    a = torch.sort(torch.randint(-2**31, 2**31 - 1, (1 << len_log2,), generator=g, dtype=torch.int32)).values.to(dev)
    v = torch.randint(-2**31, 2**31 - 1, (1 << needles_log2,), generator=g, dtype=torch.int32).to(dev)
    torch.cuda.synchronize()
    wall = _timed_steps(lambda: tcuscan.run_searchsorted(a, v), steps, warmup)
    del a, v
    torch.cuda.empty_cache()
    return {"wall_s": wall,
            "gneedles_per_s": (1 << needles_log2) * steps / wall / 1e9}


def _sort_workload(n: int, dtype_name: str, steps: int, warmup: int):
    """run_radix_sort on N random keys (fp16 standard normals or int32 over
    the full range) with int32 positions; torch.sort(stable=True) of the
    same keys timed beside it.  Reported in G keys/s."""
    import torch

    import tcuscan

    dev = torch.device("cuda:0")
    g = torch.Generator(device="cpu").manual_seed(0)
    if dtype_name == "fp16":
        x = torch.randn(n, generator=g).half().to(dev)
    else:
        x = torch.randint(-2**31, 2**31 - 1, (n,), generator=g, dtype=torch.int32).to(dev)
    torch.cuda.synchronize()
    wall = _timed_steps(lambda: tcuscan.run_radix_sort(x, 128), steps, warmup)
    twall = _timed_steps(lambda: torch.sort(x, stable=True), max(3, steps // 2), 2, spin_s=0.1)
    tsteps = max(3, steps // 2)
    del x
    torch.cuda.empty_cache()
    return {"wall_s": wall, "gkeys_per_s": n * steps / wall / 1e9,
            "torch_sort_ms": twall * 1e3 / tsteps, "torch_gkeys_per_s": n * tsteps / twall / 1e9}


def _top_p_workload(n: int, steps: int, warmup: int, p: float = 0.1):
    """tcuscan.sample_top_p on N fp16 probabilities uniform in [0, 1)
    (p = 0.1, one draw), the reference's top-p benchmark input, with the same
    recipe on torch.sort(descending) and torch.cumsum timed beside it."""
    import torch

    import tcuscan

    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev).manual_seed(0)
    probs = torch.rand(n, generator=g, device=dev).half()

    def torch_top_p():
        srt, idx = torch.sort(probs, descending=True)
        mass = torch.cumsum(srt, 0, dtype=torch.float32)
        kept = srt.float()
        kept = torch.where(mass - kept > p, torch.zeros_like(kept), kept)
        kept = kept / kept.sum()
        return idx.gather(0, torch.multinomial(kept, 1))

    torch.cuda.synchronize()
    wall = _timed_steps(lambda: tcuscan.sample_top_p(probs, p), steps, warmup)
    twall = _timed_steps(torch_top_p, steps, warmup, spin_s=0.1)
    del probs
    torch.cuda.empty_cache()
    return {"wall_s": wall, "ms_per_call": wall * 1e3 / steps,
            "torch_top_p_ms": twall * 1e3 / steps}


def _topk_workload(n: int, k: int, steps: int, warmup: int):
    """run_topk_fp16 of the k largest of N fp16 standard normals (values and
    positions, sorted), with torch.topk of the same keys timed beside it."""
    import torch

    import tcuscan

    dev = torch.device("cuda:0")
    g = torch.Generator(device="cpu").manual_seed(0)
    x = torch.randn(n, generator=g).half().to(dev)
    torch.cuda.synchronize()
    wall = _timed_steps(lambda: tcuscan.run_topk_fp16(x, k, -10.0, 10.0, 128), steps, warmup)
    twall = _timed_steps(lambda: torch.topk(x, k), steps, warmup, spin_s=0.1)
    del x
    torch.cuda.empty_cache()
    return {"wall_s": wall, "ms_per_call": wall * 1e3 / steps,
            "gbs_read_once": 2 * n * steps / wall / 1e9, "torch_topk_ms": twall * 1e3 / steps}


def spmv_matrix(rows: int, seed: int = 0):
    """The SpMV workload's matrix on the GPU: rows = cols, lognormal row
    lengths (mean 16, sigma 1.2, clipped to [0, 20000]), columns uniform and
    sorted within each row, fp16 values and x uniform in [-1, 1]."""
    import math

    import numpy as np
    import torch

    dev = torch.device("cuda:0")
    rng = np.random.default_rng(seed)
    mu = math.log(16.0) - 1.2 * 1.2 / 2.0
    lens = np.clip(np.rint(rng.lognormal(mu, 1.2, size=rows)), 0, 20000).astype(np.int64)
    indptr = torch.from_numpy(np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)).to(dev)
    nnz = int(lens.sum())
    g = torch.Generator(device=dev).manual_seed(seed)
    row_id = torch.repeat_interleave(torch.arange(rows, device=dev), torch.from_numpy(lens).to(dev))
    cols = torch.randint(0, rows, (nnz,), generator=g, device=dev)
    key, _ = torch.sort(row_id * rows + cols)
    cols = (key % rows).to(torch.int32)
    del key, row_id
    vals = (torch.rand(nnz, generator=g, device=dev) * 2 - 1).half()
    x = (torch.rand(rows, generator=g, device=dev) * 2 - 1).half()
    return vals, indptr, cols, x, nnz


def _gather_ceiling(cols, x, parts: int, iters: int = 100):
    """tcs_calib_gather_f16 over the SpMV matrix's own column indices and
    vector: the random gathers alone (no products, rows or look-back), with
    the gather split across the XCDs (parts = 2) or not; mean ms per call
    over event-timed calls with warm caches."""
    import torch

    lib = ctypes.CDLL(os.path.join(PKG, "libtcuscan_hip.so"))
    lib.tcs_calib_gather_out_elems.restype = ctypes.c_size_t
    lib.tcs_calib_gather_out_elems.argtypes = [ctypes.c_int64, ctypes.c_int]
    lib.tcs_calib_gather_f16.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64,
                                         ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    n = cols.numel()
    out = torch.empty(lib.tcs_calib_gather_out_elems(n, parts), dtype=torch.float32, device=cols.device)

    def call():
        rc = lib.tcs_calib_gather_f16(cols.data_ptr(), n, x.data_ptr(), x.numel(), parts, out.data_ptr(),
                                      ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
        assert rc == 0, rc

    _spin(call, 0.1)
    return _dist(_event_samples(call, iters))["mean_ms"]


def _stream_ceilings(gib: int = 2, iters: int = 50):
    """tcs_calib_stream over `gib` GiB buffers: the read, write and copy
    rates of HBM with nontemporal 16-byte accesses (GB/s, mean over
    event-timed calls), the practical ceilings for streaming kernels."""
    import torch

    lib = ctypes.CDLL(os.path.join(PKG, "libtcuscan_hip.so"))
    lib.tcs_calib_stream.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int,
                                     ctypes.c_void_p]
    nbytes = gib << 30
    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    out = {}
    for mode, name, moved in ((0, "read", 1), (1, "write", 1), (2, "copy", 2)):
        def call(mode=mode):
            rc = lib.tcs_calib_stream(a.data_ptr(), b.data_ptr(), nbytes, mode,
                                      ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
            assert rc == 0, rc
        _spin(call, 0.1)
        ms = _dist(_event_samples(call, iters))["mean_ms"]
        out[name] = round(moved * nbytes / (ms * 1e-3) / 1e9, 1)
    del a, b
    torch.cuda.empty_cache()
    return out


def _spmv_workload(rows: int, steps: int, warmup: int):
    import torch

    import tcuscan

    vals, indptr, cols, x, nnz = spmv_matrix(rows)
    torch.cuda.synchronize()
    m = _measure(lambda: tcuscan.run_spmv_v2(vals, indptr, cols, x, 128), steps, warmup)
    per = m["wall_s"] / steps
    min_bytes = nnz * (2 + 4) + 4 * (rows + 1) + 2 * rows + 4 * rows
    out = {"nnz": nnz, "wall_s": m["wall_s"], "event_avg_ms": m["hot"]["mean_ms"],
           "gflops": 2 * nnz / per / 1e9, "gbs_min_bytes": min_bytes / per / 1e9,
           "hot": m["hot"], "flushed": m["flushed"],
           "gflops_flushed_p50": 2 * nnz / (m["flushed"]["p50_ms"] * 1e-3) / 1e9,
           "gather_ceiling_ms": {"one_part": _gather_ceiling(cols, x, 1), "split": _gather_ceiling(cols, x, 2)}}
    del vals, indptr, cols, x
    torch.cuda.empty_cache()
    return out


def _segsum_workload(nnz_log2: int, steps: int, warmup: int, dtype_name: str = "int8"):
    """run_seg_sum_multi_core over 2^nnz_log2 values (int8 in [-4, 4] or fp16
    U[-1, 1]) in segments of geometric length (mean 64), full CSR indptr.
    Bytes: the values, plus 4 per row end read and 4 per row written."""
    import numpy as np
    import torch

    import tcuscan

    dev = torch.device("cuda:0")
    rng = np.random.default_rng(0)
    n = 1 << nnz_log2
    lens = rng.geometric(1 / 64, size=int(n / 64 * 1.1))
    ends = np.cumsum(lens)
    ends = ends[ends < n]
    indptr = torch.from_numpy(np.concatenate([[0], ends, [n]]).astype(np.int32)).to(dev)
    g = torch.Generator(device=dev).manual_seed(0)
    if dtype_name == "int8":
        x = torch.randint(-4, 5, (n,), generator=g, device=dev, dtype=torch.int8)
    else:
        x = (torch.rand(n, generator=g, device=dev) * 2 - 1).half()
    rows = indptr.numel() - 1
    torch.cuda.synchronize()
    m = _measure(lambda: tcuscan.run_seg_sum_multi_core(x, indptr, 128), steps, warmup, flush=False)
    bytes_per_step = n * x.element_size() + 8 * rows + 4
    del x, indptr
    torch.cuda.empty_cache()
    return {"nnz": n, "rows": rows, "bytes_per_step": bytes_per_step, "wall_s": m["wall_s"],
            "hot": m["hot"], "gbs": bytes_per_step / (m["hot"]["mean_ms"] * 1e-3) / 1e9}


def _copy_workload(n: int, steps: int, warmup: int):
    """A plain float32 copy of n elements (PMC calibration: 4n bytes read
    and 4n written per launch)."""
    import torch

    dev = torch.device("cuda:0")
    a = torch.rand(n, device=dev)
    b = torch.empty_like(a)
    m = _measure(lambda: b.copy_(a), steps, warmup, flush=False)
    return {"gbs_wall": 8 * n * steps / m["wall_s"] / 1e9, "event_avg_ms": m["hot"]["mean_ms"]}


def _batched_workload(steps: int, warmup: int, num_scans: int = 10_000, num_segsums: int = 2_000,
                      per_call_leg: bool = True, tune_cache: str | None = None, branches: int = 1):
    """The batched uneven workload: `num_scans` fp16 scans with lengths
    log-uniform in [2^8, 2^20], concatenated with int8 start flags, plus
    `num_segsums` indptr seg-sums over int8 data (geometric segment lengths,
    mean 64; 2^U(4, 12) segments per problem).  The batched form is one
    BatchPlan (tuned, then captured as one HIP graph, ops in sequence) launched per step; the
    per-call form launches run_scan_multi_core on every scan and
    run_seg_sum_multi_core on every seg-sum problem, each its own tensor.
    With per_call_leg False only the graph replays run (for profiling);
    with tune_cache the tuner's persisted cache is loaded from (and saved
    to) that file, so a second run does no timing runs."""
    import numpy as np
    import torch

    import tcuscan
    from tcuscan import workloads

    dev = torch.device("cuda:0")
    sc = workloads.uneven_scans(num_scans, 8, 20, seed=0, device=dev)
    ss = workloads.geometric_segsums(num_segsums, 64.0, 4, 12, seed=1, device=dev)
    plan = tcuscan.BatchPlan()
    y_scan = plan.add_seg_scan(sc.x, sc.flags)
    y_sum = plan.add_seg_sum(ss.x, ss.indptr)
    t0 = time.perf_counter()
    loaded = None
    if tune_cache:
        loaded = tcuscan.tuning.load(tune_cache)
        plan.tune(reps=2)
        tcuscan.tuning.save(tune_cache)
    else:
        plan.tune(reps=2)
    tune_s = time.perf_counter() - t0
    plan.instantiate(max_branches=branches)
    mb = _measure(plan.launch, steps, warmup, flush=False)
    wall_b = mb["wall_s"]

    total = int(sc.offsets[-1])
    nnz = int(ss.indptr_host[-1])
    rows = int(ss.row_off[-1])
    bytes_per_step = 7 * total + nnz + 4 * (rows + 1) + 4 * rows
    cfg = plan.configs()
    tuned = {"seg_scan": {"S": cfg[0][0], "tiles_per_wg": cfg[0][1], "method": cfg[0][2], "window": cfg[0][3],
                          "ms": round(cfg[0][4], 4)}}
    ms_b = wall_b * 1e3 / steps
    if not per_call_leg:
        return {"scans": num_scans, "scan_elems": total, "segsums": num_segsums, "segsum_nnz": nnz,
                "bytes_per_step": bytes_per_step, "batched_ms": ms_b, "batched_event": mb["hot"],
                "gbs_batched": bytes_per_step / (ms_b * 1e-3) / 1e9, "tuned": tuned, "tune_s": tune_s,
                "tune_cache_entries_loaded": loaded, "device_key": tcuscan.tuning.device_key()}

    scan_pieces = sc.pieces()
    sum_pieces = ss.pieces()

    def per_call():
        for p in scan_pieces:
            tcuscan.run_scan_multi_core(p, 128)
        for xi, ipi in sum_pieces:
            tcuscan.run_seg_sum_multi_core(xi, ipi, 128)

    pc_steps = max(2, min(steps, 5))
    wall_p = _timed_steps(per_call, pc_steps, 1)
    # Same results both ways (spot check of 50 scans and 50 seg-sums).
    rng = np.random.default_rng(0)
    same = True
    for i in rng.choice(num_scans, 50, replace=False):
        o0, o1 = sc.offsets[i], sc.offsets[i + 1]
        same &= bool(torch.equal(tcuscan.run_scan_multi_core(scan_pieces[i], 128), y_scan[o0:o1]))
    for i in rng.choice(num_segsums, 50, replace=False):
        r0, r1 = ss.row_off[i], ss.row_off[i + 1]
        same &= bool(torch.equal(tcuscan.run_seg_sum_multi_core(*sum_pieces[i], 128), y_sum[r0:r1]))
    del sc, ss, plan, scan_pieces, sum_pieces, y_scan, y_sum
    torch.cuda.empty_cache()
    ms_p = wall_p * 1e3 / pc_steps
    return {
        "scans": num_scans, "scan_elems": total, "segsums": num_segsums, "segsum_nnz": nnz,
        "segsum_rows": rows, "bytes_per_step": bytes_per_step,
        "batched_ms": ms_b, "batched_event_avg_ms": mb["hot"]["mean_ms"],
        "per_call_ms": ms_p, "speedup": ms_p / ms_b,
        "gbs_batched": bytes_per_step / (ms_b * 1e-3) / 1e9,
        "tuned": tuned, "tune_s": tune_s, "same_results": same,
    }


def _serve_workload(seconds: float, rate: float, seed: int = 0):
    """The serving stream: open-loop Poisson arrivals at `rate` req/s for
    `seconds`; mix 40% scans (fp16 and int8, N from 2^10 to 2^26), 30% CSR
    SpMV (fp16, lognormal rows of mean 16 clipped to 20000, nnz from 10^4 to
    10^7, at least 2^16 columns), 20% seg-sums (fp16, geometric segments of
    mean 64, nnz from 10^4 to 10^7), 10% compaction (fp32, N from 2^10 to
    2^24, Bernoulli mask p from 0.05 to 0.9).  Inputs are a device-resident
    pool of distinct requests per kind whose sizes are log-spaced over each
    range, both ends included (requests draw pool entries uniformly);
    admission under a 64 GB HBM budget."""
    import numpy as np
    import torch

    import tcuscan
    from tcuscan import workloads

    dev = torch.device("cuda:0")
    rng = np.random.default_rng(seed)
    srv = tcuscan.Server(64 << 30, 4, 32)
    kinds = {"scan": [], "spmv": [], "segsum": [], "compress": []}
    hold = []
    sizes = {"scan": 48, "spmv": 16, "segsum": 16, "compress": 16}
    for i, lg in enumerate(np.linspace(10, 26, sizes["scan"])):
        n = int(round(2 ** lg))
        g = torch.Generator(device=dev).manual_seed(1000 + i)
        if i % 2 == 0:
            x = (torch.rand(n, generator=g, device=dev) * 2 - 1).half()
        else:
            x = torch.randint(-128, 128, (n,), generator=g, device=dev, dtype=torch.int8)
        hold.append(x)
        kinds["scan"].append(srv.add_scan(x))
    spmv_nnz = []
    for i, lg in enumerate(np.linspace(4, 7, sizes["spmv"])):
        rows = max(16, int(round(10 ** lg / 16)))
        cols = max(rows, 1 << 16)  # room for distinct columns in the longest rows
        m = workloads.lognormal_csr(rows, cols, 16.0, 1.2, 20000, seed=2000 + i, device=dev)
        hold.append(m)
        spmv_nnz.append(int(m[0].numel()))
        kinds["spmv"].append(srv.add_spmv(*m))
    segsum_nnz = []
    for i, lg in enumerate(np.linspace(4, 7, sizes["segsum"])):
        lgs = float(np.log2(max(2.0, 10 ** lg / 64)))
        ss = workloads.geometric_segsums(1, 64.0, lgs, lgs + 1e-6, seed=3000 + i, device=dev)
        x = ss.x.half()
        hold.append((x, ss.indptr))
        segsum_nnz.append(int(x.numel()))
        kinds["segsum"].append(srv.add_seg_sum(x, ss.indptr))
    for i, lg in enumerate(np.linspace(10, 24, sizes["compress"])):
        n = int(round(2 ** lg))
        p = float(np.linspace(0.05, 0.9, sizes["compress"])[(i * 7) % sizes["compress"]])
        g = torch.Generator(device=dev).manual_seed(4000 + i)
        x = torch.rand(n, generator=g, device=dev)
        mk = (torch.rand(n, generator=g, device=dev) < p).to(torch.int8)
        hold.append((x, mk))
        kinds["compress"].append(srv.add_compress(x, mk))
    torch.cuda.synchronize()
    n_req = int(rate * seconds)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n_req))
    pick = rng.choice(4, n_req, p=[0.4, 0.3, 0.2, 0.1])
    names = ["scan", "spmv", "segsum", "compress"]
    entries = [int(rng.choice(kinds[names[k]])) for k in pick]
    srv.run(arrivals[: min(200, n_req)].tolist(), entries[: min(200, n_req)])  # warm-up
    st = srv.run(arrivals.tolist(), entries)
    lat = np.asarray(st.pop("latency_s"))
    per_kind = {}
    for k, nm in enumerate(names):
        sel = lat[pick == k]
        if sel.size:
            per_kind[nm] = {"n": int(sel.size), "p50_ms": round(float(np.percentile(sel, 50)) * 1e3, 4),
                            "p99_ms": round(float(np.percentile(sel, 99)) * 1e3, 4)}
    st["per_kind"] = per_kind
    st["offered_rps"] = rate
    st["seconds"] = seconds
    st["pool"] = {"scan_n": [1 << 10, 1 << 26], "spmv_nnz": [min(spmv_nnz), max(spmv_nnz)],
                  "segsum_nnz": [min(segsum_nnz), max(segsum_nnz)], "compress_n": [1 << 10, 1 << 24],
                  "entries": {k: len(v) for k, v in kinds.items()}}
    del srv, hold
    torch.cuda.empty_cache()
    return st


def _cpu_baseline(n: int, budget_s: float = 10.0):
    """Times the oracle's C port of run_scan_cpu on the full fp16 workload."""
    import numpy as np

    lib_path = os.path.join(REPO, "oracle", "libscan_cpu_port.so")
    if not os.path.exists(lib_path):
        import subprocess
        subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle")], check=True)
    lib = ctypes.CDLL(lib_path)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1
    threads = max(1, min(threads, os.cpu_count() or threads))
    x = ((np.random.default_rng(0).random(n, dtype=np.float32) * 2) - 1).astype(np.float16)
    y = np.empty(n, np.float32)
    args = (x.ctypes.data_as(ctypes.c_void_p), y.ctypes.data_as(ctypes.c_void_p), ctypes.c_int64(n), threads)
    lib.port_scan_cpu_f16(*args)  # warm (page-in)
    reps, t0 = 0, time.perf_counter()
    while True:
        lib.port_scan_cpu_f16(*args)
        reps += 1
        el = time.perf_counter() - t0
        if el >= budget_s or reps >= 200:
            break
    per = el / reps
    return {
        "value": round(n * 6 / per / 1e9, 3),
        "unit": "GB/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{reps} full runs of the N=2^{n.bit_length() - 1} fp16->fp32 scan "
                  f"(run_scan_cpu semantics, OpenMP three-phase blocked scan), {per * 1e3:.1f} ms each",
    }


def _host_info():
    """This host as the CPU baselines see it: core count, thread settings
    and the lscpu model line."""
    import subprocess

    import torch

    info = {"os_cpu_count": os.cpu_count(), "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
            "torch_threads": torch.get_num_threads()}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in ("Model name", "CPU(s)", "Thread(s) per core", "Socket(s)"):
                info["lscpu_" + k.strip().lower().replace(" ", "_").replace("(s)", "s")] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return info


def _time_cpu(fn, budget_s: float, max_reps: int = 50):
    fn()  # warm (page-in, thread pool start)
    reps, t0 = 0, time.perf_counter()
    while True:
        fn()
        reps += 1
        el = time.perf_counter() - t0
        if el >= budget_s or reps >= max_reps:
            return el / reps, reps


def _cpu_library_baselines(budget_s: float = 2.0):
    """The host libraries beside the GPU ops, each on a bounded sample of
    its workload: torch.cumsum for the scan, torch.masked_select for
    run_compress_ind, np.add.reduceat for an indptr seg-sum and scipy's CSR
    product (single-threaded) for the SpMV matrix itself."""
    import numpy as np
    import torch

    out = {}
    threads = torch.get_num_threads()
    g = torch.Generator().manual_seed(0)

    n = 1 << 26
    x = (torch.rand(n, generator=g) * 2 - 1).half()
    per, reps = _time_cpu(lambda: torch.cumsum(x, 0, dtype=torch.float32), budget_s)
    out["torch_cumsum_f16_to_f32"] = {"value": round(6 * n / per / 1e9, 3), "unit": "GB/s", "cores": threads,
                                      "sample": f"N=2^26, {reps} runs, {per * 1e3:.1f} ms each"}
    del x

    xf = torch.rand(n, generator=g)
    m = torch.rand(n, generator=g) < 0.3
    idx = torch.arange(n, dtype=torch.int32)
    k = int(m.sum())
    per, reps = _time_cpu(lambda: (torch.masked_select(xf, m), torch.masked_select(idx, m)), budget_s)
    out["torch_masked_select_compress_ind_f32"] = {
        "value": round((9 * n + 8 * k) / per / 1e9, 3), "unit": "GB/s", "cores": threads,
        "sample": f"N=2^26 fp32, Bernoulli(0.3) mask, values and positions, {reps} runs, {per * 1e3:.1f} ms each"}
    del xf, m, idx

    rng = np.random.default_rng(0)
    lens = rng.geometric(1 / 64, size=1 << 20)
    starts = np.concatenate([[0], np.cumsum(lens)[:-1]]).astype(np.int64)
    nnz = int(lens.sum())
    xs = rng.integers(-4, 5, nnz).astype(np.int8)
    per, reps = _time_cpu(lambda: np.add.reduceat(xs, starts, dtype=np.int32), budget_s)
    out["numpy_reduceat_seg_sum_i8"] = {
        "value": round((nnz + 8 * lens.size + 4) / per / 1e9, 3), "unit": "GB/s", "cores": 1,
        "sample": f"{lens.size} segments of geometric length (mean 64), {nnz} int8 values, {reps} runs, "
                  f"{per * 1e3:.1f} ms each"}
    del xs, starts

    try:
        import scipy.sparse as sp
    except ImportError:
        sp = None
    if sp is not None and torch.cuda.is_available():
        vals, indptr, cols, xv, nnz = spmv_matrix(1 << 22)
        rows = indptr.numel() - 1
        A = sp.csr_matrix((vals.float().cpu().numpy(), cols.cpu().numpy(), indptr.cpu().numpy()),
                          shape=(rows, rows))
        xh = xv.float().cpu().numpy()
        del vals, indptr, cols, xv
        torch.cuda.empty_cache()
        per, reps = _time_cpu(lambda: A @ xh, budget_s, max_reps=20)
        out["scipy_csr_spmv_f32"] = {"value": round(2 * nnz / per / 1e9, 3), "unit": "GFLOP/s", "cores": 1,
                                     "sample": f"the 2^22 x 2^22 lognormal matrix ({nnz} nonzeros, values as "
                                               f"float32), {reps} runs, {per * 1e3:.1f} ms each"}
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--log2n", type=int, default=28)
    ap.add_argument("--S", type=int, default=128)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-extras", action="store_true", help="skip the extra workloads")
    ap.add_argument("--serve-seconds", type=float, default=60.0, help="length of the serving stream")
    ap.add_argument("--serve-rate", type=float, default=2000.0, help="Poisson arrival rate, req/s")
    ap.add_argument("--branches", type=int, default=1,
                    help="with --only batched: graph branches of the plan (one measured faster: the persistent "
                         "seg-scan holds every CU, so a seg-sum beside it only finishes after it)")
    ap.add_argument("--tune-cache", default=None,
                    help="with --only batched: load the batch tuner's cache from this JSON file (and save it)")
    ap.add_argument("--only", choices=["scan_f16", "scan_i8", "scan_bf16", "scan_f16_mfma", "scan_i8_mfma",
                                       "scan_f16_dpp", "scan_i8_dpp", "compress_ind", "spmv", "copy_f32",
                                       "batched", "serve", "filter_ge", "searchsorted", "sort_f16", "sort_i32",
                                       "topk_f16", "top_p", "stream_ceilings", "segsum_i8", "segsum_f16",
                                       "split_f16"],
                    help="run one workload and print its timing only (for profilers)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is not supported", file=sys.stderr)
        return 2
    if int(os.environ.get("RANK", "0")) != 0:
        return 0

    n = 1 << a.log2n
    if a.only:
        if a.only.startswith("scan_"):
            parts = a.only.split("_")
            r = _scan_workload({"f16": "fp16", "i8": "int8", "bf16": "bf16"}[parts[1]], n, a.S,
                               a.steps, a.warmup, parts[2] if len(parts) > 2 else None)
        elif a.only == "compress_ind":
            r = _compress_workload(n, a.S, a.steps, a.warmup)
        elif a.only == "spmv":
            r = _spmv_workload(1 << 22, a.steps, a.warmup)
        elif a.only in ("segsum_i8", "segsum_f16"):
            r = _segsum_workload(26, a.steps, a.warmup, "int8" if a.only == "segsum_i8" else "fp16")
        elif a.only == "stream_ceilings":
            r = _stream_ceilings()
        elif a.only == "batched":
            r = _batched_workload(a.steps, a.warmup, per_call_leg=False, tune_cache=a.tune_cache,
                                  branches=a.branches)
        elif a.only == "serve":
            r = _serve_workload(a.serve_seconds, a.serve_rate)
        elif a.only == "filter_ge":
            r = _filter_workload(n, a.S, a.steps, a.warmup)
        elif a.only == "searchsorted":
            r = _searchsorted_workload(24, 24, a.steps, a.warmup)
        elif a.only == "split_f16":
            r = _split_workload(n, a.S, a.steps, a.warmup)
        elif a.only == "topk_f16":
            r = _topk_workload(n, 2048, a.steps, a.warmup)
        elif a.only == "top_p":
            r = _top_p_workload(1 << 24, a.steps, a.warmup)
        elif a.only in ("sort_f16", "sort_i32"):
            r = _sort_workload(n, "fp16" if a.only == "sort_f16" else "int32", a.steps, a.warmup)
        else:
            r = _copy_workload(n, a.steps, a.warmup)
        print(json.dumps({"only": a.only, **{k: (round(v, 4) if isinstance(v, float) else v) for k, v in r.items()}}))
        return 0
    _progress(f"headline fp16 scan N=2^{a.log2n}")
    main_res = _scan_workload("fp16", n, a.S, a.steps, a.warmup)
    extras = {}
    if not a.no_extras:
        for name, method in (("int8", None), ("bf16", None), ("fp16", "mfma"), ("int8", "mfma")):
            _progress(f"scan {name} {method or 'default'}")
            r = _scan_workload(name, n, a.S, max(10, a.steps // 2), max(3, a.warmup // 2), method)
            key = f"scan_{name}_2p{a.log2n}" + (f"_{method}" if method else "")
            extras[key] = {
                "value": round(r["gbs_wall"], 2), "unit": "GB/s",
                "kernel_gbs": round(r["gbs_kernel"], 2),
                "frac_of_peak": round(r["gbs_kernel"] / HBM_PEAK_GBS, 4),
                "ms_per_step": round(r["wall_s"] * 1e3 / max(10, a.steps // 2), 4),
                "event_p50_ms": round(r["event_p50_ms"], 4), "event_p99_ms": round(r["event_p99_ms"], 4),
                "flushed": _rnd(r["flushed"]),
            }
        ksteps = max(10, a.steps // 2)
        _progress("compress_ind")
        c = _compress_workload(n, a.S, ksteps, max(3, a.warmup // 2))
        extras[f"compress_ind_f32_2p{a.log2n}_p0.3"] = {
            "value": round(c["gbs_wall"], 2), "unit": "GB/s",
            "bytes_per_step": c["bytes_per_step"], "kept": c["kept"],
            "ms_per_step": round(c["wall_s"] * 1e3 / ksteps, 4),
            "hot": _rnd(c["hot"]), "flushed": _rnd(c["flushed"]),
            "note": "includes the device-to-host count read that sizes the outputs",
        }
        _progress("N=2^20 latency")
        lt = _latency_workload()
        extras["scan_f16_2p20_single_core_latency"] = {
            "value": round(lt["p50_us"], 2), "unit": "us p50 per synchronised call",
            "p99_us": round(lt["p99_us"], 2), "back_to_back_us": round(lt["back_to_back_us"], 2),
            "gbs_back_to_back": round(lt["gbs_back_to_back"], 1),
            "event_hot": _rnd(lt["event_hot"]), "event_flushed": _rnd(lt["event_flushed"]),
            "note": "N = 2^20 fp16 -> fp32, S = 128; p50/p99 from the host around synchronised calls; "
                    "event_* from HIP events around each call, event_flushed after a 512 MiB cache flush",
        }
        _progress("filter")
        fl = _filter_workload(n, a.S, ksteps, max(3, a.warmup // 2))
        extras[f"filter_ge_f16_2p{a.log2n}_pivot0.5"] = {
            "value": round(fl["gbs_wall"], 2), "unit": "GB/s",
            "bytes_per_step": fl["bytes_per_step"], "kept": fl["kept"],
            "ms_per_step": round(fl["wall_s"] * 1e3 / ksteps, 4),
            "note": "count pass (chunk counts) + counted compare-and-compact pass, with the count read back to size the output",
        }
        _progress("split")
        sp = _split_workload(n, a.S, ksteps, max(3, a.warmup // 2))
        extras[f"split_f16_2p{a.log2n}_p0.3"] = {
            "value": round(sp["gbs_wall"], 2), "unit": "GB/s",
            "bytes_per_step": sp["bytes_per_step"],
            "ms_per_step": round(sp["wall_s"] * 1e3 / ksteps, 4),
            "note": "count pass (chunk counts, K) + counted placing pass; 6 B per element (mask read twice)",
        }
        _progress("radix sort")
        so = _sort_workload(n, "fp16", max(5, ksteps // 2), 2)
        extras[f"radix_sort_f16_2p{a.log2n}"] = {
            "value": round(so["gkeys_per_s"], 2), "unit": "G keys/s",
            "ms_per_step": round(so["wall_s"] * 1e3 / max(5, ksteps // 2), 4),
            "torch_sort_stable_ms": round(so["torch_sort_ms"], 4),
        }
        _progress("top-k")
        tk = _topk_workload(n, 2048, ksteps, 2)
        extras[f"topk_f16_2p{a.log2n}_k2048"] = {
            "value": round(tk["ms_per_call"], 4), "unit": "ms per call",
            "torch_topk_ms": round(tk["torch_topk_ms"], 4),
        }
        _progress("top-p")
        tp = _top_p_workload(1 << 24, ksteps, 2)
        extras["sample_top_p_f16_2p24"] = {
            "value": round(tp["ms_per_call"], 4), "unit": "ms per call",
            "torch_top_p_ms": round(tp["torch_top_p_ms"], 4),
        }
        _progress("searchsorted")
        ss = _searchsorted_workload(24, 24, ksteps, max(3, a.warmup // 2))
        extras["searchsorted_i32_2p24_into_2p24"] = {
            "value": round(ss["gneedles_per_s"], 3), "unit": "G needles/s",
            "ms_per_step": round(ss["wall_s"] * 1e3 / ksteps, 4),
        }
        _progress("seg-sum")
        sg = _segsum_workload(26, ksteps, max(3, a.warmup // 2))
        extras["seg_sum_i8_2p26_geometric64"] = {
            "value": round(sg["gbs"], 1), "unit": "GB/s", "nnz": sg["nnz"], "rows": sg["rows"],
            "ms_per_call": round(sg["hot"]["mean_ms"], 4), "p99_ms": round(sg["hot"]["p99_ms"], 4),
            "note": "run_seg_sum_multi_core, int8 values, geometric segments (mean 64)",
        }
        _progress("spmv")
        sp = _spmv_workload(1 << 22, ksteps, max(3, a.warmup // 2))
        extras["spmv_v2_2p22_lognormal16"] = {
            "value": round(sp["gflops"], 2), "unit": "GFLOP/s", "nnz": sp["nnz"],
            "gbs_min_bytes": round(sp["gbs_min_bytes"], 2),
            "ms_per_step": round(sp["wall_s"] * 1e3 / ksteps, 4),
            "hot": _rnd(sp["hot"]), "flushed": _rnd(sp["flushed"]),
            "gflops_flushed_p50": round(sp["gflops_flushed_p50"], 2),
            "gather_ceiling_ms": _rnd(sp["gather_ceiling_ms"]),
            "frac_of_gather_ceiling": round(min(sp["gather_ceiling_ms"].values()) / sp["hot"]["mean_ms"], 4),
            "note": "gather_ceiling_ms: the matrix's random gathers of x alone (calibration kernel), unsplit "
                    "and split across the XCDs by column half",
        }
        _progress("batched plan")
        bt = _batched_workload(max(5, a.steps // 5), 2)
        extras["batched_uneven_10k_scans_2k_segsums"] = {
            "value": round(bt["speedup"], 2), "unit": "x vs per-call launches",
            "batched_ms": round(bt["batched_ms"], 3), "per_call_ms": round(bt["per_call_ms"], 2),
            "gbs_batched": round(bt["gbs_batched"], 1), "scan_elems": bt["scan_elems"],
            "segsum_nnz": bt["segsum_nnz"], "tuned": bt["tuned"], "same_results": bt["same_results"],
        }
        _progress(f"serving stream ({a.serve_seconds:g} s)")
        sv = _serve_workload(a.serve_seconds, a.serve_rate)
        extras["serve_mixed_poisson"] = {
            "value": round(sv["p99_ms"], 3), "unit": "ms p99 latency",
            "p50_ms": round(sv["p50_ms"], 3), "p99_ms": round(sv["p99_ms"], 3),
            "throughput_rps": round(sv["throughput_rps"], 1), "offered_rps": sv["offered_rps"],
            "gbs": round(sv["gbs"], 2), "requests": sv["requests"], "seconds": sv["seconds"],
            "batches": sv["batches"], "per_kind": sv["per_kind"],
            "peak_arena_bytes": sv["peak_arena_bytes"], "mean_queue_ms": round(sv["mean_queue_ms"], 4),
            "host_us_per_batch": {"mean": round(sv["host_us_per_batch_mean"], 2),
                                  "p50": round(sv["host_us_per_batch_p50"], 2),
                                  "p99": round(sv["host_us_per_batch_p99"], 2)},
            "host_us_per_request": round(sv["host_us_per_request"], 2),
            "host_busy_frac": round(sv["host_busy_frac"], 4), "pool": sv["pool"],
        }
    cpu = None
    if not a.no_cpu:
        _progress("CPU baselines")
        cpu = _cpu_baseline(n)
        cpu["host"] = _host_info()
        cpu["libraries"] = _cpu_library_baselines()

    _progress("streaming ceilings")
    ceilings = _stream_ceilings()
    ms_per_step = main_res["wall_s"] * 1e3 / a.steps
    out = {
        "metric": "scan GB/s & %HBM peak (N=2^28 fp16/int8); SpMV GFLOP/s; p50/p99 mixed-stream",
        "value": round(main_res["gbs_wall"], 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp16->f32 (f64 accumulate)",
        "data": "synthetic: U[-1,1] rounded to fp16, torch.Generator seed 0",
        "config": {"workload": f"large 1-D inclusive scan, N=2^{a.log2n} fp16 -> fp32, S={a.S}",
                   "op": "tcuscan.run_scan_multi_core", "N": n, "S": a.S,
                   "bytes_per_step": main_res["bytes_per_step"]},
        "roofline": {
            "bound": "hbm",
            "achieved": round(main_res["gbs_kernel"], 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(main_res["gbs_kernel"] / HBM_PEAK_GBS, 4),
            "traffic": _pmc_traffic(),
            "kernel_avg_ms": round(main_res["event_avg_ms"], 4),
            "kernel_p50_ms": round(main_res["event_p50_ms"], 4),
            "kernel_p99_ms": round(main_res["event_p99_ms"], 4),
            "kernel_samples": main_res["event_n"],
            "flushed": _rnd(main_res["flushed"]),
            "stream_ceilings_gbs": ceilings,
            "note": "stream_ceilings_gbs: HBM read / write / copy rates of a calibration kernel (nontemporal "
                    "16-byte accesses), the practical ceilings beside the 8 TB/s peak",
        },
        "cpu_baseline": cpu,
        "extras": extras,
    }
    print(json.dumps(out))
    return 0


def _pmc_traffic():
    """HBM bytes per launch of the scan kernel from the committed rocprofv3
    --pmc summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel
    calibration factor recorded beside them), or None when absent."""
    path = os.path.join(REPO, "profiles", "scan_f16_2p28_pmc.json")
    try:
        with open(path) as f:
            d = json.load(f)
        return d.get("traffic_bytes_per_launch")
    except (OSError, ValueError):
        return None


if __name__ == "__main__":
    sys.exit(main())
