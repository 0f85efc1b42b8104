"""GPU tests of the serving loop: mixed scan / seg-sum / SpMV / compaction
requests admitted under an HBM budget and overlapped on several streams.
Each request's result must be identical in the mix and alone, and alone it
must match the oracle."""
import numpy as np
import pytest
import torch

import oracle
from _util import assert_exact, gen, to_np, uniform_pm1
from tcuscan import workloads

pytestmark = pytest.mark.gpu


def _pool(ops, srv, dev):
    """Registers a small mixed pool; returns [(entry, kind, inputs)]."""
    out = []
    for i, n in enumerate([1000, 70_001, 1 << 20]):
        x = uniform_pm1(n, torch.float16, 10 + i).to(dev)
        out.append((srv.add_scan(x), "scan", (x,)))
    for i, n in enumerate([513, 300_007]):
        x = torch.randint(-128, 128, (n,), generator=gen(20 + i), dtype=torch.int8).to(dev)
        out.append((srv.add_scan(x), "scan", (x,)))
    for i in range(2):
        ss = workloads.geometric_segsums(1, 64.0, 8 + 3 * i, 8 + 3 * i + 0.01, seed=30 + i, device=dev)
        x = ss.x.half() if i == 0 else ss.x
        out.append((srv.add_seg_sum(x, ss.indptr), "segsum", (x, ss.indptr)))
    for i, rows in enumerate([3000, 40_000]):
        vals, ip, cols, xv = workloads.lognormal_csr(rows, rows, 16.0, 1.2, 20000, seed=40 + i, device=dev)
        out.append((srv.add_spmv(vals, ip, cols, xv), "spmv", (vals, ip, cols, xv)))
    for i, (n, p) in enumerate([(5000, 0.05), (400_003, 0.9)]):
        x = torch.rand(n, generator=gen(50 + i)).to(dev)
        m = (torch.rand(n, generator=gen(60 + i)) < p).to(torch.int8).to(dev)
        out.append((srv.add_compress(x, m), "compress", (x, m)))
    return out


def _oracle(kind, inputs):
    a = [to_np(t) for t in inputs]
    if kind == "scan":
        return oracle.scan(a[0])
    if kind == "segsum":
        return oracle.seg_sum_indptr(a[0], a[1])
    if kind == "spmv":
        return oracle.spmv(*a)
    return oracle.compress(a[0], a[1])


def _check_alone(kind, got, inputs):
    exp = _oracle(kind, inputs)
    if kind == "spmv":
        np.testing.assert_allclose(got, exp, rtol=1e-6, atol=1e-6)
    else:
        assert_exact(got, exp, kind)


def test_mix_equals_alone(ops, dev):
    srv = ops.Server(1 << 30, 4, 8)
    pool = _pool(ops, srv, dev)
    alone = {}
    for e, kind, inputs in pool:
        alone[e] = srv.run_alone(e)
        _check_alone(kind, to_np(alone[e]), inputs)
    rng = np.random.default_rng(1)
    n_req = 400
    arrivals = np.cumsum(rng.exponential(1 / 50_000.0, n_req))  # dense: requests overlap
    entries = rng.integers(0, len(pool), n_req)
    st = srv.run(arrivals.tolist(), entries.tolist(), list(range(n_req)))
    assert st["requests"] == n_req and st["batches"] >= 1
    assert 0 < st["p50_ms"] <= st["p99_ms"] <= st["max_ms"]
    assert st["peak_arena_bytes"] <= 1 << 30
    for r in range(n_req):
        assert torch.equal(srv.kept(r), alone[int(entries[r])]), f"request {r} (entry {entries[r]})"


def test_budget_makes_requests_wait(ops, dev):
    # Budget just above the largest request: at most one large request holds
    # the arena at a time, so others queue; every request still completes
    # with its alone result.
    srv = ops.Server(6 << 20, 2, 4)
    x = uniform_pm1(1 << 20, torch.float16, 70).to(dev)      # 4 MiB of output
    e_big = srv.add_scan(x)
    y = torch.randint(-128, 128, (10_000,), generator=gen(71), dtype=torch.int8).to(dev)
    e_small = srv.add_scan(y)
    ref_big, ref_small = srv.run_alone(e_big), srv.run_alone(e_small)
    entries = [e_big, e_small] * 40
    arrivals = [0.0] * len(entries)  # all at once
    st = srv.run(arrivals, entries, list(range(len(entries))))
    assert st["peak_arena_bytes"] <= 6 << 20
    assert st["mean_queue_ms"] > 0
    for r, e in enumerate(entries):
        assert torch.equal(srv.kept(r), ref_big if e == e_big else ref_small)


def test_server_rejects_bad_input(ops, dev):
    srv = ops.Server(1 << 20, 1, 1)
    with pytest.raises(RuntimeError):
        srv.add_scan(torch.ones(10, dtype=torch.float16))           # host tensor
    with pytest.raises(RuntimeError):
        srv.add_scan(torch.ones(1 << 20, dtype=torch.float16, device=dev))  # outputs exceed budget
    with pytest.raises(RuntimeError):
        srv.run([0.0, 1.0], [0], [])
