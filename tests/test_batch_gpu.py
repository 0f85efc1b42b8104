"""GPU tests of the batched form: plans of many uneven scans and segmented
sums, tuned and replayed as one HIP graph, against the oracle and against
per-call launches of the plain operators; and torch graph capture of the
plain operators (launch tags come from the workspace, so replays of one
captured graph stay correct).

Every graph test replays several times with fresh input data, eager work and
synchronisation between replays, and checks every replay.
"""
import numpy as np
import pytest
import torch

import oracle
from _util import assert_exact, gen, to_np, uniform_pm1
from tcuscan import workloads

pytestmark = pytest.mark.gpu


def _refill(t: torch.Tensor, seed: int):
    g = torch.Generator(device=t.device).manual_seed(seed)
    if t.dtype == torch.int8:
        t.copy_(torch.randint(-128, 128, t.shape, generator=g, device=t.device, dtype=torch.int8))
    else:
        t.copy_((torch.rand(t.shape, generator=g, device=t.device) * 2 - 1).to(t.dtype))


def test_plan_of_uneven_scans_and_segsums(ops, dev):
    sc = workloads.uneven_scans(300, 8, 14, seed=1, device=dev)
    ss = workloads.geometric_segsums(60, 64.0, 4, 9, seed=2, device=dev)
    plan = ops.BatchPlan()
    y_scan = plan.add_seg_scan(sc.x, sc.flags)
    y_sum = plan.add_seg_sum(ss.x, ss.indptr)
    assert plan.num_ops() == 2
    plan.tune(reps=1)
    cfgs = plan.configs()
    assert cfgs[0][0] in (32, 64, 128) and cfgs[0][4] > 0  # the scan op was timed
    plan.instantiate(max_branches=2)
    flags_np = to_np(sc.flags)
    for rep in range(3):
        _refill(sc.x, 100 + rep)
        _refill(ss.x, 200 + rep)
        # Eager work on the same stream between replays.
        eager = ops.run_seg_scan(sc.x, sc.flags, 128)
        plan.launch()
        torch.cuda.synchronize()
        assert_exact(to_np(y_scan), oracle.seg_scan_flags(to_np(sc.x), flags_np), f"replay {rep}")
        assert torch.equal(y_scan, eager)
        assert_exact(to_np(y_sum), oracle.seg_sum_indptr(to_np(ss.x), ss.indptr_host), f"replay {rep}")
    # Batched results equal per-call launches of the plain operators.
    for i, piece in enumerate(sc.pieces()[:40]):
        o0, o1 = sc.offsets[i], sc.offsets[i + 1]
        assert torch.equal(ops.run_scan_multi_core(piece, 128), y_scan[o0:o1])
    for i, (xi, ipi) in enumerate(ss.pieces()[:20]):
        r0, r1 = ss.row_off[i], ss.row_off[i + 1]
        assert torch.equal(ops.run_seg_sum_multi_core(xi, ipi, 128), y_sum[r0:r1])


def test_plan_scans_spmv_and_eager_run(ops, dev):
    x16 = uniform_pm1(1_000_003, torch.float16, 5).to(dev)
    x8 = torch.randint(-128, 128, (700_001,), generator=gen(6), dtype=torch.int8).to(dev)
    vals, indptr, cols, xv = workloads.lognormal_csr(20_000, 30_000, 16.0, 1.2, 2000, seed=7, device=dev)
    plan = ops.BatchPlan()
    y16 = plan.add_scan(x16)
    y8 = plan.add_scan(x8)
    ysp = plan.add_spmv(vals, indptr, cols, xv)
    plan.run_eager()  # untuned defaults
    torch.cuda.synchronize()
    assert_exact(to_np(y16), oracle.scan(to_np(x16)))
    assert_exact(to_np(y8), oracle.scan(to_np(x8)))
    ref = oracle.spmv(to_np(vals), to_np(indptr), to_np(cols), to_np(xv))
    np.testing.assert_allclose(to_np(ysp), ref, rtol=1e-6, atol=1e-6)
    plan.tune(reps=1)
    plan.instantiate(max_branches=3)
    for rep in range(3):
        _refill(x16, 300 + rep)
        _refill(x8, 400 + rep)
        y16.zero_()
        plan.launch()
        torch.cuda.synchronize()
        assert_exact(to_np(y16), oracle.scan(to_np(x16)), f"replay {rep}")
        assert_exact(to_np(y8), oracle.scan(to_np(x8)), f"replay {rep}")
        np.testing.assert_allclose(to_np(ysp), ref, rtol=1e-6, atol=1e-6)


def test_plan_rejects_unbindable_inputs(ops, dev):
    plan = ops.BatchPlan()
    x = uniform_pm1(1000, torch.float16, 8).to(dev)
    with pytest.raises(RuntimeError):
        plan.add_scan(x[1:])          # misaligned view: the plan binds addresses
    with pytest.raises(RuntimeError):
        plan.add_scan(x.cpu())
    with pytest.raises(RuntimeError):
        plan.add_seg_scan(x, torch.zeros(999, dtype=torch.int8, device=dev))
    with pytest.raises(RuntimeError):
        plan.launch()                 # nothing instantiated


def _capture(fn, stream):
    with torch.cuda.stream(stream):
        fn()  # warm-up on the capture stream: its workspace exists before capture
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=stream):
        out = fn()
    return g, out


@pytest.mark.parametrize("dtype", [torch.float16, torch.int8], ids=str)
def test_torch_graph_replays_of_plain_ops(ops, dev, dtype):
    n = 3_000_017
    x = (uniform_pm1(n, dtype, 9) if dtype == torch.float16
         else torch.randint(-128, 128, (n,), generator=gen(9), dtype=torch.int8)).to(dev)
    f = (torch.rand(n, generator=gen(10)) < 0.001).to(torch.int8).to(dev)
    ip = torch.from_numpy(np.sort(np.random.default_rng(11).integers(0, n, 5001)).astype(np.int32))
    ip[0], ip[-1] = 0, n
    ip = ip.to(dev)
    mask = (torch.rand(n, generator=gen(12)) < 0.4).to(torch.int8).to(dev)
    s = torch.cuda.Stream()

    def work():
        return (ops.run_scan_multi_core(x, 128), ops.run_seg_scan(x, f, 128),
                ops.run_seg_sum_multi_core(x, ip, 128))

    g, (ya, yb, yc) = _capture(work, s)
    if dtype == torch.float16:
        xs = x.float().contiguous()
        kept = int(to_np(mask).sum())  # outside the capture: a host read cannot be captured
        gm, zm = _capture(lambda: ops.run_compress_pos(xs, mask, kept, 128), s)
    for rep in range(3):
        _refill(x, 500 + rep)
        eager = ops.run_scan_multi_core(x, 128)
        g.replay()
        torch.cuda.synchronize()
        xn, fn_ = to_np(x), to_np(f)
        assert_exact(to_np(ya), oracle.scan(xn), f"replay {rep}")
        assert torch.equal(ya, eager)
        assert_exact(to_np(yb), oracle.seg_scan_flags(xn, fn_), f"replay {rep}")
        assert_exact(to_np(yc), oracle.seg_sum_indptr(xn, to_np(ip)), f"replay {rep}")
        if dtype == torch.float16:
            xs.copy_(x.float())
            gm.replay()
            torch.cuda.synchronize()
            assert_exact(to_np(zm), oracle.compress(to_np(xs), to_np(mask)), f"replay {rep}")


def test_tuning_cache_persists(ops, dev, tmp_path):
    """The tuner's cache goes to a JSON file and back: loaded entries apply
    to plans built afterwards with no timing runs, and a file written for
    another device is ignored."""
    import json
    import math

    from tcuscan import tuning

    sc = workloads.uneven_scans(40, 16, 18, seed=9, device=dev)
    n = sc.x.numel()
    log2n = math.ceil(math.log2(n))
    ops._set_tuned(1, 0, log2n, 64, 2, 1, 32, 0.5)  # seg-scan, fp16: S=64, 2 tiles, DPP, window 32
    path = tmp_path / "tune.json"
    tuning.save(str(path))
    d = json.loads(path.read_text())
    assert d["device"]["compute_units"] == torch.cuda.get_device_properties(0).multi_processor_count
    assert any(e["kind"] == 1 and e["dtype"] == 0 and e["log2n"] == log2n and e["S"] == 64 for e in d["entries"])
    other = tmp_path / "other.json"
    d2 = dict(d, device=dict(d["device"], name="another GPU"))
    other.write_text(json.dumps(d2))
    assert tuning.load(str(other)) == 0
    ops._set_tuned(1, 0, log2n, 32, 16, 1, 8, 0.25)  # overwritten in this process ...
    assert tuning.load(str(path)) == len(d["entries"])  # ... and restored from the file
    plan = ops.BatchPlan()
    y = plan.add_seg_scan(sc.x, sc.flags)
    plan.tune(reps=1)  # a cached key: no timing runs
    S, tpw, method, window, ms = plan.configs()[0]
    assert (S, tpw, method, window) == (64, 2, "dpp", 32) and abs(ms - 0.5) < 1e-6
    plan.instantiate(max_branches=1)
    plan.launch()
    torch.cuda.synchronize()
    assert_exact(to_np(y), oracle.seg_scan_flags(to_np(sc.x), to_np(sc.flags)), "plan with a loaded config")
