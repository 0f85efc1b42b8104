"""run_spmv_v2_multi_cube and run_seg_scan_mc_revert.

CPU: the revert oracle against the reference test's per-segment cumsum loop
(tests/test_seg_scan_mc_revert.py:37-49) on integer data, where float32 is
exact.  GPU: the revert bit for bit against the oracle on the reference's
density x length grid (and within the reference's atol 1e-2 of the loop);
the multi-cube SpMV on the reference's grid (tests/test_spmv_v2_multi_cube.py:
39-88) against the float64 oracle (rtol = atol = 1e-6) and scipy, equal to
run_spmv_v2, and rejecting the dtypes the reference rejects.
"""
import numpy as np
import pytest
import torch

import oracle
from _util import assert_exact, sparse_random, to_np

MULT = [1, 2, 3, 5, 11, 23]  # tests/test_seg_scan_mc_revert.py:35
NROW = [64 * 64 - 1, 24 * 128 + 1, 16 * 128 - 13, 64 * 128 + 13, 32 * 128 - 133]  # tests/test_spmv_v2_multi_cube.py:30


def ref_segscan(x, f):
    # tests/test_seg_scan_mc_revert.py:37-49
    z = np.zeros(len(x))
    scan_f = np.cumsum(f)
    index = 0
    for seg_id in range(0, scan_f[-1] + 1):
        seg_size = np.sum(scan_f == seg_id)
        z[index:index + seg_size] = np.cumsum(x[scan_f == seg_id])
        index += seg_size
    return z


def revert_inputs(n, density, seed=42):
    # tests/test_seg_scan_mc_revert.py:52-70
    rng = np.random.default_rng(seed)
    x = rng.integers(0, 10, n).astype(np.float32)
    f = np.zeros(n, dtype=np.int32)
    f[rng.choice(n, int(n * density), replace=False)] = 1
    f[0] = 0
    sx = np.cumsum(x).astype(np.float32)
    sf = np.cumsum(f).astype(np.int32)
    d = np.compress(np.append(f[1:], 1), sx).astype(np.float32)
    return x, f, sx, sf, d


def test_oracle_revert_matches_reference_loop():
    for n, dens in ((5000, 0.01), (20000, 0.1), (3000, 0.0)):
        x, f, sx, sf, d = revert_inputs(n, dens, n)
        assert_exact(oracle.seg_scan_revert(sx, sf, d), ref_segscan(x, f).astype(np.float32))


@pytest.mark.gpu
@pytest.mark.parametrize("density,per_core", [(0.01, 2048), (0.1, 1024), (0.001, 1024), (0.003, 1024)])
def test_seg_scan_mc_revert_reference_grid(ops, dev, density, per_core):
    for m in MULT:
        n = 40 * per_core * m
        x, f, sx, sf, d = revert_inputs(n, density)
        z = ops.run_seg_scan_mc_revert(torch.from_numpy(sx).to(dev), torch.from_numpy(sf).to(dev),
                                       torch.from_numpy(d).to(dev))
        assert z.dtype == torch.float32 and z.shape == (n,)
        assert_exact(to_np(z), oracle.seg_scan_revert(sx, sf, d), f"n={n}")
        if n <= 150_000:  # the loop is quadratic; the oracle is pinned to it on the CPU
            np.testing.assert_allclose(to_np(z), ref_segscan(x, f), atol=1e-2)


@pytest.mark.gpu
def test_seg_scan_mc_revert_edges(ops, dev):
    sx = torch.arange(1, 10001, dtype=torch.float32, device=dev)
    sf = torch.zeros(10000, dtype=torch.int32, device=dev)
    assert torch.equal(ops.run_seg_scan_mc_revert(sx, sf, torch.empty(0, device=dev)), sx)
    assert ops.run_seg_scan_mc_revert(sx[:0], sf[:0], sx[:0]).numel() == 0
    with pytest.raises(RuntimeError):
        ops.run_seg_scan_mc_revert(sx.half(), sf, sx)
    with pytest.raises(RuntimeError):
        ops.run_seg_scan_mc_revert(sx, sf[:-1], sx)


@pytest.mark.gpu
@pytest.mark.parametrize("s", [32, 64, 128])
@pytest.mark.parametrize("density", [0.01, 0.001, 0.0001])
def test_spmv_v2_multi_cube_reference_grid(ops, dev, s, density):
    ones = torch.ones((s, s), dtype=torch.float16, device=dev)
    upper, lower = torch.triu(ones), torch.tril(ones, -1)
    for nrow in NROW:
        B = sparse_random(nrow - 1, nrow - 1, density, nrow + s)
        vals = B.data.astype(np.float16)
        ip = B.indptr.astype(np.int32)
        cols = B.indices.astype(np.int32)
        vec = np.random.default_rng(42).uniform(1, 9, nrow - 1).astype(np.float16)
        args = [torch.from_numpy(a).to(dev) for a in (vals, ip, cols, vec)]
        y = ops.run_spmv_v2_multi_cube(*args, upper, lower)
        assert y.dtype == torch.float32 and y.shape == (nrow - 1,)
        np.testing.assert_allclose(to_np(y), oracle.spmv(vals, ip, cols, vec), rtol=1e-6, atol=1e-6)
        np.testing.assert_allclose(to_np(y), B.astype(np.float32) @ vec.astype(np.float32), atol=1.0)
        assert torch.equal(y, ops.run_spmv_v2(*args, s))


@pytest.mark.gpu
def test_spmv_v2_multi_cube_rejects(ops, dev):
    v = torch.ones(4, dtype=torch.float16, device=dev)
    ip = torch.tensor([0, 2, 4], dtype=torch.int32, device=dev)
    c = torch.zeros(4, dtype=torch.int32, device=dev)
    up = torch.ones(32, 32, dtype=torch.float16, device=dev)
    with pytest.raises(RuntimeError):
        ops.run_spmv_v2_multi_cube(v.float(), ip, c, v.float(), up, up)
    with pytest.raises(RuntimeError):
        ops.run_spmv_v2_multi_cube(v, ip.to(torch.uint32), c, v, up, up)
