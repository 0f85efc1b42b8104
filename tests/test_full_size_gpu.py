"""Parity at the full sizes of the SpMV and batched workloads, against the
oracle (GPU).

* SpMV: 2^22 x 2^22, lognormal row lengths (mean 16, sigma 1.2, clipped to
  [0, 20000]), distinct sorted int32 columns per row, fp16 values and x in
  [-1, 1], about 6.7e7 nonzeros.  The reference cannot run this size: its
  gather keeps x in the 192 KB on-chip buffer (src/kernels/kernel_csr_gather.h:73).
  Here x (8 MB) is larger than one XCD's L2, which the smaller SpMV tests
  never reach.
* The batched plan: 10,000 fp16 scans with lengths floor(2^U(8, 20))
  concatenated with int8 segment-start flags (one flag-segmented scan op)
  plus 2,000 int8 seg-sum problems (2^U(4, 12) geometric segments of mean
  64), tuned and replayed as one HIP graph; every scan piece and every
  segment sum is checked against the oracle, piece by piece so host memory
  stays bounded.
"""
import numpy as np
import pytest
import torch

import oracle
from _util import assert_exact, to_np
from tcuscan import workloads

pytestmark = pytest.mark.gpu


def test_spmv_full_size_lognormal(ops, dev):
    rows = 1 << 22
    vals, indptr, cols, x = workloads.lognormal_csr(rows, rows, 16.0, 1.2, 20000, seed=2, device=dev)
    nnz = vals.numel()
    assert 6.0e7 < nnz < 7.5e7
    y = ops.run_spmv_v2(vals, indptr, cols, x, 128, None)
    torch.cuda.synchronize()
    ip = to_np(indptr)
    c = to_np(cols)
    # Distinct sorted columns in every row (the workload's definition).
    row_start = np.zeros(nnz, dtype=bool)
    row_start[ip[:-1][ip[1:] > ip[:-1]]] = True
    assert np.all((np.diff(c.astype(np.int64)) > 0) | row_start[1:]), "columns not strictly increasing in a row"
    exp = oracle.spmv(to_np(vals), ip, c, to_np(x))
    # Products are exact and summed in double, then rounded once to float32.
    np.testing.assert_allclose(to_np(y), exp, rtol=1e-6, atol=1e-6)


def test_batched_plan_full_size(ops, dev):
    sc = workloads.uneven_scans(10_000, 8, 20, seed=0, device=dev)
    ss = workloads.geometric_segsums(2_000, 64.0, 4, 12, seed=1, device=dev)
    total = int(sc.offsets[-1])
    assert 1.0e9 < total < 1.6e9
    plan = ops.BatchPlan()
    y_scan = plan.add_seg_scan(sc.x, sc.flags)
    y_sum = plan.add_seg_sum(ss.x, ss.indptr)
    plan.tune(reps=1)
    plan.instantiate(max_branches=2)
    plan.launch()
    torch.cuda.synchronize()
    # Every scan piece against the oracle, in chunks of about 2^26 elements.
    o = sc.offsets
    i = 0
    checked = 0
    while i < len(o) - 1:
        j = i
        while j < len(o) - 1 and o[j + 1] - o[i] <= (1 << 26):
            j += 1
        j = max(j, i + 1)
        xs = to_np(sc.x[o[i]:o[j]])
        ys = to_np(y_scan[o[i]:o[j]])
        for k in range(i, j):
            a, b = o[k] - o[i], o[k + 1] - o[i]
            assert_exact(ys[a:b], oracle.scan(xs[a:b]), f"scan piece {k} (length {b - a})")
        checked += j - i
        i = j
    assert checked == 10_000
    exp_sum = oracle.seg_sum_indptr(to_np(ss.x), ss.indptr_host)
    assert_exact(to_np(y_sum), exp_sum, "seg-sums")
    # A second replay after eager work on the same stream gives the same bits.
    first = y_scan[: 1 << 20].clone()
    ops.run_scan_multi_core(sc.x[: 1 << 20], 128)
    plan.launch()
    torch.cuda.synchronize()
    assert torch.equal(y_scan[: 1 << 20], first)
