"""Pins the CPU oracle before it is trusted as the checker.

Each oracle function is checked against something independent of its own
formula: a closed form, an invariant, a library routine (numpy / scipy /
torch on the CPU) or a brute-force loop over a tiny input.  The checks are
chosen so that a dropped term, a wrong sign, an off-by-one segment boundary
or a transposed operand fails at least one of them.  The reference ships no
golden vectors; its only known-answer example (kernel_seg_sum_vec_revert.h:
22-24 of the reference) is internally inconsistent and is checked here in
corrected form.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

import oracle
from _util import assert_exact


def brute_scan(x):
    out, run = [], 0
    for v in x.tolist():
        run += v
        out.append(run)
    return out


# --- scan ------------------------------------------------------------------
def test_scan_closed_forms():
    n = 1000
    assert_exact(oracle.scan(np.ones(n, np.float16)), np.arange(1, n + 1, dtype=np.float32))
    a = np.arange(n, dtype=np.int32)
    assert_exact(oracle.scan(a), (a * (a + 1) // 2).astype(np.int32))
    assert_exact(oracle.scan(np.ones(n, np.int8), start=5), np.arange(6, n + 6, dtype=np.int32))


def test_scan_brute_force_tiny(rng):
    for n in (1, 2, 3, 7, 33):
        x = rng.integers(-128, 128, n).astype(np.int8)
        assert oracle.scan(x).tolist() == brute_scan(x)
        xf = rng.standard_normal(n).astype(np.float16)
        exp = np.array(brute_scan(xf.astype(np.float64)), np.float64).astype(np.float32)
        assert_exact(oracle.scan(xf), exp)


def test_scan_invariants(rng):
    x = rng.integers(-100, 100, 10_001).astype(np.int8)
    y = oracle.scan(x).astype(np.int64)
    assert np.array_equal(np.diff(y), x[1:].astype(np.int64)) and y[0] == x[0]
    a, b = x[:4321], x[4321:]
    cat = np.concatenate([oracle.scan(a), oracle.scan(b) + np.int32(a.astype(np.int64).sum())])
    assert_exact(oracle.scan(x), cat.astype(np.int32))


def test_scan_matches_torch_cumsum(rng):
    xf = rng.standard_normal(100_000).astype(np.float16)
    exp = torch.cumsum(torch.from_numpy(xf).double(), 0).float().numpy()
    assert_exact(oracle.scan(xf), exp)
    xi = rng.integers(-2**31, 2**31 - 1, 1000, dtype=np.int64).astype(np.int32)
    exp = torch.cumsum(torch.from_numpy(xi), 0, dtype=torch.int32).numpy()  # wraps
    assert_exact(oracle.scan(xi), exp)


def test_scan_start_is_rounded_to_float32():
    y = oracle.scan(np.zeros(3, np.float16), start=0.1)
    assert y[0] == np.float32(0.1)
    y = oracle.scan(np.zeros(3, np.int8), start=7.9)
    assert y.tolist() == [7, 7, 7]


# --- rows / segments ----------------------------------------------------------
def test_scan_rows_against_per_row_loop(rng):
    x = rng.integers(-5, 5, 1000).astype(np.int8)
    for L in (1, 7, 64, 999, 1000, 5000):
        exp = []
        for s in range(0, x.size, L):
            exp += brute_scan(x[s:s + L])
        assert oracle.scan_rows(x, L).tolist() == exp
    assert_exact(oracle.scan_rows(x, x.size), oracle.scan(x))


def test_seg_scan_flags_against_reference_loop(rng):
    # Restates the segment walk of tests/test_seg_scan_single_core.py:40-63.
    x = rng.integers(0, 10, 3000).astype(np.int8)
    f = (rng.random(3000) < 0.02).astype(np.int8)
    seg = np.cumsum(f)
    exp = np.empty(x.size, np.int64)
    for sid in range(int(seg[-1]) + 1):
        idx = np.flatnonzero(seg == sid)
        if idx.size:
            exp[idx] = np.cumsum(x[idx].astype(np.int64))
    assert oracle.seg_scan_flags(x, f).tolist() == exp.tolist()
    assert_exact(oracle.seg_scan_flags(x, np.zeros_like(f)), oracle.scan(x))
    assert oracle.seg_scan_flags(x, np.ones_like(f)).tolist() == x.astype(np.int32).tolist()


def test_seg_scan_flags_fp16_exact(rng):
    x = rng.standard_normal(5000).astype(np.float16)
    f = (rng.random(5000) < 0.01).astype(np.int8)
    y = oracle.seg_scan_flags(x, f)
    starts = np.concatenate([[0], np.flatnonzero(f), [x.size]])
    for a, b in zip(starts[:-1], starts[1:]):
        if b > a:
            assert_exact(y[a:b], np.cumsum(x[a:b].astype(np.float64)).astype(np.float32))


# --- reduce_tiles / complete_rows / complete_blocks ----------------------------
def test_work_distribution_properties():
    for n, t, nb in ((1000, 32, 7), (10, 4, 20), (16384 * 40 + 5, 8192, 40), (0, 16, 3)):
        r = oracle.work_distribution(n, t, nb)
        assert len(r) == nb and r[0][0] == 0 and r[-1][1] == n
        for (s0, e0), (s1, e1) in zip(r[:-1], r[1:]):
            assert e0 == s1 and s0 <= e0
        per = -(-(-(-n // t)) // nb) * t
        assert all(e - s <= per for s, e in r)


def test_reduce_tiles_properties(rng):
    x = rng.integers(-3, 3, 40 * 128 * 128 + 11).astype(np.int8)
    out = oracle.reduce_tiles(x, 128 * 128 // 2, 40)
    assert out.dtype == np.int32 and out.size == 40
    assert int(out.astype(np.int64).sum()) == int(x.astype(np.int64).sum())
    assert oracle.reduce_tiles(x, 64, 1).tolist() == [int(x.astype(np.int64).sum())]
    small = rng.integers(-3, 3, 12).astype(np.int8)
    assert oracle.reduce_tiles(small, 1, 12).tolist() == small.astype(np.int32).tolist()
    # 10 elements, tiles of 3 -> 4 tiles -> 2 tiles per block -> [0,6),[6,10),[]
    t = np.arange(10, dtype=np.int8)
    assert oracle.reduce_tiles(t, 3, 3).tolist() == [15, 30, 0]


def test_complete_rows_closes_the_scan(rng):
    """Row scans + per-block sums through complete_rows give the full scan."""
    for w, nb, mult in ((32, 2, 3), (128, 10, 3), (512, 4, 3)):
        n = nb * w * w * mult
        x = rng.integers(-3, 3, n).astype(np.int32)
        sums = x.reshape(nb, -1).astype(np.int64).sum(1).astype(np.int32)
        rows = oracle.scan_rows(x, w)
        assert_exact(oracle.complete_rows(rows, sums, w, w), oracle.scan(x))


def test_complete_rows_misaligned_blocks_close_the_scan(rng):
    """When len(x) is not a multiple of the tile grid, blocks follow the
    reference's GetWorkDistribution, not len(x) / len(sums)."""
    w, nb, n = 4, 3, 60  # tiles of 16 -> 4 tiles -> blocks [0,32), [32,60), []
    x = rng.integers(-3, 3, n).astype(np.int32)
    r = oracle.work_distribution(n, 16, nb)
    assert r == [(0, 32), (32, 60), (60, 60)]
    sums = np.array([x[s:e].sum() for s, e in r], np.int32)
    rows = np.concatenate([oracle.scan_rows(x[:32], w), oracle.scan_rows(x[32:], w)])
    assert_exact(oracle.complete_rows(rows, sums, w, w), oracle.scan(x))


def test_complete_rows_reference_loop(rng):
    # Restates ref_complete_rows of tests/test_complete_rows.py:27-55 on a tiny
    # case aligned to the tile grid (tiles of w*w, as in that test's grid).
    w, nb = 4, 3
    x = rng.integers(-3, 3, nb * w * w * 2).astype(np.int32)
    sums = rng.integers(-9, 9, nb).astype(np.int32)
    exp = x.astype(np.int64).copy()
    bl = x.size // nb
    pre = np.concatenate([[0], np.cumsum(sums)[:-1]])
    for b in range(nb):
        acc = pre[b]
        for h in range(bl // w):
            s = b * bl + h * w
            exp[s:s + w] += acc
            acc = exp[s + w - 1]
    assert oracle.complete_rows(x, sums, w, w).tolist() == exp.tolist()


def test_complete_blocks_closes_the_scan(rng):
    for s, nb, mult in ((32, 2, 3), (64, 10, 5), (128, 4, 7)):
        n = nb * s * s * mult
        x = rng.integers(0, 3, n).astype(np.int32)
        sums = x.reshape(nb, -1).astype(np.int64).sum(1).astype(np.int32)
        blocks = oracle.scan_rows(x, n // nb)
        assert_exact(oracle.complete_blocks(blocks, sums, s * s), oracle.scan(x))
    xf = (0.1 * rng.standard_normal(4 * 1024)).astype(np.float32)
    out = oracle.complete_blocks(xf, np.array([1, 2, 3, 4], np.float32), 256)
    assert_exact(out[1024:2048], (xf[1024:2048].astype(np.float64) + 1).astype(np.float32))
    assert_exact(out[3072:], (xf[3072:].astype(np.float64) + 6).astype(np.float32))


# --- segmented sums / SpMV / gather -----------------------------------------
def test_seg_sum_known_answer_corrected():
    # Docstring example of kernel_seg_sum_vec_revert.h:22-24 states
    # [1..10] with segment starts [0,4,7] -> (10, 11, 34); the sums of
    # [0,4), [4,7), [7,10) are (10, 18, 27).
    x = np.arange(1, 11, dtype=np.int8)
    assert oracle.seg_sum_indptr(x, np.array([0, 4, 7, 10])).tolist() == [10, 18, 27]
    assert oracle.seg_sum_indptr(x, np.array([0, 4, 7]), ends_implicit=True).tolist() == [10, 18, 27]


def test_seg_sum_matches_scipy(rng):
    A = sp.random(513, 128 * 128 - 13, density=0.01, format="csr", dtype=np.float32,
                  random_state=np.random.RandomState(1))
    A.data = (A.data - 0.5).astype(np.float16).astype(np.float32)
    exp = (A.astype(np.float64) @ np.ones(A.shape[1])).astype(np.float32)
    assert_exact(oracle.seg_sum_indptr(A.data.astype(np.float16), A.indptr), exp)
    Ai = A.copy()
    Ai.data = rng.integers(-2, 2, Ai.nnz).astype(np.float32)
    expi = (Ai @ np.ones(A.shape[1])).astype(np.int32)
    assert_exact(oracle.seg_sum_indptr(Ai.data.astype(np.int8), Ai.indptr), expi)


def test_seg_sum_empty_rows_and_edges():
    x = np.arange(1, 7, dtype=np.int8)
    ip = np.array([0, 0, 2, 2, 2, 6, 6])
    assert oracle.seg_sum_indptr(x, ip).tolist() == [0, 3, 0, 0, 18, 0]
    assert oracle.seg_sum_indptr(x, np.array([0, 6])).tolist() == [21]
    assert oracle.seg_sum_indptr(np.zeros(0, np.int8), np.array([0, 0, 0])).tolist() == [0, 0]


def test_seg_sum_flags_against_reference_loop(rng):
    # Restates ref_segsum of tests/test_segmented_sum.py:39-66 with the start
    # flags shifted into end flags as that test does (:70-75).
    n = 5000
    x = rng.integers(-2, 2, n).astype(np.float16)
    f = (rng.random(n) < 0.01).astype(np.int8)
    f[0] = 0
    scan_f = np.cumsum(f)
    nseg = int(scan_f[-1])
    exp = np.zeros(nseg + 1, np.float64)
    off = 0
    for sid in range(nseg + 1):
        size = int((scan_f == sid).sum())
        exp[sid] = x[off:off + size].astype(np.float64).sum()
        off += size
    ends = np.concatenate([f[1:], [1]]).astype(np.int8)
    assert_exact(oracle.seg_sum_flags(x, ends), exp.astype(np.float32))


def test_csr_gather_matches_torch(rng):
    vals = rng.standard_normal(10_000).astype(np.float16)
    x = rng.standard_normal(777).astype(np.float16)
    cols = rng.integers(0, 777, 10_000).astype(np.int32)
    exp = (torch.from_numpy(vals) * torch.from_numpy(x)[torch.from_numpy(cols).long()]).numpy()
    assert_exact(oracle.csr_gather(vals, cols, x), exp)
    vi = rng.integers(-30000, 30000, 10_000).astype(np.int16)
    xi = rng.integers(-30000, 30000, 777).astype(np.int16)
    exp = (torch.from_numpy(vi) * torch.from_numpy(xi)[torch.from_numpy(cols).long()]).numpy()
    assert_exact(oracle.csr_gather(vi, cols, xi), exp)


def test_spmv_matches_scipy(rng):
    A = sp.random(2000, 1500, density=0.01, format="csr", random_state=np.random.RandomState(2))
    A.data = (2 * A.data - 1).astype(np.float16).astype(np.float64)
    x = (2 * rng.random(1500) - 1).astype(np.float16)
    exp = (A @ x.astype(np.float64)).astype(np.float32)
    assert_exact(oracle.spmv(A.data.astype(np.float16), A.indptr, A.indices, x), exp)


# --- compaction ------------------------------------------------------------------
def test_compress_matches_masked_select(rng):
    x = rng.standard_normal(10_007).astype(np.float32)
    m = (rng.random(10_007) < 0.3).astype(np.int8)
    exp = torch.masked_select(torch.from_numpy(x), torch.from_numpy(m).bool()).numpy()
    assert_exact(oracle.compress(x, m), exp)
    v, i = oracle.compress_ind(x, np.arange(x.size, dtype=np.int32), m)
    assert_exact(v, exp)
    assert_exact(i, torch.nonzero(torch.from_numpy(m)).flatten().int().numpy())
    v2, i2 = oracle.compress_ind_no_arange(x, m)
    assert_exact(v2, exp)
    assert_exact(i2, i)


def test_compress_keeps_only_ones():
    x = np.arange(6, dtype=np.int32)
    m = np.array([1, 2, 0, 1, -1, 1], np.int8)
    assert oracle.compress(x, m).tolist() == [0, 3, 5]
