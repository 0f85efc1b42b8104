"""The decoupled look-back's forward-progress fallback and launch-tag
hygiene, on the GPU.

Fallback: every look-back kernel, when a predecessor tile has not published
its descriptor after a bounded number of polls, computes that tile's
aggregate from the input itself.  With the workspace's poll limit set to 0
every walk takes that path at the first missing predecessor, so the parity
grids below run the scan, seg-scan, matrix-unit scan, compaction, filter,
split, seg-sum, SpMV and sort kernels through their fallbacks and must
still match the oracle bit for bit.  The waits stay bounded (the limit only
shortens them), so every kernel exits normally.

Tags: descriptors carry a 30-bit launch tag taken from an epoch word in the
workspace.  Words with flag 0 or 3 are never accepted, scratch that ops
leave behind never reads as a live descriptor, and the owner of a
workspace re-zeroes it before tags can repeat.  These tests plant stale
words in the workspace (ordinary tensor writes; no kernel is made to fault
or wait) and check that results stay exact.
"""
import numpy as np
import pytest
import torch

import oracle
from _util import assert_exact, gen, to_np, uniform_pm1

pytestmark = pytest.mark.gpu

HEADER_BYTES = 18 * 256          # workspace header: tickets, epoch, arrival counters
EPOCH_OFFSET = 8 * 256           # the epoch word
POLL_OFFSET = 8 * 256 + 128      # the poll-limit word
TAG_CYCLE = (1 << 30) - 1
MAX_CALLS = 1 << 24              # TCS_WORKSPACE_MAX_CALLS


@pytest.fixture
def fallback(ops, dev):
    """Poll limit 0 on the current stream's workspace for one test."""
    like = torch.empty(1, device=dev)
    ops._set_lookback_poll_limit(like, 0)
    torch.cuda.synchronize()
    words = ops._workspace_tensor(like)[POLL_OFFSET:POLL_OFFSET + 8].view(torch.int64)
    assert int(words.item()) == 1, "poll-limit word not set"
    yield
    ops._set_lookback_poll_limit(like, -1)
    torch.cuda.synchronize()


# --- scans -------------------------------------------------------------------
@pytest.mark.parametrize("n", [5000, 1 << 20, 3_000_017])
@pytest.mark.parametrize("dtype", [torch.float16, torch.bfloat16, torch.float32, torch.int8], ids=str)
def test_scan_fallback(ops, dev, fallback, n, dtype):
    if dtype == torch.int8:
        x = torch.randint(-4, 5, (n,), generator=gen(n), dtype=torch.int8)
    else:
        x = uniform_pm1(n, dtype, n)
    y = ops.run_scan_single_core(x.to(dev), 128, 3.0)
    assert_exact(to_np(y), oracle.scan(to_np(x), 3.0), f"single_core n={n}")
    if dtype != torch.float32:
        y = ops.run_scan_multi_core(x.to(dev), 128)
        assert_exact(to_np(y), oracle.scan(to_np(x)), f"multi_core n={n}")


@pytest.mark.parametrize("tiles_per_wg", [1, 4, 0])
@pytest.mark.parametrize("method", ["dpp", "mfma"])
def test_scan_methods_fallback(ops, dev, fallback, tiles_per_wg, method):
    n = (1 << 21) + 333
    for dtype in (torch.float16, torch.int8):
        x = (torch.randint(-4, 5, (n,), generator=gen(n), dtype=torch.int8) if dtype == torch.int8
             else uniform_pm1(n, dtype, n + 1))
        y = ops.run_scan_ex(x.to(dev), 32, tiles_per_wg, method, 0.0, 0)
        assert_exact(to_np(y), oracle.scan(to_np(x)), f"{method} tpw={tiles_per_wg} {dtype}")


@pytest.mark.parametrize("density", [0.0, 0.0001, 0.01, 0.5])
def test_seg_scan_and_rows_fallback(ops, dev, fallback, density):
    n = 2_000_003
    x = torch.randint(-2, 3, (n,), generator=gen(n), dtype=torch.int8).to(torch.float16)
    f = (torch.rand(n, generator=gen(n + 1)) < density).to(torch.int8)
    y = ops.run_seg_scan(x.to(dev), f.to(dev), 32)
    assert_exact(to_np(y), oracle.seg_scan_flags(to_np(x), to_np(f)), f"seg_scan density={density}")
    L = 100_003
    xb = x[: (n // L) * L].reshape(-1, L)
    yb = ops.run_scan_batch(xb.to(dev), 64)
    assert_exact(to_np(yb).reshape(-1), oracle.scan_rows(to_np(xb).reshape(-1), L), "scan_batch")


# --- compaction, filters, split ---------------------------------------------
@pytest.mark.parametrize("p", [0.0, 0.01, 0.3, 1.0])
def test_compaction_fallback(ops, dev, fallback, p):
    n = 3_000_001
    x = torch.rand(n, generator=gen(n))
    m = (torch.rand(n, generator=gen(n + 7)) < p).to(torch.int8)
    idx = torch.arange(n, dtype=torch.int32)
    z, zi = ops.run_compress_ind(x.to(dev), idx.to(dev), m.to(dev), 128)
    ez, ei = oracle.compress_ind(to_np(x), to_np(idx), to_np(m))
    assert_exact(to_np(z), ez, f"compress_ind values p={p}")
    assert_exact(to_np(zi), ei, f"compress_ind indices p={p}")
    z2, zi2 = ops.run_compress_ind_no_arange(x.to(dev), m.to(dev), 128)
    ez2, ei2 = oracle.compress_ind_no_arange(to_np(x), to_np(m))
    assert_exact(to_np(z2), ez2, "no_arange values")
    assert_exact(to_np(zi2), ei2, "no_arange indices")
    # run_compress_pos has no count pass: its offsets come by look-back.
    z3 = ops.run_compress_pos(x.to(dev), m.to(dev), ez.size, 128)
    assert_exact(to_np(z3), ez, f"compress_pos values p={p}")


@pytest.mark.parametrize("pivot", [-3.0, 0.0, 0.5, 3.0])
def test_filter_fallback(ops, dev, fallback, pivot):
    n = 4_000_037
    x = torch.randn(n, generator=gen(n)).half()
    got = ops.run_filter_greater_eq(x.to(dev), pivot, 128)
    assert_exact(to_np(got), oracle.filter_cmp(to_np(x), pivot, 4), f"filter_ge pivot={pivot}")
    pos = ops.run_where(x.to(dev), pivot, 128)
    assert_exact(to_np(pos), oracle.filter_cmp(to_np(x), pivot, 4, positions=True), f"where pivot={pivot}")


@pytest.mark.parametrize("p", [0.0, 0.2, 0.9])
def test_split_fallback(ops, dev, fallback, p):
    n = 2_500_013
    x = torch.randint(-30000, 30000, (n,), generator=gen(n), dtype=torch.int16)
    m = (torch.rand(n, generator=gen(n + 3)) < p).to(torch.int8)
    idx = torch.arange(1, n + 1, dtype=torch.int32)
    z, zi = ops.run_split_ind(x.to(dev), m.to(dev), idx.to(dev), 64)
    ez, ei = oracle.split(to_np(x), to_np(m), to_np(idx))
    assert_exact(to_np(z), ez, f"split values p={p}")
    assert_exact(to_np(zi), ei, f"split indices p={p}")


# --- segmented sums and SpMV -------------------------------------------------
@pytest.mark.parametrize("mean_len", [1, 64, 5000])
def test_seg_sum_fallback(ops, dev, fallback, mean_len):
    rng = np.random.default_rng(mean_len)
    rows = max(1, 2_000_000 // mean_len)
    lens = rng.geometric(1.0 / mean_len, size=rows) - (1 if mean_len > 1 else 0)
    indptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    nnz = int(indptr[-1])
    for dtype in (torch.float16, torch.int8):
        x = (torch.randint(-5, 6, (nnz,), generator=gen(nnz), dtype=torch.int8) if dtype == torch.int8
             else uniform_pm1(nnz, dtype, nnz))
        y = ops.run_seg_sum_multi_core(x.to(dev), torch.from_numpy(indptr).to(dev), 128, None)
        assert_exact(to_np(y), oracle.seg_sum_indptr(to_np(x), indptr), f"seg_sum {dtype} mean={mean_len}")


def test_spmv_fallback(ops, dev, fallback):
    rng = np.random.default_rng(5)
    rows = cols = 1 << 16
    lens = oracle.lognormal_row_lengths(rows, 16.0, 1.2, 20000, rng)
    indptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    nnz = int(indptr[-1])
    row_id = np.repeat(np.arange(rows), lens)
    col = rng.integers(0, cols, nnz)
    col = col[np.lexsort((col, row_id))].astype(np.int32)  # sorted within each row
    vals = uniform_pm1(nnz, torch.float16, 11)
    xv = uniform_pm1(cols, torch.float16, 12)
    y = ops.run_spmv_v2(vals.to(dev), torch.from_numpy(indptr).to(dev), torch.from_numpy(col).to(dev),
                        xv.to(dev), 128, None)
    exp = oracle.spmv(to_np(vals), indptr, col, to_np(xv))
    np.testing.assert_allclose(to_np(y), exp, rtol=1e-6, atol=1e-6)


def test_spmv_split_fallback(ops, dev, fallback):
    """The split gather (a 2^22-column fp16 vector): both parts' chains
    through the fallback, which sums only its own part's products."""
    rng = np.random.default_rng(6)
    rows, cols = 200_003, 1 << 22
    lens = oracle.lognormal_row_lengths(rows, 16.0, 1.2, 20000, rng)
    lens[11] = 20000
    indptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    nnz = int(indptr[-1])
    row_id = np.repeat(np.arange(rows), lens)
    col = rng.integers(0, cols, nnz)
    col = col[np.lexsort((col, row_id))].astype(np.int32)
    vals = uniform_pm1(nnz, torch.float16, 13)
    xv = uniform_pm1(cols, torch.float16, 14)
    y = ops.run_spmv_v2(vals.to(dev), torch.from_numpy(indptr).to(dev), torch.from_numpy(col).to(dev),
                        xv.to(dev), 128, None)
    exp = oracle.spmv(to_np(vals), indptr, col, to_np(xv))
    np.testing.assert_allclose(to_np(y), exp, rtol=1e-6, atol=1e-6)


# --- sort ----------------------------------------------------------------------
@pytest.mark.parametrize("dtype", [torch.float16, torch.int32], ids=str)
def test_sort_fallback(ops, dev, fallback, dtype):
    n = 1_000_003
    x = (torch.randn(n, generator=gen(n)).half() if dtype == torch.float16
         else torch.randint(-2**31, 2**31 - 1, (n,), generator=gen(n), dtype=torch.int32))
    v, i = ops.run_radix_sort(x.to(dev), 128)
    ev, ei = oracle.sort(to_np(x))
    assert_exact(to_np(v), ev, "sorted values")
    assert_exact(to_np(i), ei, "sorted positions")


# --- launch tags ---------------------------------------------------------------
def _words(ops, dev):
    return ops._workspace_tensor(torch.empty(1, device=dev)).view(torch.int64)


def _epoch(ops, dev):
    torch.cuda.synchronize()
    return int(_words(ops, dev)[EPOCH_OFFSET // 8].item())


def _plant(ops, dev, status_of_word, count=8192):
    """Writes `count` descriptor words right after the header: word k gets
    status status_of_word(k) in its high half and a payload that is no
    plausible prefix."""
    w = _words(ops, dev)
    k = torch.arange(count, dtype=torch.int64)
    status = torch.tensor([status_of_word(int(j)) for j in range(count)], dtype=torch.int64)
    w[HEADER_BYTES // 8: HEADER_BYTES // 8 + count] = ((status << 32) | (0x3C00_0000 + k)).to(dev)
    torch.cuda.synchronize()


def test_words_with_invalid_flags_are_never_descriptors(ops, dev):
    n = 1 << 21
    x = uniform_pm1(n, torch.float16, 1)
    exp = oracle.scan(to_np(x))
    xi = torch.randint(-4, 5, (n,), generator=gen(2), dtype=torch.int8)
    expi = oracle.scan(to_np(xi))
    for _ in range(3):
        e = _epoch(ops, dev)
        # Words carrying the next launch's tag (and the one after) with flag 0 or 3.
        tags = [(e % TAG_CYCLE) + 1, ((e + 1) % TAG_CYCLE) + 1]
        _plant(ops, dev, lambda j: (tags[(j // 2) % 2] << 2) | (0 if j % 4 < 2 else 3))
        assert_exact(to_np(ops.run_scan_multi_core(x.to(dev), 128)), exp, "fp16 after planted words")
        e = _epoch(ops, dev)
        tag = (e % TAG_CYCLE) + 1
        _plant(ops, dev, lambda j: (tag << 2) | (0 if j % 2 else 3))
        assert_exact(to_np(ops.run_scan_multi_core(xi.to(dev), 128)), expi, "int8 after planted words")


def test_workspace_rezeroed_before_tags_repeat(ops, dev):
    """Near the end of the tag cycle, with stale descriptors of the tags the
    next cycle starts with, the owner's call counter reaches its limit and
    the workspace is re-zeroed before the scan runs."""
    n = (1 << 22) + 5
    x = uniform_pm1(n, torch.float16, 3)
    exp = oracle.scan(to_np(x))
    like = torch.empty(1, device=dev)
    w = _words(ops, dev)
    w[EPOCH_OFFSET // 8] = TAG_CYCLE - 2
    # Tags 1..4 (the first ones after a wrap), flags 1 and 2, payload garbage.
    _plant(ops, dev, lambda j: ((1 + (j // 2) % 4) << 2) | (1 + (j // 2) % 2))
    ops._workspace_set_calls(like, MAX_CALLS - 1)
    y = ops.run_scan_multi_core(x.to(dev), 128)
    assert_exact(to_np(y), exp, "scan after the re-zeroing")
    assert _epoch(ops, dev) == 1, "the workspace was not re-zeroed"
    # The epoch keeps advancing from zero afterwards.
    assert_exact(to_np(ops.run_scan_multi_core(x.to(dev), 128)), exp, "second scan")
    assert _epoch(ops, dev) == 2


def test_scratch_left_behind_is_never_a_live_descriptor(ops, dev):
    """After ops whose scratch holds doubles and signed integers
    (reduce_tiles, complete_blocks, complete_rows), every word of the
    workspace past the header is zero in its high half or a descriptor of
    an earlier launch (flag 1 or 2, tag at most the last one used)."""
    like = torch.empty(1, device=dev)
    ops._workspace_set_calls(like, MAX_CALLS - 1)
    ops.run_scan_multi_core(torch.ones(1000, dtype=torch.int8, device=dev), 128)  # re-zeroes first
    assert _epoch(ops, dev) == 1
    n = 40 * 128 * 128
    xf = (1000 * torch.randn(n, generator=gen(9))).float()
    xi = torch.randint(-2**20, 2**20, (n,), generator=gen(10), dtype=torch.int32)
    ops.run_reduce_tiles(xf.half().to(dev), 128 * 64, 40)
    ops.run_reduce_tiles(xi.to(dev), 128 * 64, 40)
    sums = -torch.ones(40, dtype=torch.float32) * 12345.5
    ops.run_complete_blocks(xf.to(dev), sums.to(dev), 128)
    ops.run_complete_blocks(xi.to(dev), (-sums).int().to(dev), 128)
    ops.run_complete_rows(xf.to(dev), sums.to(dev), 128, 64)
    ops.run_complete_rows(xi.to(dev), (-sums).int().to(dev), 128, 64)
    ops.run_complete_rows(xi.long().to(dev), (-sums).long().to(dev), 128, 64)
    e = _epoch(ops, dev)
    last_tag = ((e - 1) % TAG_CYCLE) + 1
    w = _words(ops, dev)[HEADER_BYTES // 8:].cpu().numpy().view(np.uint64)
    hi = (w >> np.uint64(32)).astype(np.uint64)
    flag = hi & np.uint64(3)
    tag = hi >> np.uint64(2)
    ok = (hi == 0) | (((flag == 1) | (flag == 2)) & (tag >= 1) & (tag <= last_tag))
    bad = np.flatnonzero(~ok)
    assert bad.size == 0, f"{bad.size} scratch words read as descriptors, first at {bad[:4]}: {hex(int(w[bad[0]]))}"


# --- int64 input of run_complete_rows (src/torch/torch_reduce.h:165-170) -----
@pytest.mark.parametrize("tile", [32, 128])
def test_complete_rows_int64(ops, dev, tile):
    num_blocks = 10
    n = num_blocks * tile * tile * 3
    raw = torch.randint(-1000, 1000, (n,), generator=gen(n), dtype=torch.int64)
    # Row scans of the input and its block totals: complete_rows then gives
    # the full scan (blocks are n / num_blocks long at these sizes).
    x = raw.reshape(-1, tile).cumsum(1).reshape(-1)
    sums = raw.reshape(num_blocks, -1).sum(1)
    y = ops.run_complete_rows(x.to(dev), sums.to(dev), tile, tile)
    assert y.dtype == torch.int32 and y.shape == (n,)
    exp = oracle.complete_rows(x.int().numpy(), sums.int().numpy(), tile, tile)
    assert_exact(to_np(y), exp, "int64 complete_rows")
    y32 = ops.run_complete_rows(x.int().to(dev), sums.int().to(dev), tile, tile)
    assert_exact(to_np(y), to_np(y32), "int64 vs int32 input")
    assert_exact(to_np(y), np.cumsum(raw.numpy()).astype(np.int32), "closure: full scan")
