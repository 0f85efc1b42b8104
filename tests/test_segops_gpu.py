"""GPU parity of the reduction, segmented-sum, SpMV, gather and compaction
operators against the oracle.

Grids follow the reference tests (tests/test_reduce_tiles.py,
test_complete_rows.py, test_complete_blocks.py, test_seg_sum_single_core.py,
test_seg_sum_multi_core.py, test_seg_sum_single_cube.py,
test_seg_sum_multi_cube.py, test_segmented_sum.py, test_csr_gather.py,
test_spmv_v2.py, test_compress*.py), thinned where the reference sweeps
hundreds of large cases.  Integer results, compaction, gathers and fp16 sums
(exact in the double accumulator) are compared bit for bit; complete_rows
and SpMV (whose sums are not exact in double) within a stated tolerance.
"""
from functools import partial

import numpy as np
import pytest
import scipy.sparse as sp
import torch

import oracle
from _util import assert_exact, gen, sparse_random, to_np

pytestmark = pytest.mark.gpu


# --- reduce_tiles (tests/test_reduce_tiles.py:24-68) -------------------------
@pytest.mark.parametrize("m", [1, 2, 7, 15])
@pytest.mark.parametrize("s", [16, 32, 64, 128])
@pytest.mark.parametrize("offset", [-127, -34, -11, -1, 0, 1, 11, 17, 34])
@pytest.mark.parametrize("dtype", [torch.int8, torch.float16], ids=str)
def test_reduce_tiles(ops, dev, m, s, offset, dtype):
    n = m * 40 * 128 * 128 + offset
    g = gen(n + s)
    if dtype == torch.int8:
        x = torch.randint(-3, 3, (n,), generator=g, dtype=torch.int8)
    else:
        x = (0.1 * torch.randn(n, generator=g)).half()
    tile = s * s // 2
    y = ops.run_reduce_tiles(x.to(dev), tile, 40)
    assert y.shape == (40,)
    assert_exact(to_np(y), oracle.reduce_tiles(to_np(x), tile, 40), f"n={n} s={s}")


def test_reduce_tiles_more_blocks_than_tiles(ops, dev):
    x = torch.arange(10, dtype=torch.int8)
    assert to_np(ops.run_reduce_tiles(x.to(dev), 3, 3)).tolist() == [15, 30, 0]


# --- complete_rows (tests/test_complete_rows.py:24-84) -----------------------
@pytest.mark.parametrize("num_blocks", [2, 10, 20, 40])
@pytest.mark.parametrize("tile_len", [32, 128, 512])
def test_complete_rows_fp32(ops, dev, num_blocks, tile_len):
    n = num_blocks * tile_len * tile_len * 3
    x = (0.1 * torch.randn(n, generator=gen(n))).float()
    sums = x.reshape(num_blocks, -1).double().sum(1).float()
    y = ops.run_complete_rows(x.to(dev), sums.to(dev), tile_len, tile_len)
    exp = oracle.complete_rows(to_np(x), to_np(sums), tile_len, tile_len)
    assert y.dtype == torch.float32 and y.shape == (n,)
    # Carries pass through one float32 rounding (the row-end scan), so the
    # result is within a few float32 ulps of the float64 oracle.
    np.testing.assert_allclose(to_np(y), exp, rtol=2e-6, atol=2e-6)
    # The reference's own tolerance.
    np.testing.assert_allclose(to_np(y), exp, rtol=1e-2, atol=1.0)


@pytest.mark.parametrize("num_blocks", [2, 10, 40])
@pytest.mark.parametrize("tile_len", [32, 128, 512])
def test_complete_rows_int32(ops, dev, num_blocks, tile_len):
    n = num_blocks * tile_len * tile_len * 3
    x = torch.randint(-3, 3, (n,), generator=gen(n), dtype=torch.int32)
    sums = x.reshape(num_blocks, -1).sum(1, dtype=torch.int32)
    y = ops.run_complete_rows(x.to(dev), sums.to(dev), tile_len, tile_len)
    assert_exact(to_np(y), oracle.complete_rows(to_np(x), to_np(sums), tile_len, tile_len))


def test_complete_rows_closes_a_row_scan(ops, dev):
    """Row scan, block sums, complete_rows: the full scan (the reference's
    MCSCAN pipeline assembled from its building blocks)."""
    w, nb = 128, 20
    n = nb * w * 64 * 7
    x = torch.randint(-3, 4, (n,), generator=gen(3), dtype=torch.int8)
    xd = x.to(dev)
    rows = ops.run_scan_batch(xd.reshape(-1, w), 32).reshape(-1)
    tile = w * min(w, -(-8192 // w))
    sums = ops.run_reduce_tiles(xd, tile, nb)
    full = ops.run_complete_rows(rows, sums, w, w)
    assert_exact(to_np(full), oracle.scan(to_np(x)))


# --- complete_blocks (tests/test_complete_blocks.py:24-81) -------------------
@pytest.mark.parametrize("num_blocks", [2, 10, 20, 40, 80])
@pytest.mark.parametrize("multiplier", [3, 5, 7])
@pytest.mark.parametrize("tile_ratio", [32, 4, 2])
@pytest.mark.parametrize("dtype", [torch.float32, torch.int32], ids=str)
@pytest.mark.parametrize("s", [32, 64, 128])
def test_complete_blocks(ops, dev, num_blocks, multiplier, tile_ratio, dtype, s):
    n = num_blocks * s * s * multiplier
    g = gen(n + tile_ratio)
    if dtype == torch.float32:
        x = 0.1 * torch.randn(n, generator=g)
    else:
        x = torch.randint(0, 3, (n,), generator=g, dtype=torch.int32)
    sums = x.reshape(num_blocks, -1).sum(1, dtype=dtype)
    y = ops.run_complete_blocks(x.to(dev), sums.to(dev), s * s // tile_ratio)
    assert y.dtype == dtype and y.shape == (n,)
    assert_exact(to_np(y), oracle.complete_blocks(to_np(x), to_np(sums), s * s // tile_ratio))


def test_complete_blocks_rejects_ragged(ops, dev):
    with pytest.raises(RuntimeError):
        ops.run_complete_blocks(torch.zeros(10, device=dev), torch.zeros(3, device=dev), 4)


# --- segmented sums by indptr -------------------------------------------------
_NUM_SEGMENTS = [513, 1025, 2011]
_NUM_COLUMNS = [64 * 64 - 1, 128 * 128, 128 * 128 - 13, 128 * 128 + 13, 128 * 128 - 133]


def _csr(rows, cols, density, seed, integer):
    rs = np.random.RandomState(seed)
    if integer:
        rvs = lambda k: rs.randint(-2, 2, size=k).astype(np.float64)  # noqa: E731
    else:
        rvs = lambda k: 2 * rs.uniform(0, 1, size=k) - 1  # noqa: E731
    return sp.random(rows, cols, density=density, format="csr", random_state=rs, data_rvs=rvs)


def _seg_inputs(rows, cols, dtype, seed):
    A = _csr(rows, cols, 0.01, seed, dtype == torch.int8)
    data = A.data.astype(np.float32)
    x = torch.from_numpy(data).to(dtype)
    indptr = torch.from_numpy(A.indptr.astype(np.int32))
    return A, x, indptr


@pytest.mark.parametrize("num_segments", _NUM_SEGMENTS)
@pytest.mark.parametrize("num_cols", _NUM_COLUMNS)
@pytest.mark.parametrize("s", [32, 64, 128])
@pytest.mark.parametrize("dtype", [torch.int8, torch.float16], ids=str)
def test_seg_sum_single_and_multi_core(ops, dev, num_segments, num_cols, s, dtype):
    A, x, indptr = _seg_inputs(num_segments, num_cols, dtype, num_segments * 7 + num_cols + s)
    exp = oracle.seg_sum_indptr(to_np(x), to_np(indptr))
    out_dtype = torch.int32 if dtype == torch.int8 else torch.float32
    xd = x.to(dev)
    # single_core binds indptr as uint32 in the reference test, multi_core as int32.
    y1 = ops.run_seg_sum_single_core(xd, indptr.to(torch.uint32).to(dev), s)
    y2 = ops.run_seg_sum_multi_core(xd, indptr.to(dev), s)
    for y in (y1, y2):
        assert y.dtype == out_dtype and y.shape == (num_segments,)
        assert_exact(to_np(y), exp, f"rows={num_segments} cols={num_cols}")
    # Reference tolerance against scipy on the (unrounded) float32 data.
    ref = (A.astype(np.float64) @ np.ones(num_cols)).astype(np.float32)
    np.testing.assert_allclose(to_np(y2).astype(np.float64), ref, atol=1e-2)


def test_seg_sum_multi_core_with_segm_offsets(ops, dev):
    A, x, indptr = _seg_inputs(1025, 128 * 128, torch.float16, 5)
    nnz = A.nnz
    block_len = -(-(-(-nnz // 4096)) // 20) * 4096
    sstart = torch.clamp(torch.arange(0, 21, dtype=torch.int32) * block_len, max=nnz)
    so = torch.searchsorted(indptr, sstart, out_int32=True)
    y = ops.run_seg_sum_multi_core(x.to(dev), indptr.to(dev), 64, so.to(dev))
    assert_exact(to_np(y), oracle.seg_sum_indptr(to_np(x), to_np(indptr)))


@pytest.mark.parametrize("num_segments", _NUM_SEGMENTS)
@pytest.mark.parametrize("num_cols", _NUM_COLUMNS[:3])
@pytest.mark.parametrize("s", [16, 32, 64, 128])
def test_seg_sum_single_cube(ops, dev, num_segments, num_cols, s):
    A, x, indptr = _seg_inputs(num_segments, num_cols, torch.float16, num_segments + s)
    starts = indptr[:-1]
    ones = torch.ones(s, s, dtype=torch.float16, device=dev)
    y = ops.run_seg_sum_single_cube(x.to(dev), torch.triu(ones), torch.tril(ones, -1),
                                    starts.to(torch.uint32).to(dev), s)
    assert y.shape == (num_segments,)
    assert_exact(to_np(y), oracle.seg_sum_indptr(to_np(x), to_np(starts), ends_implicit=True))


@pytest.mark.parametrize("num_segments", _NUM_SEGMENTS)
@pytest.mark.parametrize("s", [32, 64, 128])
def test_seg_sum_multi_cube(ops, dev, num_segments, s):
    A, x, indptr = _seg_inputs(num_segments, 128 * 128 - 13, torch.float16, num_segments * 3 + s)
    ones = torch.ones(s, s, dtype=torch.float16, device=dev)
    y = ops.run_seg_sum_multi_cube(x.to(dev), torch.triu(ones), torch.tril(ones, -1), indptr.to(dev))
    assert_exact(to_np(y), oracle.seg_sum_indptr(to_np(x), to_np(indptr)))
    with pytest.raises(RuntimeError):
        ops.run_seg_sum_multi_cube(x.float().to(dev), torch.triu(ones), torch.tril(ones, -1),
                                   indptr.to(dev))


@pytest.mark.parametrize("case", ["empty_rows", "one_row", "long_rows", "all_empty", "tail_rows"])
def test_seg_sum_edge_cases(ops, dev, case):
    rng = np.random.default_rng(1)
    if case == "empty_rows":
        lens = rng.integers(0, 3, 50_000) * rng.integers(0, 2, 50_000)
    elif case == "one_row":
        lens = np.array([200_003])
    elif case == "long_rows":
        lens = rng.integers(0, 40_000, 300)
    elif case == "all_empty":
        lens = np.zeros(1000, np.int64)
    else:
        lens = np.concatenate([rng.integers(1, 30, 20_000), np.zeros(5000, np.int64)])
    indptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    nnz = int(indptr[-1])
    x = rng.integers(-128, 128, nnz).astype(np.int8)
    y = ops.run_seg_sum_multi_core(torch.from_numpy(x).to(dev), torch.from_numpy(indptr).to(dev), 128)
    assert_exact(to_np(y), oracle.seg_sum_indptr(x, indptr), case)
    xh = rng.standard_normal(nnz).astype(np.float16)
    yh = ops.run_seg_sum_multi_core(torch.from_numpy(xh).to(dev), torch.from_numpy(indptr).to(dev), 128)
    assert_exact(to_np(yh), oracle.seg_sum_indptr(xh, indptr), case)


def test_seg_sum_config4_shape(ops, dev):
    """2,000 segments over int8 data, geometric segment lengths (mean 64)."""
    rng = np.random.default_rng(4)
    lens = rng.geometric(1 / 64, 2000)
    indptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    x = rng.integers(-4, 5, int(indptr[-1])).astype(np.int8)
    y = ops.run_seg_sum_multi_core(torch.from_numpy(x).to(dev), torch.from_numpy(indptr).to(dev), 64)
    assert_exact(to_np(y), oracle.seg_sum_indptr(x, indptr))


@pytest.mark.parametrize("p_empty", [0.5, 0.99, 0.999])
@pytest.mark.parametrize("dtype", [np.int8, np.float16], ids=["i8", "f16"])
def test_seg_sum_tiles_with_many_rows(ops, dev, p_empty, dtype):
    """Mostly empty rows: with p_empty = 0.99 and above a 16K-element int8
    tile overlaps more than 65536 rows, so the kernel finds each row end's
    row by binary search instead of its 16-bit LDS record; every tile size
    and row pattern must give the oracle's sums."""
    rng = np.random.default_rng(int(p_empty * 1000))
    rows = 3_000_000
    lens = np.where(rng.random(rows) < p_empty, 0, rng.geometric(1 / 8, rows))
    lens[-1] = 5
    indptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    nnz = int(indptr[-1])
    x = (rng.integers(-128, 128, nnz) if dtype == np.int8 else rng.standard_normal(nnz)).astype(dtype)
    y = ops.run_seg_sum_multi_core(torch.from_numpy(x).to(dev), torch.from_numpy(indptr).to(dev), 128)
    assert_exact(to_np(y), oracle.seg_sum_indptr(x, indptr), f"p_empty={p_empty} nnz={nnz}")


def test_seg_sum_int8_large_geometric(ops, dev):
    """2^25 int8 values in geometric segments (mean 64): many 16K-element tiles
    with rows spanning tile boundaries."""
    rng = np.random.default_rng(25)
    n = 1 << 25
    lens = rng.geometric(1 / 64, n // 32)
    ends = np.cumsum(lens)
    ends = ends[ends < n]
    indptr = np.concatenate([[0], ends, [n]]).astype(np.int32)
    x = rng.integers(-128, 128, n).astype(np.int8)
    y = ops.run_seg_sum_multi_core(torch.from_numpy(x).to(dev), torch.from_numpy(indptr).to(dev), 128)
    assert_exact(to_np(y), oracle.seg_sum_indptr(x, indptr))


# --- flag segmented sum (tests/test_segmented_sum.py:36-112) -------------------
@pytest.mark.parametrize("multiplier", [1, 2, 3, 5, 6, 7, 8, 9, 10, 11, 12])
def test_seg_sum_flags_s32(ops, dev, multiplier):
    s = 32
    n = multiplier * 20 * s * s
    g = gen(n)
    x = torch.randint(-2, 2, (n,), generator=g).half()
    f = (torch.rand(n, generator=g) < 0.01).to(torch.int8)
    f[0] = 0
    ends = torch.cat([f[1:], torch.ones(1, dtype=torch.int8)])
    y = ops.run_seg_sum(x.to(dev), ends.to(dev), s)
    assert y.dtype == torch.float32
    assert_exact(to_np(y), oracle.seg_sum_flags(to_np(x), to_np(ends)))


@pytest.mark.parametrize("density", [0.0, 0.3, 1.0])
def test_seg_sum_flags_density(ops, dev, density):
    n = 128 * 128 * 3 + 5
    g = gen(int(density * 100) + 1)
    x = torch.randn(n, generator=g).half()
    f = (torch.rand(n, generator=g) < density).to(torch.int8)
    y = ops.run_seg_sum(x.to(dev), f.to(dev), 128)
    assert_exact(to_np(y), oracle.seg_sum_flags(to_np(x), to_np(f)))


# --- csr gather (tests/test_csr_gather.py:28-93) ------------------------------
_GATHER_SIZES = [16 * 1024, 61 * 1024 - 13, 63 * 1024 - 7, 80 * 1024 - 63, 1024 * 1024 - 31]


@pytest.mark.parametrize("col_len", _GATHER_SIZES)
@pytest.mark.parametrize("x_len", [256, 1024, 70000, 1 << 22])
@pytest.mark.parametrize("dtype", [torch.int16, torch.float16, torch.float32], ids=str)
def test_csr_gather(ops, dev, col_len, x_len, dtype):
    g = gen(col_len + x_len)
    if dtype == torch.int16:
        vals = torch.randint(-3000, 3000, (col_len,), generator=g, dtype=dtype)
        x = torch.randint(-3000, 3000, (x_len,), generator=g, dtype=dtype)
    else:
        vals = torch.randn(col_len, generator=g).to(dtype)
        x = torch.randn(x_len, generator=g).to(dtype)
    cols = torch.randint(0, x_len, (col_len,), generator=g, dtype=torch.int32)
    z = ops.run_csr_gather(vals.to(dev), cols.to(dev), x.to(dev))
    assert z.dtype == dtype and z.shape == (col_len,)
    assert_exact(to_np(z), oracle.csr_gather(to_np(vals), to_np(cols), to_np(x)))
    assert torch.equal(z.cpu(), vals * x[cols.long()])


def test_csr_gather_rejects_bad_args(ops, dev):
    v = torch.ones(8, dtype=torch.float16, device=dev)
    with pytest.raises(RuntimeError):
        ops.run_csr_gather(v, torch.zeros(7, dtype=torch.int32, device=dev), v)
    with pytest.raises(RuntimeError):
        ops.run_csr_gather(v, torch.zeros(8, dtype=torch.int64, device=dev), v)
    with pytest.raises(RuntimeError):
        ops.run_csr_gather(v, torch.zeros(8, dtype=torch.int32, device=dev), v.float())


# --- SpMV v2 (tests/test_spmv_v2.py:30-114) ------------------------------------
@pytest.mark.parametrize("s", [32, 128])
@pytest.mark.parametrize("density", [0.01, 0.001, 0.0001])
@pytest.mark.parametrize("nrow", [1024, 1024 * 3, 1024 * 8, 1024 * 16])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float16], ids=str)
def test_spmv_v2(ops, dev, s, density, nrow, dtype):
    B = sparse_random(nrow - 1, nrow - 1, density, nrow + int(1 / density))
    vec = np.random.default_rng(42).uniform(1, 9, nrow - 1)
    vals = torch.from_numpy(B.data.astype(np.float32)).to(dtype)
    xv = torch.from_numpy(vec.astype(np.float32)).to(dtype)
    indptr = torch.from_numpy(B.indptr.astype(np.uint32))
    cols = torch.from_numpy(B.indices.astype(np.int32))
    y = ops.run_spmv_v2(vals.to(dev), indptr.to(dev), cols.to(dev), xv.to(dev), s)
    assert y.dtype == torch.float32 and y.shape == (nrow - 1,)
    exp = oracle.spmv(to_np(vals), to_np(indptr), to_np(cols), to_np(xv))
    # Products are exact and summed in double: at most a last-place difference
    # from the float64 oracle when a row sum is not exact in double.
    np.testing.assert_allclose(to_np(y), exp, rtol=1e-6, atol=1e-6)
    ref = B @ vec.astype(np.float32)
    np.testing.assert_allclose(to_np(y), ref, atol=1e-1 if dtype == torch.float32 else 0.5)


def test_spmv_lognormal_rows(ops, dev):
    """Lognormal row lengths (mean 16, sigma 1.2, capped at 20000), as the
    SpMV workload, at 2^16 rows."""
    rng = np.random.default_rng(7)
    rows = 1 << 16
    lens = oracle.lognormal_row_lengths(rows, 16.0, 1.2, 20000, rng)
    lens[5] = 20000
    indptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    nnz = int(indptr[-1])
    row_id = np.repeat(np.arange(rows), lens)
    cols = rng.integers(0, rows, nnz)
    cols = cols[np.lexsort((cols, row_id))].astype(np.int32)  # sorted within each row
    vals = (2 * rng.random(nnz) - 1).astype(np.float16)
    x = (2 * rng.random(rows) - 1).astype(np.float16)
    y = ops.run_spmv_v2(torch.from_numpy(vals).to(dev), torch.from_numpy(indptr).to(dev),
                        torch.from_numpy(cols).to(dev), torch.from_numpy(x).to(dev), 128)
    np.testing.assert_allclose(to_np(y), oracle.spmv(vals, indptr, cols, x), rtol=1e-6, atol=1e-6)


def test_spmv_rejects_bad_dtypes(ops, dev):
    v = torch.ones(4, dtype=torch.float16, device=dev)
    ip = torch.tensor([0, 2, 4], dtype=torch.int32, device=dev)
    c = torch.zeros(4, dtype=torch.int32, device=dev)
    with pytest.raises(RuntimeError):
        ops.run_spmv_v2(v, ip, c, v.float(), 32)
    with pytest.raises(RuntimeError):
        ops.run_spmv_v2(v, ip.long(), c, v, 32)


# --- compaction (tests/test_compress*.py) --------------------------------------
_OFFSETS = [-127, -13, 13, 65, 129]


def _cmp_inputs(n, dtype, seed):
    g = gen(seed)
    x = torch.randint(0, 2**7 - 1, (n,), generator=g).to(dtype)
    mask = (torch.randn(n, generator=g) > 0).to(torch.int8)
    return x, mask


@pytest.mark.parametrize("s,mults", [(32, (1, 4, 11)), (64, (1, 3, 7)), (128, (1, 2, 5))])
@pytest.mark.parametrize("dtype", [torch.float16, torch.float32], ids=str)
@pytest.mark.parametrize("offset", [0] + _OFFSETS)
def test_compress_family(ops, dev, s, mults, dtype, offset):
    for mlt in mults:
        n = mlt * 20 * s * s - offset
        x, mask = _cmp_inputs(n, dtype, n)
        xd, md = x.to(dev), mask.to(dev)
        exp = torch.masked_select(x, mask.bool())
        z = ops.run_compress(xd, md, s)
        assert torch.equal(z.cpu(), exp)
        idx_in = torch.arange(n, dtype=torch.int32)
        zv, zi = ops.run_compress_ind(xd, idx_in.to(dev), md, s)
        assert torch.equal(zv.cpu(), exp) and torch.equal(zi.cpu(), torch.masked_select(idx_in, mask.bool()))
        zv2, zi2 = ops.run_compress_ind_no_arange(xd, md, s)
        assert torch.equal(zv2.cpu(), exp) and zi2.dtype == torch.int32
        assert torch.equal(zi2.cpu(), torch.nonzero(mask).flatten().int())
        pos = torch.sum(mask).to(torch.int32)
        zp = ops.run_compress_pos(xd, md, pos, s)
        assert torch.equal(zp.cpu(), exp)


def test_compress_bytes_and_indices_dtypes(ops, dev):
    n = 100_003
    g = gen(9)
    mask = (torch.rand(n, generator=g) < 0.37).to(torch.int8)
    for dt in (torch.int8, torch.int16, torch.bfloat16, torch.int32, torch.float64, torch.int64):
        x = torch.randint(-100, 100, (n,), generator=g).to(dt)
        z = ops.run_compress(x.to(dev), mask.to(dev), 64)
        assert torch.equal(z.cpu(), x[mask.bool()])
    idx = torch.randint(0, 1 << 40, (n,), generator=g, dtype=torch.int64)
    zv, zi = ops.run_compress_ind(torch.ones(n).to(dev), idx.to(dev), mask.to(dev), 64)
    assert zi.dtype == torch.int64 and torch.equal(zi.cpu(), idx[mask.bool()])


def test_compress_mask_keeps_only_ones(ops, dev):
    x = torch.arange(6, dtype=torch.float32)
    m = torch.tensor([1, 2, 0, 1, -1, 1], dtype=torch.int8)
    assert ops.run_compress(x.to(dev), m.to(dev), 32).cpu().tolist() == [0.0, 3.0, 5.0]


def test_compress_edge_densities(ops, dev):
    n = 1_000_003
    x = torch.arange(n, dtype=torch.int32)
    for p in (0.0, 1.0, 1e-6):
        m = (torch.rand(n, generator=gen(1)) < p).to(torch.int8) if 0 < p < 1 else \
            torch.full((n,), int(p), dtype=torch.int8)
        z, i = ops.run_compress_ind_no_arange(x.to(dev), m.to(dev), 128)
        assert torch.equal(z.cpu(), x[m.bool()]) and torch.equal(i.cpu(), x[m.bool()])


def test_compress_ind_full_size_2p28(ops, dev):
    """The compaction workload at its full size: N = 2^28 fp32 values
    U[0,1), int8 Bernoulli(0.3) mask, arange indices; against the oracle."""
    n = 1 << 28
    g = gen(0)
    x = torch.rand(n, generator=g)
    mask = (torch.rand(n, generator=g) < 0.3).to(torch.int8)
    idx = torch.arange(n, dtype=torch.int32)
    zv, zi = ops.run_compress_ind(x.to(dev), idx.to(dev), mask.to(dev), 128)
    ev, ei = oracle.compress_ind(to_np(x), to_np(idx), to_np(mask))
    assert_exact(to_np(zv), ev)
    assert_exact(to_np(zi), ei)
