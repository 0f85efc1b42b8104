"""Filters (x[x >= p], x[x <= p]), `where`, count_if and searchsorted.

CPU: the oracle is pinned against torch's own operators on the CPU
(masked_select, argwhere, count_nonzero of the six comparisons,
searchsorted) and numpy's searchsorted, on the reference tests' data
(integers 0..126 with pivot 13.0, tests/test_filter_greater_equal.py:40-54),
on ties at a pivot that the input type cannot hold, on NaN/inf, and on the
hand-checked searchsorted edge vector of tests/test_searchsorted.py:58-63.

GPU: the HIP operators against the oracle, bit for bit, on the reference
tests' length grids (multiples of 20 * S * S, fp32 with ragged offsets),
on every input type, on empty / all / none kept, on unaligned views and at
2^26 elements.
"""
import numpy as np
import pytest
import torch

import oracle
from _util import assert_exact, gen, to_np

OFFSETS = [-127, -13, 13, 65, 129]  # tests/test_filter_greater_equal.py:27
TORCH_CMP = [torch.lt, torch.gt, torch.eq, torch.le, torch.ge, torch.ne]


def lengths(s, max_iters):
    # tests/test_filter_greater_equal.py:30-33 (20 AI cores there)
    return [m * 20 * s * s for m in range(1, max_iters)]


def ints_0_126(n, dtype, seed):
    return torch.randint(0, 2**7 - 1, (n,), generator=gen(seed)).to(dtype)


def mixed(n, dtype, seed, pivot):
    """Random normals, exact ties at the pivot (as the type rounds it),
    +-inf and NaN."""
    x = torch.randn(n, generator=gen(seed)).to(dtype)
    if n >= 8:
        pick = torch.randint(0, n, (max(n // 50, 4),), generator=gen(seed + 1))
        x[pick] = torch.tensor(pivot, dtype=torch.float32).to(dtype)
        x[pick[:1]] = float("inf")
        x[pick[1:2]] = float("-inf")
        x[pick[2:3]] = float("nan")
    return x


# ---------------------------------------------------------------------------
# Oracle pinning (CPU)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("dtype", [torch.float16, torch.bfloat16, torch.float32], ids=str)
@pytest.mark.parametrize("pivot", [13.0, 0.1, -0.7, 1e-3])
def test_oracle_filters_match_torch(dtype, pivot):
    for seed, x in enumerate([ints_0_126(5000, dtype, 1), mixed(5000, dtype, 2, pivot)]):
        xn = to_np(x) if dtype != torch.bfloat16 else None
        ge = torch.masked_select(x, x >= pivot)
        le = torch.masked_select(x, x <= pivot)
        pos = torch.argwhere(x >= pivot).flatten().to(torch.int32)
        if dtype == torch.bfloat16:
            # numpy has no bfloat16: compare in float32, exact for bf16 values
            # once the pivot is rounded to bf16 as the operator rounds it.
            pb = float(torch.tensor(pivot, dtype=torch.float32).to(dtype))
            xf = x.float().numpy()
            assert_exact(oracle.filter_cmp(xf, pb, oracle.CMP_GE), ge.float().numpy())
            assert_exact(oracle.filter_cmp(xf, pb, oracle.CMP_LE), le.float().numpy())
            assert_exact(oracle.filter_cmp(xf, pb, oracle.CMP_GE, positions=True), pos.numpy())
            continue
        assert_exact(oracle.filter_cmp(xn, pivot, oracle.CMP_GE), to_np(ge), f"ge seed={seed}")
        assert_exact(oracle.filter_cmp(xn, pivot, oracle.CMP_LE), to_np(le), f"le seed={seed}")
        assert_exact(oracle.filter_cmp(xn, pivot, oracle.CMP_GE, positions=True), to_np(pos))


def test_oracle_pivot_is_rounded_to_the_input_type():
    # half(0.1) = 0.0999755859375 < 0.1: it passes both x >= 0.1 and x <= 0.1
    # in half, as torch compares a half tensor with a Python float.
    h = np.array([0.1, 0.2, 0.05], dtype=np.float16)
    assert oracle.filter_cmp(h, 0.1, oracle.CMP_GE).tolist() == [h[0], h[1]]
    assert oracle.filter_cmp(h, 0.1, oracle.CMP_LE).tolist() == [h[0], h[2]]
    assert bool((torch.from_numpy(h[:1]) >= 0.1).item()) and bool((torch.from_numpy(h[:1]) <= 0.1).item())
    # float32 holds 0.1 as 0.100000001490116: float(0.1) itself compares equal.
    f = np.array([np.float32(0.1)], dtype=np.float32)
    assert oracle.count_if(f, 0.1, oracle.CMP_EQ) == 1


@pytest.mark.parametrize("dtype", [torch.float16, torch.float32], ids=str)
def test_oracle_count_if_every_mode_matches_torch(dtype):
    for pivot in (0.1, 0.0, -1.5):
        x = mixed(4099, dtype, 7, pivot)
        for mode, fn in enumerate(TORCH_CMP):
            exp = int(torch.count_nonzero(fn(x, pivot)))
            assert oracle.count_if(to_np(x), pivot, mode) == exp, (mode, pivot)
    # NaN counts only for "!="
    nan = np.array([np.nan, 1.0], dtype=np.float32)
    assert [oracle.count_if(nan, 1.0, m) for m in range(6)] == [0, 0, 1, 1, 1, 1]


def test_oracle_searchsorted_known_answer_and_libraries(rng):
    a = np.array([0, 0, 2, 2, 2, 5, 9, 9, 10], dtype=np.int32)
    v = np.array([-1, 0, 1, 2, 3, 5, 8, 9, 10, 11], dtype=np.int32)
    assert oracle.searchsorted(a, v).tolist() == [0, 0, 2, 2, 5, 5, 6, 6, 8, 9]
    assert oracle.searchsorted(np.zeros(0, np.int32), v).tolist() == [0] * v.size
    for n in (1, 2, 7, 1024, 65537):
        a = np.sort(rng.integers(-50, 50, n)).astype(np.int32)
        v = rng.integers(-60, 60, 500).astype(np.int32)
        got = oracle.searchsorted(a, v)
        assert_exact(got, np.searchsorted(a, v, side="left").astype(np.int32))
        assert_exact(got, torch.searchsorted(torch.from_numpy(a), torch.from_numpy(v), out_int32=True).numpy())
        # brute force: the number of elements below each needle
        assert_exact(got, (a[None, :] < v[:, None]).sum(1).astype(np.int32))
    big = np.array([-2**31, -1, 0, 2**31 - 1], dtype=np.int32)
    assert oracle.searchsorted(big, big).tolist() == [0, 1, 2, 3]


def test_oracle_catches_a_planted_off_by_one(rng):
    # upper bound instead of lower bound must fail the brute-force check
    a = np.sort(rng.integers(0, 5, 100)).astype(np.int32)
    v = rng.integers(0, 5, 50).astype(np.int32)
    wrong = (a[None, :] <= v[:, None]).sum(1).astype(np.int32)
    assert not np.array_equal(oracle.searchsorted(a, v), wrong)


# ---------------------------------------------------------------------------
# HIP operators (GPU)
# ---------------------------------------------------------------------------

def _check_filters(ops, dev, x, pivot, what):
    xd = x.to(dev)
    xn = to_np(x)
    if x.dtype == torch.bfloat16:
        pivot_np = float(torch.tensor(pivot, dtype=torch.float32).to(torch.bfloat16))
    else:
        pivot_np = pivot
    for fn, mode in ((ops.run_filter_greater_eq, oracle.CMP_GE), (ops.run_filter_less_eq, oracle.CMP_LE)):
        got = fn(xd, float(pivot), 32)
        assert got.dtype == x.dtype and got.device == xd.device
        assert_exact(to_np(got), oracle.filter_cmp(xn, pivot_np, mode), f"{what} mode={mode}")
    got = ops.run_where(xd, float(pivot), 32)
    assert got.dtype == torch.int32
    assert_exact(to_np(got), oracle.filter_cmp(xn, pivot_np, oracle.CMP_GE, positions=True), f"{what} where")


@pytest.mark.gpu
@pytest.mark.parametrize("S", [32, 64, 128])
def test_filters_reference_grid_fp16(ops, dev, S):
    # tests/test_filter_greater_equal.py:56-68 and tests/test_where.py:52-64
    for n in lengths(S, {32: 12, 64: 8, 128: 6}[S])[::2]:
        x = ints_0_126(n, torch.float16, n)
        xd = x.to(dev)
        exp = torch.masked_select(xd, xd >= 13.0)
        got = ops.run_filter_greater_eq(xd, 13.0, S)
        assert torch.equal(got, exp), n
        assert_exact(to_np(ops.run_where(xd, 13.0, S)),
                     oracle.filter_cmp(to_np(x), 13.0, oracle.CMP_GE, positions=True), f"where n={n}")
        assert_exact(to_np(ops.run_filter_less_eq(xd, 13.0, S)),
                     oracle.filter_cmp(to_np(x), 13.0, oracle.CMP_LE), f"le n={n}")


@pytest.mark.gpu
@pytest.mark.parametrize("offset", OFFSETS)
def test_filters_reference_grid_fp32_offsets(ops, dev, offset):
    # tests/test_filter_greater_equal.py:71-81
    for n in lengths(32, 12)[::3] + lengths(64, 8)[::3]:
        _check_filters(ops, dev, ints_0_126(n - offset, torch.float32, n + offset), 13.0, f"n={n - offset}")


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [torch.float16, torch.bfloat16, torch.float32], ids=str)
@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 1000, 4096, 4097, 65535, 65536 + 3, 1_000_003])
def test_filters_mixed_values(ops, dev, dtype, n):
    for pivot in (0.1, -0.7):
        _check_filters(ops, dev, mixed(n, dtype, n, pivot), pivot, f"n={n} pivot={pivot}")


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [torch.float16, torch.float32], ids=str)
def test_filters_all_none_empty_unaligned(ops, dev, dtype):
    x = ints_0_126(100_000, dtype, 5).to(dev)
    assert torch.equal(ops.run_filter_greater_eq(x, -1.0, 64), x)                  # all kept
    assert ops.run_filter_greater_eq(x, 1000.0, 64).numel() == 0                  # none kept
    assert torch.equal(ops.run_where(x, -1.0, 64), torch.arange(x.numel(), device=dev, dtype=torch.int32))
    assert ops.run_where(x, 1000.0, 64).numel() == 0
    e = torch.empty(0, dtype=dtype, device=dev)
    assert ops.run_filter_less_eq(e, 1.0, 64).numel() == 0
    assert ops.run_where(e, 1.0, 64).dtype == torch.int32
    # a view one element in: not 16-byte aligned
    v = x[1:]
    _check_filters(ops, dev, v.cpu(), 13.0, "aligned copy")
    assert torch.equal(ops.run_filter_greater_eq(v, 13.0, 64), torch.masked_select(v, v >= 13.0))
    assert torch.equal(ops.run_where(v, 13.0, 64), torch.argwhere(v >= 13.0).flatten().int())


@pytest.mark.gpu
def test_filters_at_scale(ops, dev):
    # 2^26 + 7 fp16 values: the outputs equal torch's masked_select / argwhere
    n = (1 << 26) + 7
    x = torch.randn(n, generator=gen(9)).half().to(dev)
    for p in (0.0, 1.5):
        assert torch.equal(ops.run_filter_greater_eq(x, p, 128), torch.masked_select(x, x >= p))
        assert torch.equal(ops.run_filter_less_eq(x, p, 128), torch.masked_select(x, x <= p))
        assert torch.equal(ops.run_where(x, p, 128), torch.argwhere(x >= p).flatten().int())


@pytest.mark.gpu
def test_filters_reject(ops, dev):
    with pytest.raises(RuntimeError):
        ops.run_where(torch.zeros(10, dtype=torch.int32, device=dev), 1.0, 32)
    with pytest.raises(RuntimeError):
        ops.run_filter_greater_eq(torch.zeros(10), 1.0, 32)  # host tensor
    with pytest.raises(RuntimeError):
        ops.run_filter_greater_eq(torch.zeros(10, device=dev), 1.0, 0)


# tests/test_count_if.py:31-44
COUNT_SIZES = [1024 * 8, 1024 * 16, 2048 * 32, 1024 * 128, 1024 * 8 - 16, 1024 * 16 - 16, 2048 * 32 - 16,
               1024 * 128 - 16, 1024 * 8 + 16, 1024 * 16 + 16, 2048 * 32 + 16, 1024 * 128 + 15]


@pytest.mark.gpu
@pytest.mark.parametrize("tile_len", [128, 1024])
def test_count_if_reference_grid(ops, dev, tile_len):
    for n in COUNT_SIZES:
        x = torch.randn(n, generator=gen(n)).half()
        got = ops.run_count_if(x.to(dev), 0.1, tile_len, oracle.CMP_LE)
        assert got.dtype == torch.int32 and got.shape == (1,)
        assert int(got.item()) == oracle.count_if(to_np(x), 0.1, oracle.CMP_LE), n


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [torch.float16, torch.bfloat16, torch.float32], ids=str)
def test_count_if_every_mode(ops, dev, dtype):
    for n in (0, 1, 777, 300_001):
        x = mixed(n, dtype, n + 3, 0.1)
        xd = x.to(dev)
        for mode, fn in enumerate(TORCH_CMP):
            got = int(ops.run_count_if(xd, 0.1, 1024, mode).item())
            assert got == int(torch.count_nonzero(fn(xd, 0.1))), (n, mode)
            if dtype != torch.bfloat16:
                assert got == oracle.count_if(to_np(x), 0.1, mode)
    with pytest.raises(RuntimeError):
        ops.run_count_if(torch.zeros(4, device=dev), 0.1, 1024, 6)


def _search(ops, dev, a, v):
    got = ops.run_searchsorted(torch.from_numpy(a).to(dev), torch.from_numpy(v).to(dev))
    assert got.dtype == torch.int32 and got.shape == (v.size,)
    return to_np(got)


@pytest.mark.gpu
@pytest.mark.parametrize("data_len", [1, 8, 127, 1024, 65537])
@pytest.mark.parametrize("num_values", [1, 3, 21, 64])
def test_searchsorted_random(ops, dev, data_len, num_values):
    # tests/test_searchsorted.py:48-56
    r = np.random.default_rng(data_len * 100 + num_values)
    a = np.sort(r.integers(-50, 50, data_len)).astype(np.int32)
    v = r.integers(-60, 60, num_values).astype(np.int32)
    assert_exact(_search(ops, dev, a, v), oracle.searchsorted(a, v))


@pytest.mark.gpu
def test_searchsorted_edges_and_indptr(ops, dev):
    # tests/test_searchsorted.py:58-85
    a = np.array([0, 0, 2, 2, 2, 5, 9, 9, 10], dtype=np.int32)
    v = np.array([-1, 0, 1, 2, 3, 5, 8, 9, 10, 11], dtype=np.int32)
    assert _search(ops, dev, a, v).tolist() == [0, 0, 2, 2, 5, 5, 6, 6, 8, 9]
    r = np.random.default_rng(4)
    nnz = 100000
    indptr = np.concatenate([[0], np.sort(r.integers(0, nnz, 4000))]).astype(np.int32)
    st = np.array([0, nnz // 4, nnz // 2, (3 * nnz) // 4, nnz], dtype=np.int32)
    assert_exact(_search(ops, dev, indptr, st), oracle.searchsorted(indptr, st))
    # empty haystack and no needles; int32 extremes
    assert _search(ops, dev, np.zeros(0, np.int32), v).tolist() == [0] * v.size
    assert _search(ops, dev, a, np.zeros(0, np.int32)).size == 0
    big = np.array([-2**31, -1, 0, 2**31 - 1], dtype=np.int32)
    assert _search(ops, dev, big, big).tolist() == [0, 1, 2, 3]
    with pytest.raises(RuntimeError):
        ops.run_searchsorted(torch.zeros(4, dtype=torch.int64, device=dev), torch.zeros(4, dtype=torch.int32, device=dev))


@pytest.mark.gpu
@pytest.mark.parametrize("data_len", [16383, 16384, 16385, 32769, 1_000_003])
def test_searchsorted_sample_windows(ops, dev, data_len):
    # The kernel brackets each needle between two of every B-th haystack
    # element staged in LDS, B = ceil(len / 16384): runs of equal values that
    # straddle those positions, needles equal to, just below and just above
    # haystack values, and more needles than one sweep of the grid covers.
    r = np.random.default_rng(data_len)
    a = np.sort(r.integers(-3000, 3000, data_len) * 7).astype(np.int32)
    picks = a[r.integers(0, data_len, 1_500_000)]
    v = (picks + r.integers(-1, 2, picks.size)).astype(np.int32)
    v[:3] = [a[0] - 1, a[-1], a[-1] + 1]
    assert_exact(_search(ops, dev, a, v), oracle.searchsorted(a, v))


@pytest.mark.gpu
def test_searchsorted_large_haystack_with_ties(ops, dev):
    # 2^22 needles into a 2^24 haystack with ties (B = 1024-element windows
    # below the LDS samples); every needle in one window; every needle below
    # the haystack; against the oracle.
    r = np.random.default_rng(21)
    a = np.sort(r.integers(-10**6, 10**6, 1 << 24)).astype(np.int32)
    v = np.concatenate([a[r.integers(0, a.size, 1 << 21)], r.integers(-10**6 - 5, 10**6 + 5, 1 << 21)])
    v = v.astype(np.int32)
    assert_exact(_search(ops, dev, a, v), oracle.searchsorted(a, v), "2^22 needles into 2^24")
    one = (a[5_000_000] + r.integers(-1, 2, 1 << 20)).astype(np.int32)
    assert_exact(_search(ops, dev, a, one), oracle.searchsorted(a, one), "one window")
    low = np.full(1 << 17, a[0] - 3, dtype=np.int32)
    assert _search(ops, dev, a, low).max() == 0


@pytest.mark.gpu
def test_searchsorted_at_scale(ops, dev):
    r = np.random.default_rng(8)
    a = np.sort(r.integers(-2**31, 2**31 - 1, (1 << 24) + 5)).astype(np.int32)
    v = r.integers(-2**31, 2**31 - 1, 1 << 20).astype(np.int32)
    assert_exact(_search(ops, dev, a, v), oracle.searchsorted(a, v))
