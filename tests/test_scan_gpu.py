"""GPU parity of the scan family against the oracle.

Follows the reference tests' grids (tests/test_scan_multi_core.py:27-67,
tests/test_scan_single_core.py:63-82, tests/test_scan_batch.py:27-68,
tests/test_row_scan.py:27-47, tests/test_block_scan.py:27-53,
tests/test_seg_scan_single_core.py) and tightens their tolerances: fp16
and fp32 inputs whose exact prefix fits in a double are compared bit for bit
with the float64 oracle rounded once to float32 (the kernels accumulate in
double), int8 results are exact.  The reference's own tolerance against
torch.cumsum is asserted as well.
"""
import numpy as np
import pytest
import torch

import oracle
from _util import assert_exact, gen, to_np, uniform_pm1

pytestmark = pytest.mark.gpu

NUM_AI_CORES = 20  # the unit of the reference's length grids


def mc_lengths(s, max_iters):
    return [m * NUM_AI_CORES * s * s for m in range(1, max_iters)]


def _mc_input(vec_len, dtype, seed):
    if dtype == torch.float16:
        return torch.randn(vec_len, generator=gen(seed)).half()
    return torch.randint(0, 10, (vec_len,), generator=gen(seed), dtype=torch.int8)


@pytest.mark.parametrize("dtype", [torch.int8, torch.float16], ids=str)
@pytest.mark.parametrize("s", [32, 64, 128])
@pytest.mark.parametrize("vec_len", mc_lengths(16, 16))
def test_scan_multi_core(ops, dev, dtype, s, vec_len):
    x = _mc_input(vec_len, dtype, vec_len + s)
    y = ops.run_scan_multi_core(x.to(dev), s)
    out_dtype = torch.float32 if dtype == torch.float16 else torch.int32
    assert y.dtype == out_dtype and y.shape == (vec_len,)
    assert_exact(to_np(y), oracle.scan(to_np(x)), f"scan {dtype} s={s} n={vec_len}")
    ref = torch.cumsum(x, 0, dtype=out_dtype)
    assert torch.allclose(y.cpu(), ref, atol=1e-3, rtol=1e-7)


@pytest.mark.parametrize("vec_len", mc_lengths(16, 12))
def test_scan_multi_core_fp16_s16(ops, dev, vec_len):
    x = _mc_input(vec_len, torch.float16, vec_len)
    y = ops.run_scan_multi_core(x.to(dev), 16)
    assert_exact(to_np(y), oracle.scan(to_np(x)))


@pytest.mark.parametrize("dtype", [torch.int8, torch.float16], ids=str)
@pytest.mark.parametrize("s", [32, 128])
def test_scan_multi_core_no_l2(ops, dev, dtype, s):
    for vec_len in mc_lengths(16, 16)[::5]:
        x = _mc_input(vec_len, dtype, 7 * vec_len)
        y = ops.run_scan_multi_core_no_l2(x.to(dev), s)
        assert_exact(to_np(y), oracle.scan(to_np(x)))


@pytest.mark.parametrize("n", [1, 2, 15, 16, 17, 63, 64, 65, 1023, 1024, 1025, 4095, 4097,
                               16383, 16384, 16385, 65536 * 3 + 5, 1_000_003])
@pytest.mark.parametrize("dtype", [torch.float16, torch.bfloat16, torch.int8], ids=str)
def test_scan_ragged_lengths(ops, dev, n, dtype):
    if dtype == torch.int8:
        x = torch.randint(-128, 127, (n,), generator=gen(n), dtype=torch.int8)
    else:
        x = uniform_pm1(n, dtype, n)
    for s in (32, 64, 128):
        y = ops.run_scan_multi_core(x.to(dev), s)
        exp = oracle.scan(to_np(x))
        assert_exact(to_np(y), exp, f"n={n} s={s} {dtype}")


def test_scan_empty(ops, dev):
    y = ops.run_scan_multi_core(torch.empty(0, dtype=torch.float16, device=dev), 128)
    assert y.shape == (0,) and y.dtype == torch.float32


def test_scan_noncontiguous_and_offset_views(ops, dev):
    base = uniform_pm1(100_001, torch.float16, 3).to(dev)
    for view in (base[1:], base[::2], base[7:90_007]):
        y = ops.run_scan_multi_core(view, 64)
        assert_exact(to_np(y), oracle.scan(to_np(view)))


def test_scan_rejects_bad_dtype(ops, dev):
    with pytest.raises(RuntimeError):
        ops.run_scan_multi_core(torch.zeros(10, dtype=torch.float64, device=dev), 64)
    with pytest.raises(RuntimeError):
        ops.run_scan_multi_core(torch.zeros(10, dtype=torch.float16, device=dev), 0)


@pytest.mark.parametrize("dtype", [torch.float16, torch.bfloat16, torch.int8], ids=str)
def test_scan_full_size_2p28(ops, dev, dtype):
    """The large-scan workload at its full size, against the full oracle."""
    n = 1 << 28
    if dtype != torch.int8:
        x = uniform_pm1(n, dtype, 0)
    else:
        x = torch.randint(-4, 5, (n,), generator=gen(0), dtype=torch.int8)
    y = ops.run_scan_multi_core(x.to(dev), 128)
    torch.cuda.synchronize()
    assert_exact(to_np(y), oracle.scan(to_np(x)), f"2^28 {dtype}")


# --- run_scan_single_core (tests/test_scan_single_core.py:63-82) -----------
@pytest.mark.parametrize("multiplier", [1, 2, 5, 12])
@pytest.mark.parametrize("s", [32, 64, 128])
@pytest.mark.parametrize("dtype", [torch.int8, torch.float16, torch.float32], ids=str)
@pytest.mark.parametrize("offset", [-113, -53, -7, -1, 1, 3, 17, 113])
def test_scan_single_core(ops, dev, multiplier, s, dtype, offset):
    n = multiplier * s * s + offset
    hi = 2 if dtype == torch.int8 else 6
    x = torch.randint(0, hi, (n,), generator=gen(n), dtype=torch.int8).to(dtype)
    y = ops.run_scan_single_core(x.to(dev), s)
    assert y.dtype == (torch.int32 if dtype == torch.int8 else torch.float32)
    assert_exact(to_np(y), oracle.scan(to_np(x)))


@pytest.mark.parametrize("dtype", [torch.int8, torch.float16, torch.float32], ids=str)
@pytest.mark.parametrize("start", [0, 5, -3, 1234.5])
def test_scan_single_core_starting_sum(ops, dev, dtype, start):
    n = 3 * 128 * 128 + 17
    x = torch.randint(-3, 4, (n,), generator=gen(11), dtype=torch.int8).to(dtype)
    y = ops.run_scan_single_core(x.to(dev), 128, start)
    assert_exact(to_np(y), oracle.scan(to_np(x), start))
    # starting_sum is a keyword with default 0, as in the reference binding
    y0 = ops.run_scan_single_core(x.to(dev), 128)
    assert_exact(to_np(y0), oracle.scan(to_np(x)))


# --- run_scan_batch (tests/test_scan_batch.py:27-68) ------------------------
def _batch_cases():
    for s in (32, 64, 128):
        for mult in (1, 2, 3, 11, 63):
            for b in (2, 16, 20, 32, 44, 256):
                for L in (mult * s * s, mult * s * s // 2, mult * s * s // 4,
                          mult * s * s - 1, mult * s * s + 1):
                    if b * L <= (1 << 22):
                        yield s, L, b


@pytest.mark.parametrize("s,vec_len,batch", list(_batch_cases()))
@pytest.mark.parametrize("dtype", [torch.float32, torch.float16], ids=str)
def test_scan_batch(ops, dev, s, vec_len, batch, dtype):
    x = torch.rand((batch, vec_len), generator=gen(batch * vec_len), dtype=torch.float32).to(dtype)
    y = ops.run_scan_batch(x.to(dev), s)
    assert y.shape == (batch, vec_len) and y.dtype == torch.float32
    assert_exact(to_np(y).reshape(-1), oracle.scan_rows(to_np(x).reshape(-1), vec_len))
    ref = torch.cumsum(x.float(), dim=1)
    assert torch.allclose(y.cpu(), ref, rtol=1e-3)


def test_scan_batch_large_rows_property(ops, dev):
    """A batch whose total exceeds the case cap: every row's last element
    equals the row sum and a row equals the 1-D scan of that row."""
    x = uniform_pm1(256 * 63 * 128 * 128 // 16, torch.float16, 5).reshape(256, -1).to(dev)
    y = ops.run_scan_batch(x, 128)
    last = y[:, -1].double().cpu()
    assert torch.equal(last.float(), x.double().sum(1).float().cpu())
    for r in (0, 100, 255):
        assert torch.equal(y[r], ops.run_scan_multi_core(x[r].contiguous(), 128))


# --- run_row_scan (tests/test_row_scan.py:27-47) ----------------------------
@pytest.mark.parametrize("m", [20 * 128 * i - 13 for i in range(1, 10)])
def test_row_scan(ops, dev, m):
    S = 128
    a = torch.ones(m, S, dtype=torch.float16)
    y = ops.run_row_scan(a.to(dev), S)
    assert y.shape == (m * S,) and y.dtype == torch.float32
    tril = torch.tril(torch.ones(S, S))
    assert torch.equal(y.cpu(), torch.nn.functional.linear(a.float(), tril).flatten())
    r = torch.randn(m, S, generator=gen(m)).half()
    assert_exact(to_np(ops.run_row_scan(r.to(dev), S)), oracle.scan_rows(to_np(r), S))


@pytest.mark.parametrize("s", [16, 32, 64, 128])
def test_row_scan_other_s(ops, dev, s):
    r = torch.randn(777, s, generator=gen(s)).half()
    assert_exact(to_np(ops.run_row_scan(r.to(dev), s)), oracle.scan_rows(to_np(r), s))


# --- run_block_scan (tests/test_block_scan.py:27-53) ------------------------
# S = 128 takes m in 10..12 here; its reference grid (m in 20..24) is
# test_block_scan_s128.
@pytest.mark.parametrize("m,s", [(m, s) for s in (16, 32, 64, 128) for m in range(10, 21) if s < 128 or m <= 12])
def test_block_scan(ops, dev, m, s):
    t = s * s
    a = (0.1 * torch.randn(m, t, generator=gen(m * s))).half()
    ones = torch.ones(s, s, dtype=torch.float16, device=dev)
    y = ops.run_block_scan(a.flatten().to(dev), torch.triu(ones), torch.tril(ones, -1))
    assert y.dtype == torch.float32
    assert_exact(to_np(y), oracle.scan_rows(to_np(a).reshape(-1), t))
    ref = torch.matmul(a.float(), torch.triu(torch.ones(t, t))).flatten()
    assert torch.allclose(y.cpu(), ref, atol=1e-2, rtol=1e-3)


@pytest.mark.parametrize("m", list(range(20, 25)))
def test_block_scan_s128(ops, dev, m):
    s = 128
    a = (0.1 * torch.randn(m, s * s, generator=gen(m))).half()
    ones = torch.ones(s, s, dtype=torch.float16, device=dev)
    y = ops.run_block_scan(a.flatten().to(dev), torch.triu(ones), torch.tril(ones, -1))
    assert_exact(to_np(y), oracle.scan_rows(to_np(a).reshape(-1), s * s))


def test_scan_multi_cube(ops, dev):
    for s in (32, 64, 128):
        x = uniform_pm1(20 * s * s * 3 + 11, torch.float16, s)
        ones = torch.ones(s, s, dtype=torch.float16, device=dev)
        y = ops.run_scan_multi_cube(x.to(dev), torch.triu(ones), torch.tril(ones, -1))
        assert_exact(to_np(y), oracle.scan(to_np(x)))


# --- flag-segmented scan (tests/test_seg_scan_single_core.py:40-63) ---------
@pytest.mark.parametrize("n", [1000, 4096, 20 * 32 * 32 * 3, 1_000_001])
@pytest.mark.parametrize("density", [0.0, 0.001, 0.01, 0.3, 1.0])
@pytest.mark.parametrize("dtype", [torch.float16, torch.int8], ids=str)
def test_seg_scan_flags(ops, dev, n, density, dtype):
    g = gen(n)
    x = torch.randint(-2, 3, (n,), generator=g, dtype=torch.int8).to(dtype)
    f = (torch.rand(n, generator=g) < density).to(torch.int8)
    for fn in (ops.run_seg_scan, ops.run_seg_scan_vec):
        y = fn(x.to(dev), f.to(dev), 32)
        assert_exact(to_np(y), oracle.seg_scan_flags(to_np(x), to_np(f)))
