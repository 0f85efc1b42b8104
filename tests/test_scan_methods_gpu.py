"""GPU parity of the two local scan methods (DPP wavefront scans and the
matrix-unit scan of 32 x 32 units) against the oracle and each other.

The matrix-unit method is accepted for fp16, bf16 and int8 plain scans.
Its results must be bit-identical to the oracle's (float64 prefix rounded
once to float32; int32 with wrap) on every tile geometry it instantiates, on
ragged lengths, on units that take its sequential path (fp16 |x| >= 256;
bf16 rows spanning more than its three bands, or reaching 2^122; inf, NaN)
and with a starting sum.
"""
import numpy as np
import pytest
import torch

import oracle
from _util import assert_exact, gen, to_np, uniform_pm1

pytestmark = pytest.mark.gpu

# tiles_per_wg at S = 32 (1024-element matrix tiles) covering every
# (data waves, groups per wave) geometry of both methods: tiles of 1K to 64K
# (64K is int8 only; other dtypes cap at 32K).
TPW_S32 = [1, 2, 4, 8, 16, 32, 64]


def _input(n, dtype, seed):
    if dtype == torch.int8:
        return torch.randint(-128, 128, (n,), generator=gen(seed), dtype=torch.int8)
    return uniform_pm1(n, dtype, seed)


@pytest.mark.parametrize("method", ["dpp", "mfma"])
@pytest.mark.parametrize("tpw", TPW_S32)
@pytest.mark.parametrize("dtype", [torch.float16, torch.bfloat16, torch.int8], ids=str)
def test_methods_every_geometry(ops, dev, method, tpw, dtype):
    for n in (1, 1000, 1024, 1025, 32768 * 3 + 517, 1_000_003):
        x = _input(n, dtype, n + tpw)
        y = ops.run_scan_ex(x.to(dev), 32, tpw, method)
        assert_exact(to_np(y), oracle.scan(to_np(x)), f"{method} tpw={tpw} n={n} {dtype}")


@pytest.mark.parametrize("dtype", [torch.float16, torch.bfloat16, torch.int8], ids=str)
@pytest.mark.parametrize("n", [1, 31, 32, 33, 511, 512, 513, 1023, 1024, 1025, 2047, 2049,
                               16383, 16385, 65536 * 3 + 5])
def test_mfma_ragged_lengths(ops, dev, dtype, n):
    x = _input(n, dtype, 3 * n)
    y = ops.run_scan_ex(x.to(dev), 128, 0, "mfma")
    assert_exact(to_np(y), oracle.scan(to_np(x)), f"n={n} {dtype}")


def test_mfma_large_values_take_the_sequential_path(ops, dev):
    # Units holding |x| >= 256 are scanned from LDS; others on the matrix
    # cores.  Mix both, including values at the fp16 extremes.
    n = 1 << 20
    x = uniform_pm1(n, torch.float16, 11).float()
    idx = torch.randint(0, n, (300,), generator=gen(12))
    x[idx] = torch.tensor([256.0, -256.0, 65504.0, -65504.0, 300.5, 1000.0], dtype=torch.float32).repeat(50)
    x[5000:6000] = 255.875  # largest value below 256: stays on the matrix cores
    x = x.half()
    for tpw in (1, 16):
        y = ops.run_scan_ex(x.to(dev), 32, tpw, "mfma")
        assert_exact(to_np(y), oracle.scan(to_np(x)), f"tpw={tpw}")


def test_mfma_subnormals_exact(ops, dev):
    # fp16 subnormals (multiples of 2^-24) mixed with values near 1: the
    # split into 2^-6-grid and remainder parts must keep every bit.
    n = 1 << 18
    bits = torch.randint(0, 0x0400, (n,), generator=gen(21), dtype=torch.int32)
    sign = torch.randint(0, 2, (n,), generator=gen(22), dtype=torch.int32) << 15
    sub = (bits | sign).to(torch.int16).view(torch.float16)
    big = uniform_pm1(n, torch.float16, 23)
    pick = torch.rand(n, generator=gen(24)) < 0.5
    x = torch.where(pick, sub, big)
    y = ops.run_scan_ex(x.to(dev), 64, 0, "mfma")
    assert_exact(to_np(y), oracle.scan(to_np(x)))


def test_mfma_inf_nan(ops, dev):
    n = 40000
    x = uniform_pm1(n, torch.float16, 31)
    x[1500] = float("inf")      # inf inside a unit: outputs before it stay finite
    x[20000] = float("-inf")    # inf - inf = nan from here on
    y = to_np(ops.run_scan_ex(x.to(dev), 32, 4, "mfma"))
    exp = oracle.scan(to_np(x))
    assert_exact(y, exp)
    assert np.isfinite(y[:1500]).all() and np.isinf(y[1500:20000]).all() and np.isnan(y[20000:]).all()
    x2 = uniform_pm1(n, torch.float16, 32)
    x2[7777] = float("nan")
    y2 = to_np(ops.run_scan_ex(x2.to(dev), 32, 4, "mfma"))
    assert_exact(y2, oracle.scan(to_np(x2)))
    assert np.isfinite(y2[:7777]).all() and np.isnan(y2[7777:]).all()


@pytest.mark.parametrize("dtype", [torch.float16, torch.bfloat16, torch.int8], ids=str)
def test_mfma_starting_sum(ops, dev, dtype):
    n = 300_001
    x = _input(n, dtype, 41)
    for start in (0.0, 2.5, -1234.0, 1e6):
        y = ops.run_scan_ex(x.to(dev), 128, 0, "mfma", start)
        assert_exact(to_np(y), oracle.scan(to_np(x), start), f"start={start}")


@pytest.mark.parametrize("dtype", [torch.float16, torch.bfloat16, torch.int8], ids=str)
def test_methods_agree_at_scale(ops, dev, dtype):
    n = (1 << 26) + 12345
    x = _input(n, dtype, 51).to(dev)
    ya = ops.run_scan_ex(x, 128, 0, "dpp")
    yb = ops.run_scan_ex(x, 128, 0, "mfma")
    assert torch.equal(ya, yb)
    # int8 wraps in int32 like the sequential scan: spot-check the tail.
    if dtype == torch.int8:
        ref = torch.cumsum(x.long(), 0).to(torch.int64)
        tail = ((ref[-1000:] + 2**31) % 2**32 - 2**31).to(torch.int32)
        assert torch.equal(yb[-1000:], tail)


def test_auto_method_and_rejections(ops, dev):
    x = uniform_pm1(100_000, torch.float16, 61).to(dev)
    assert torch.equal(ops.run_scan_ex(x, 128), ops.run_scan_multi_core(x, 128))
    with pytest.raises(RuntimeError):
        ops.run_scan_ex(x.float(), 128, 0, "mfma")
    with pytest.raises(RuntimeError):
        ops.run_scan_ex(x.int(), 128, 0, "mfma")
    with pytest.raises(RuntimeError):
        ops.run_scan_ex(x, 128, 0, "tensor-cores")


# --- bf16 on the matrix units: three exponent bands per row ---------------
def _bf16_from_bits(bits: torch.Tensor) -> torch.Tensor:
    return bits.to(torch.int16).view(torch.bfloat16)


def test_mfma_bf16_exponent_spans(ops, dev):
    """Rows whose values span up to 28 binades below their largest stay on
    the matrix cores; wider spans (bits below the lowest band) and tiny
    values take the sequential path.  Every case is exact."""
    n = 1 << 20
    base = uniform_pm1(n, torch.float32, 71)
    scale = torch.pow(2.0, -torch.randint(0, 40, (n,), generator=gen(72)).float())
    x = (base * scale).bfloat16()
    for tpw in (1, 8):
        assert_exact(to_np(ops.run_scan_ex(x.to(dev), 32, tpw, "mfma")), oracle.scan(to_np(x)), f"tpw={tpw}")
    # Every unit within 20 binades: all on the matrix cores.
    narrow = (base * torch.pow(2.0, -torch.randint(0, 20, (n,), generator=gen(73)).float())).bfloat16()
    assert_exact(to_np(ops.run_scan_ex(narrow.to(dev), 32, 8, "mfma")), oracle.scan(to_np(narrow)))


def test_mfma_bf16_extremes(ops, dev):
    """Subnormals, the smallest normals, values near 2^122 and 2^127 (rows
    that could overflow fp32 row sums), inf and NaN."""
    n = 1 << 18
    # bf16 subnormals mixed with tiny normals.
    bits = torch.randint(0, 0x0100, (n,), generator=gen(81), dtype=torch.int32)
    sign = torch.randint(0, 2, (n,), generator=gen(82), dtype=torch.int32) << 15
    tiny = _bf16_from_bits(bits | sign)
    assert_exact(to_np(ops.run_scan_ex(tiny.to(dev), 64, 0, "mfma")), oracle.scan(to_np(tiny)), "subnormals")
    # Large magnitudes: alternate signs so that prefix sums stay finite.
    e = torch.randint(110, 127, (n,), generator=gen(83)).float()
    big = (torch.pow(2.0, e) * torch.where(torch.arange(n) % 2 == 0, 1.0, -1.0)).bfloat16()
    assert_exact(to_np(ops.run_scan_ex(big.to(dev), 64, 0, "mfma")), oracle.scan(to_np(big)), "large")
    x = uniform_pm1(40000, torch.bfloat16, 84)
    x[1500] = float("inf")
    x[20000] = float("nan")
    y = to_np(ops.run_scan_ex(x.to(dev), 32, 4, "mfma"))
    assert_exact(y, oracle.scan(to_np(x)), "inf/nan")
    assert np.isfinite(y[:1500]).all() and np.isinf(y[1500:20000]).all() and np.isnan(y[20000:]).all()
