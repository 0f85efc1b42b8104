"""Radix sort: run_radix_sort (ascending, the reference's op) and
run_radix_sort_ex (either direction); values and input positions.

CPU: the oracle against torch.sort(stable=True) on the CPU, values and
indices, with NaN, +-0, +-inf and many ties, and against a brute-force
comparator sort.  GPU: the HIP sort against the oracle bit for bit (values
and indices) on the reference tests' size grids (fp16 U[0,1), int16 over
its full range, tests/test_radix_sort.py:27-95), on every supported dtype,
both directions, ragged tails, special values; at 2^26 and 2^28 elements
against torch.sort(stable=True) on the GPU.
"""
import functools
import math

import numpy as np
import pytest
import torch

import oracle
from _util import assert_exact, gen, to_np

# tests/test_radix_sort.py:27-42
SORT_SIZES = [1024, 2048, 4096, 8192, 16384, 32768, 65536, 131072, 262144, 524288, 1179648, 4325376,
              16908288, 25165824]


def int16_sizes(s):
    # tests/test_radix_sort.py:45-52
    max_iters = math.ceil(1e7 / (20 * s * s)) - 5
    return [i * 20 * s * s for i in range(1, max_iters, 16 * 128 // s)]


def special(n, dtype, seed):
    x = torch.randn(n, generator=gen(seed)).to(dtype)
    if n >= 16:
        pick = torch.randint(0, n, (max(n // 20, 8),), generator=gen(seed + 1))
        vals = torch.tensor([0.0, -0.0, float("inf"), float("-inf"), float("nan"), 1.0, -1.0, 0.5])
        x[pick] = vals.repeat(len(pick) // 8 + 1)[: len(pick)].to(dtype)
    return x


def _torch_sort(x, descending):
    v, i = torch.sort(x, descending=descending, stable=True)
    return v, i.to(torch.int32)


# ---------------------------------------------------------------------------
# Oracle pinning (CPU)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("dtype", [torch.float16, torch.float32, torch.int16, torch.int32, torch.int8], ids=str)
@pytest.mark.parametrize("descending", [False, True])
def test_oracle_sort_matches_torch_stable(dtype, descending):
    for n in (1, 2, 100, 5000):
        if dtype.is_floating_point:
            x = special(n, dtype, n)
        else:
            info = torch.iinfo(dtype)
            x = torch.randint(info.min, info.max, (n,), generator=gen(n)).to(dtype)
            x[: n // 3] = x[n // 2]  # many ties
        v, i = oracle.sort(to_np(x), descending)
        ev, ei = _torch_sort(x, descending)
        assert_exact(i, ei.numpy(), f"indices n={n}")
        assert_exact(v, to_np(ev), f"values n={n}")


def test_oracle_sort_brute_force():
    def cmp(a, b):  # NaN last, -0 == +0, ties by position
        (va, ia), (vb, ib) = a, b
        ka = (math.isnan(va), 0.0 if math.isnan(va) else va)
        kb = (math.isnan(vb), 0.0 if math.isnan(vb) else vb)
        return -1 if (ka, ia) < (kb, ib) else 1
    x = np.array([0.5, -0.0, float("nan"), 0.0, -1.0, float("inf"), 0.5, float("-inf"), float("nan"), -0.0],
                 dtype=np.float32)
    exp = sorted([(float(v), k) for k, v in enumerate(x)], key=functools.cmp_to_key(cmp))
    v, i = oracle.sort(x)
    assert i.tolist() == [k for _, k in exp]
    # descending: NaN first, then values high to low, ties by position
    v, i = oracle.sort(x, descending=True)
    assert i.tolist() == [2, 8, 5, 0, 6, 1, 3, 9, 4, 7]


# ---------------------------------------------------------------------------
# HIP sort (GPU)
# ---------------------------------------------------------------------------

def _check(ops, dev, x, descending, what):
    if descending:
        v, i = ops.run_radix_sort_ex(x.to(dev), True)
    else:
        v, i = ops.run_radix_sort(x.to(dev), 128)
    assert v.dtype == x.dtype and i.dtype == torch.int32 and v.shape == x.shape
    ev, ei = oracle.sort(to_np(x), descending)
    assert_exact(to_np(i), ei, f"{what} indices")
    assert_exact(to_np(v).view(np.uint8), ev.view(np.uint8), f"{what} value bits")


@pytest.mark.gpu
@pytest.mark.parametrize("S", [32, 64, 128])
def test_sort_reference_grid_fp16(ops, dev, S):
    for n in SORT_SIZES[:12]:
        if n < S * S:
            continue
        x = torch.rand(n, generator=gen(n + S)).half()
        v, i = ops.run_radix_sort(x.to(dev), S)
        ev, ei = oracle.sort(to_np(x))
        assert_exact(to_np(v), ev, f"n={n}")
        assert_exact(to_np(i), ei, f"n={n} indices")


@pytest.mark.gpu
@pytest.mark.parametrize("n", SORT_SIZES[12:])
def test_sort_reference_grid_fp16_large(ops, dev, n):
    x = torch.rand(n, generator=gen(n)).half().to(dev)
    v, i = ops.run_radix_sort(x, 128)
    ev, ei = _torch_sort(x, False)
    assert torch.equal(v, ev) and torch.equal(i, ei)


@pytest.mark.gpu
@pytest.mark.parametrize("S", [64, 128])
def test_sort_reference_grid_int16(ops, dev, S):
    for n in int16_sizes(S) + [1000, 8191, 8193, 123457]:
        x = torch.randint(-2**15, 2**15 - 1, (n,), generator=gen(n + S), dtype=torch.int16)
        _check(ops, dev, x, False, f"n={n}")


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [torch.float16, torch.bfloat16, torch.float32, torch.int8, torch.int16,
                                   torch.int32], ids=str)
@pytest.mark.parametrize("descending", [False, True])
def test_sort_every_dtype(ops, dev, dtype, descending):
    for n in (1, 2, 63, 64, 65, 8191, 8192, 8193, 3 * 8192 + 17, 300_001):
        if dtype.is_floating_point:
            x = special(n, dtype, n)
        else:
            info = torch.iinfo(dtype)
            x = torch.randint(info.min, info.max, (n,), generator=gen(n), dtype=torch.int64).to(dtype)
        if dtype == torch.bfloat16:
            # numpy has no bfloat16: compare indices with the oracle on float32
            # (exact for bf16 values) and values through the indices.
            v, i = ops.run_radix_sort_ex(x.to(dev), descending)
            ev, ei = oracle.sort(x.float().numpy(), descending)
            assert_exact(to_np(i), ei, f"n={n}")
            assert torch.equal(v.cpu().view(torch.int16), x[i.cpu().long()].view(torch.int16))
            continue
        _check(ops, dev, x, descending, f"n={n}")


@pytest.mark.gpu
def test_sort_uint32_and_ties(ops, dev):
    n = 200_000
    x = torch.randint(0, 2**32, (n,), generator=gen(1), dtype=torch.int64)
    v, i = ops.run_radix_sort(x.to(torch.uint32).to(dev), 64)
    xi = x.numpy()
    order = np.lexsort((np.arange(n), xi))
    assert_exact(to_np(i), order.astype(np.int32))
    assert_exact(v.cpu().to(torch.int64).numpy(), xi[order])
    # all equal: the identity permutation
    z = torch.full((100_000,), 7, dtype=torch.int16, device=dev)
    v, i = ops.run_radix_sort(z, 64)
    assert torch.equal(i, torch.arange(100_000, dtype=torch.int32, device=dev))


@pytest.mark.gpu
def test_sort_edges(ops, dev):
    e = torch.empty(0, dtype=torch.float16, device=dev)
    v, i = ops.run_radix_sort(e, 64)
    assert v.numel() == 0 and i.dtype == torch.int32
    x = torch.rand(10_001, generator=gen(2)).half().to(dev)
    xv = x[1:]  # unaligned view
    v, i = ops.run_radix_sort(xv, 64)
    ev, ei = _torch_sort(xv, False)
    assert torch.equal(v, ev) and torch.equal(i, ei)
    with pytest.raises(RuntimeError):
        ops.run_radix_sort(torch.zeros(4, dtype=torch.float64, device=dev), 64)


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [torch.float16, torch.float32], ids=str)
def test_sort_at_scale(ops, dev, dtype):
    n = (1 << 26) + 3
    x = torch.randn(n, generator=gen(4)).to(dtype).to(dev)
    for descending in (False, True):
        v, i = ops.run_radix_sort_ex(x, descending)
        ev, ei = _torch_sort(x, descending)
        assert torch.equal(v, ev) and torch.equal(i, ei), descending


@pytest.mark.gpu
def test_workspace_shared_across_sizes_and_ops(ops, dev):
    # Regression: every op on a stream shares one workspace, so scratch a
    # call leaves behind (digit counts, row indices) later sits where a
    # larger call keeps look-back descriptors.  Alternate sizes and ops and
    # check every result.
    for n in (1 << 22, 131072, 1 << 20, 50_000, 3 << 20):
        x = torch.rand(n, generator=gen(n)).half()
        v, i = ops.run_radix_sort(x.to(dev), 64)
        ev, ei = oracle.sort(to_np(x))
        assert_exact(to_np(i), ei, f"sort n={n}")
        rows = n // 64
        indptr = torch.arange(0, rows + 1, dtype=torch.int32) * 64
        xs = torch.randint(-4, 5, (rows * 64,), generator=gen(n + 1), dtype=torch.int8)
        y = ops.run_seg_sum_multi_core(xs.to(dev), indptr.to(dev), 128)
        assert_exact(to_np(y), oracle.seg_sum_indptr(to_np(xs), indptr.numpy()), f"segsum n={n}")
        xi = torch.randint(-4, 5, (2 * n,), generator=gen(n + 2), dtype=torch.int8)
        assert_exact(to_np(ops.run_scan_multi_core(xi.to(dev), 32)), oracle.scan(to_np(xi)), f"scan n={n}")
