"""Split (stable partition by mask == 1): run_split and run_split_ind.

CPU: the oracle against torch's masked_select of mask == 1 followed by that
of mask == 0 (the reference tests' expectation, tests/test_split.py:52-66)
and against a brute-force loop.  GPU: the HIP operator against the oracle
bit for bit on the reference tests' length grid (both dtypes, every S,
1-based indices as in tests/test_split_ind.py), on masks with values other
than 0 and 1, on every element size, all / none kept, empty, unaligned
views, and 2^26 elements against torch.
"""
import numpy as np
import pytest
import torch

import oracle
from _util import assert_exact, gen, to_np

# tests/test_split.py:27-42
VEC_LENS = [256 - 1, 1024 - 1, 2048 - 1, 4096 - 1, 8192 - 1, 1024 * 1024 - 1, 256 + 1, 1024 + 1, 2048 + 1,
            4096 + 1, 8192 + 1, 1024 * 1024 + 1, 1024 * 1024 + 13, 1024 * 1024 + 27]


def _mask(n, seed):
    return (torch.randn(n, generator=gen(seed)) > 0).to(torch.int8)


def test_oracle_split_matches_torch_and_loop():
    for n in (0, 1, 7, 1000, 4097):
        x = torch.randint(0, 127, (n,), generator=gen(n)).to(torch.int16)
        m = _mask(n, n + 1)
        idx = torch.arange(1, n + 1, dtype=torch.int32)
        z, zi = oracle.split(x.numpy(), m.numpy(), idx.numpy())
        exp = torch.cat([torch.masked_select(x, m == 1), torch.masked_select(x, m == 0)])
        assert_exact(z, exp.numpy())
        expi = torch.cat([torch.nonzero(m == 1, as_tuple=True)[0], torch.nonzero(m == 0, as_tuple=True)[0]]) + 1
        assert_exact(zi, expi.to(torch.int32).numpy())
        # brute force
        left = [v for v, k in zip(x.tolist(), m.tolist()) if k == 1]
        right = [v for v, k in zip(x.tolist(), m.tolist()) if k != 1]
        assert z.tolist() == left + right
    # values other than 0/1 go to the second part
    z = oracle.split(np.arange(5), np.array([1, 2, 0, 1, -1], dtype=np.int8))
    assert z.tolist() == [0, 3, 1, 2, 4]


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [torch.int16, torch.float16], ids=str)
@pytest.mark.parametrize("S", [32, 64, 128])
def test_split_reference_grid(ops, dev, dtype, S):
    for n in VEC_LENS:
        x = torch.randint(0, 2**7 - 1, (n,), generator=gen(n + S)).to(dtype)
        m = _mask(n, 7 * n + S)
        idx = torch.cumsum(torch.ones(n), dim=-1).to(torch.int32)   # 1-based, tests/test_split_ind.py
        z = ops.run_split(x.to(dev), m.to(dev), S)
        assert z.dtype == dtype and z.shape == (n,)
        assert_exact(to_np(z), oracle.split(to_np(x), m.numpy()), f"n={n}")
        z2, zi = ops.run_split_ind(x.to(dev), m.to(dev), idx.to(dev), S)
        ez, ei = oracle.split(to_np(x), m.numpy(), idx.numpy())
        assert_exact(to_np(z2), ez, f"ind n={n}")
        assert_exact(to_np(zi), ei, f"ind idx n={n}")


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [torch.int8, torch.int16, torch.float32, torch.int32, torch.float64, torch.int64],
                         ids=str)
def test_split_every_element_size(ops, dev, dtype):
    for n in (1, 63, 8191, 8193, 100_003):
        x = torch.randint(-1000, 1000, (n,), generator=gen(n)).to(dtype)
        m = torch.randint(-1, 3, (n,), generator=gen(n + 1)).to(torch.int8)  # 1 kept; -1, 0, 2 not
        idx = torch.randperm(n, generator=gen(n + 2)).to(torch.int32)
        assert_exact(to_np(ops.run_split(x.to(dev), m.to(dev), 64)), oracle.split(to_np(x), m.numpy()))
        z, zi = ops.run_split_ind(x.to(dev), m.to(dev), idx.to(dev), 64)
        ez, ei = oracle.split(to_np(x), m.numpy(), idx.numpy())
        assert_exact(to_np(z), ez)
        assert_exact(to_np(zi), ei)


@pytest.mark.gpu
def test_split_edges(ops, dev):
    n = 50_000
    x = torch.randint(0, 100, (n,), generator=gen(3)).to(torch.float16).to(dev)
    ones = torch.ones(n, dtype=torch.int8, device=dev)
    assert torch.equal(ops.run_split(x, ones, 64), x)
    assert torch.equal(ops.run_split(x, torch.zeros_like(ones), 64), x)
    e = torch.empty(0, dtype=torch.float16, device=dev)
    assert ops.run_split(e, torch.empty(0, dtype=torch.int8, device=dev), 64).numel() == 0
    # unaligned views of x and mask
    m = _mask(n, 4).to(dev)
    xv, mv = x[3:], m[1:n - 2]
    exp = torch.cat([torch.masked_select(xv, mv == 1), torch.masked_select(xv, mv != 1)])
    assert torch.equal(ops.run_split(xv, mv, 64), exp)
    with pytest.raises(RuntimeError):
        ops.run_split(x, m.to(torch.int32), 64)
    with pytest.raises(RuntimeError):
        ops.run_split(x, m[:-1], 64)
    with pytest.raises(RuntimeError):
        ops.run_split_ind(x, m, torch.arange(n, device=dev), 64)  # int64 indices


@pytest.mark.gpu
def test_split_at_scale(ops, dev):
    n = (1 << 26) + 5
    x = torch.randn(n, generator=gen(5)).half().to(dev)
    m = (torch.rand(n, generator=gen(6)) < 0.3).to(torch.int8).to(dev)
    idx = torch.arange(n, dtype=torch.int32, device=dev)
    z, zi = ops.run_split_ind(x, m, idx, 128)
    k = int((m == 1).sum())
    assert torch.equal(z[:k], torch.masked_select(x, m == 1))
    assert torch.equal(z[k:], torch.masked_select(x, m != 1))
    assert torch.equal(zi[:k], torch.nonzero(m == 1, as_tuple=True)[0].int())
    assert torch.equal(zi[k:], torch.nonzero(m != 1, as_tuple=True)[0].int())
