"""Top-p (nucleus) filtering and sampling (tcuscan.top_p_filter,
tcuscan.sample_top_p) built on the radix sort and the scan.

CPU: the oracle against the reference's own recipe written with torch's CPU
operators (sort descending, cumsum, mask, renormalise;
scripts/profiling/profile_tcuscan_ops.py:372-392).  GPU: the kept set and
order exactly as the oracle's (the prefix sums are rounded once in both),
the renormalised probabilities within 1e-6, and draws that only ever land
in the nucleus.
"""
import numpy as np
import pytest
import torch

import oracle
from _util import assert_exact, gen, to_np


def _probs(n, seed):
    w = torch.rand(n, generator=gen(seed)) ** 4
    return (w / w.sum()).half()


def test_oracle_top_p_matches_torch_recipe():
    for n, p in ((100, 0.1), (5000, 0.5), (5000, 0.9), (33, 0.0)):
        x = _probs(n, n)
        kept, order = oracle.top_p_filter(x.numpy(), p)
        s, idx = torch.sort(x.float(), descending=True, stable=True)
        mass = torch.cumsum(s.double(), 0).float()
        ref = torch.where(mass - s > p, torch.zeros_like(s), s)
        ref = ref / ref.double().sum()
        assert_exact(order, idx.to(torch.int32).numpy())
        assert ((kept > 0) == (ref.numpy() > 0)).all()
        np.testing.assert_allclose(kept, ref.numpy(), rtol=1e-6, atol=1e-9)


@pytest.mark.gpu
@pytest.mark.parametrize("n", [1000, 50_000, 1 << 20])
@pytest.mark.parametrize("p", [0.1, 0.5, 0.95])
def test_top_p_filter_gpu(dev, n, p):
    import tcuscan
    x = _probs(n, n + int(100 * p))
    kept, order = tcuscan.top_p_filter(x.to(dev), p)
    ek, eo = oracle.top_p_filter(x.numpy(), p)
    assert_exact(to_np(order), eo)
    assert ((to_np(kept) > 0) == (ek > 0)).all()
    np.testing.assert_allclose(to_np(kept), ek, rtol=1e-6, atol=1e-9)


@pytest.mark.gpu
def test_sample_top_p_draws_from_the_nucleus(dev):
    import tcuscan
    x = _probs(10_000, 5).to(dev)
    kept, order = tcuscan.top_p_filter(x, 0.3)
    nucleus = set(order[kept > 0].long().tolist())
    g = torch.Generator(device=dev).manual_seed(0)
    draws = [int(tcuscan.sample_top_p(x, 0.3, generator=g)) for _ in range(200)]
    assert all(d in nucleus for d in draws)
    with pytest.raises(RuntimeError):
        tcuscan.top_p_filter(x.float(), 0.3)
