"""Top-p (nucleus) filtering and sampling (tcuscan.top_p_filter,
tcuscan.sample_top_p) built on the radix sort and the scan.

CPU: the oracle against the reference's own recipe written with torch's CPU
operators (sort descending, cumsum, mask, renormalise;
scripts/profiling/profile_tcuscan_ops.py:372-392).  GPU: the kept set and
order exactly as the oracle's (the prefix sums are rounded once in both),
the renormalised probabilities within 1e-6, the fused mask kernel bit-exact
against torch's evaluation of the reference's mask expression, and draws
that only ever land in the nucleus, with frequencies matching the kept
weights.
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


@pytest.mark.gpu
@pytest.mark.parametrize("n", [1, 7, 8, 1000, 65_537])
def test_top_p_keep_is_torchs_mask_expression(ops, dev, n):
    # kept = where(mass - x > p, 0, x) in float32, as torch evaluates the
    # reference's `probs_sum - probs_sort > p` (bit-exact), also on views
    # that start off the 16-byte grid.
    x = _probs(n + 1, n).to(dev)
    srt, _ = ops.run_radix_sort_ex(x, True)
    mass = ops.run_scan_multi_core(srt, 128)
    for p in (0.0, 0.3, 0.9, 2.0):
        for a, m in ((srt, mass), (srt[1:], mass[1:])):
            got = ops.run_top_p_keep(a, m, p)
            ref = torch.where(m - a.float() > p, torch.zeros_like(m), a.float())
            assert torch.equal(got, ref), (n, p)
    with pytest.raises(RuntimeError):
        ops.run_top_p_keep(srt.float(), mass, 0.1)
    with pytest.raises(RuntimeError):
        ops.run_top_p_keep(srt, mass[:-1], 0.1)


@pytest.mark.gpu
def test_sample_top_p_frequencies(dev):
    # Draw frequencies of the inverse-CDF sampler against the renormalised
    # kept weights (each within 5 standard deviations), and p = 0: the
    # nucleus is the largest probability alone.
    import tcuscan
    w = torch.tensor([8, 5, 4, 3, 2, 1, 1, 0, 6], dtype=torch.float32)
    x = (w / w.sum()).half().to(dev)
    kept, order = tcuscan.top_p_filter(x, 0.7)
    prob = {int(i): float(k) for i, k in zip(order.cpu(), kept.cpu()) if k > 0}
    g = torch.Generator(device=dev).manual_seed(3)
    draws = 3000
    counts = {}
    for _ in range(draws):
        d = int(tcuscan.sample_top_p(x, 0.7, generator=g))
        counts[d] = counts.get(d, 0) + 1
    assert set(counts) <= set(prob)
    for i, q in prob.items():
        sd = (draws * q * (1 - q)) ** 0.5
        assert abs(counts.get(i, 0) - draws * q) <= 5 * sd + 1, (i, counts.get(i, 0), draws * q)
    assert all(int(tcuscan.sample_top_p(x, 0.0, generator=g)) == 0 for _ in range(20))
