"""Top-k (run_topk_fp16, run_topk_int16, run_topk) and the k-th largest
estimator (run_topk_pivot_fp16).

CPU: the top-k oracle (first k of the stable sort) against torch.topk's
values and a brute-force count (exactly k - need elements above the k-th
value), and the estimator oracle against a direct loop and the lower-bound
property it guarantees.  GPU: values and indices bit-exact against the
oracle on the reference tests' grid (fp16 U[0,1) with its heavy ties and
k = 1, 2, 5, sqrt(n); tests/test_topk.py:27-90), on int16 over its full
range, every dtype in both directions, k = 0 and k = n, and at 2^26 against
torch.topk's values; the estimator against its oracle and on the reference
test's lower-bound check (tests/test_topk_pivot.py:40-69).
"""
import math

import numpy as np
import pytest
import torch

import oracle
from _util import assert_exact, gen, to_np

VEC_LENS = [5 * 1024, 10 * 1024, 10 * 1024 * 1024 - 1, 20 * 1024 * 1024 - 10]  # tests/test_topk.py:27


def test_oracle_topk_matches_torch_values():
    for n, k in ((1000, 1), (1000, 17), (5000, 5000), (4096, 64)):
        x = torch.rand(n, generator=gen(n + k)).half()  # many ties
        for largest in (True, False):
            v, i = oracle.topk(x.numpy(), k, largest)
            tv, _ = torch.topk(x.float(), k, largest=largest)
            assert_exact(v.astype(np.float32), tv.numpy(), f"n={n} k={k}")
            assert_exact(x.numpy()[i], v)
            # lowest positions among the ties at the k-th value
            kth = v[-1]
            strictly = (x.numpy() > kth) if largest else (x.numpy() < kth)
            eq_pos = np.flatnonzero(x.numpy() == kth)
            took = np.sort(i[v == kth])
            assert np.array_equal(took, eq_pos[: k - int(strictly.sum())])


def test_oracle_topk_pivot_direct_loop_and_bound():
    x = (10 * torch.randn(100_000, generator=gen(3))).half().numpy()
    est = oracle.topk_pivot(x, 32)
    tile = 1024
    tiles = -(-x.size // tile)
    per = -(-tiles // 64)
    for b in (0, 5, 63):
        vals = []
        for it in range(min(8, per)):
            base = (b * per + it) * tile
            if base >= x.size:
                break
            t = list(x[base:base + tile]) + [-np.inf] * max(0, base + tile - x.size)
            g = [sorted(t[32 * s:32 * s + 32], reverse=True)[7] for s in range(32)]
            vals.append(sorted(g, reverse=True)[3])
        assert est[b] == (max(vals) if vals else -np.inf)  # a block past the end holds -inf
    # at least 32 elements of the block's sample are >= its estimate
    for b in range(len(est)):
        lo = b * per * tile
        seg = x[lo: lo + min(8, per) * tile]
        if seg.size:
            assert (seg >= est[b]).sum() >= 32


@pytest.mark.gpu
@pytest.mark.parametrize("n", VEC_LENS)
def test_topk_reference_grid(ops, dev, n):
    torch.manual_seed(42)
    x = torch.rand(n, generator=gen(n)).half()
    xd = x.to(dev)
    for k in (1, 2, 5, int(np.sqrt(n))):
        v, i = ops.run_topk_fp16(xd, k, float(xd.min()), float(xd.max()), 128)
        assert v.dtype == torch.float16 and i.dtype == torch.int32 and v.shape == (k,)
        ev, ei = oracle.topk(x.numpy(), k)
        assert_exact(to_np(v), ev, f"n={n} k={k}")
        assert_exact(to_np(i), ei, f"n={n} k={k} indices")
        # the reference test's own check: sorted values agree with torch.topk
        tv, _ = torch.topk(xd, k, largest=True)
        assert torch.equal(torch.sort(v).values, torch.sort(tv).values)


@pytest.mark.gpu
def test_topk_int16(ops, dev):
    for n in (5 * 1024, 1_000_003):
        x = torch.randint(-2**15, 2**15 - 1, (n,), generator=gen(n), dtype=torch.int16)
        xd = x.to(dev)
        for k in (1, 7, 1000):
            v, i = ops.run_topk_int16(xd, k, xd.min(), xd.max(), 128)
            ev, ei = oracle.topk(x.numpy(), k)
            assert_exact(to_np(v), ev, f"n={n} k={k}")
            assert_exact(to_np(i), ei, f"n={n} k={k} indices")


def _special(n, dtype, seed):
    if not dtype.is_floating_point:
        info = torch.iinfo(dtype)
        x = torch.randint(info.min, info.max, (n,), generator=gen(seed), dtype=torch.int64).to(dtype)
        x[: n // 4] = x[n // 2]
        return x
    x = torch.randn(n, generator=gen(seed)).to(dtype)
    if n >= 16:
        pick = torch.randint(0, n, (max(n // 20, 8),), generator=gen(seed + 1))
        vals = torch.tensor([0.0, -0.0, float("inf"), float("-inf"), float("nan"), 1.0, -1.0, 0.5])
        x[pick] = vals.repeat(len(pick) // 8 + 1)[: len(pick)].to(dtype)
    return x


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [torch.float16, torch.bfloat16, torch.float32, torch.int8, torch.int16,
                                   torch.int32], ids=str)
@pytest.mark.parametrize("largest", [True, False])
def test_topk_every_dtype(ops, dev, dtype, largest):
    for n in (1, 100, 8192, 8193, 50_001, 400_000):
        x = _special(n, dtype, n)
        # 4096 and 4097: the largest k sorted in one workgroup, the smallest by the radix sort
        for k in sorted({1, min(n, 3), min(n, 1000), min(n, 4096), min(n, 4097), n}):
            v, i = ops.run_topk(x.to(dev), k, largest)
            if dtype == torch.bfloat16:
                ev, ei = oracle.topk(x.float().numpy(), k, largest)
                assert_exact(to_np(i), ei, f"n={n} k={k}")
                assert torch.equal(v.cpu().view(torch.int16), x[i.cpu().long()].view(torch.int16))
                continue
            ev, ei = oracle.topk(to_np(x), k, largest)
            assert_exact(to_np(i), ei, f"n={n} k={k} indices")
            assert_exact(to_np(v).view(np.uint8), ev.view(np.uint8), f"n={n} k={k} values")


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [torch.float16, torch.float32, torch.int16], ids=str)
@pytest.mark.parametrize("largest", [True, False])
def test_topk_recorded_and_recounted_tiles(ops, dev, dtype, largest):
    """Keys of two or more bytes: the last digit round records, per tile, the
    keys above the chosen higher digits and the low bytes of up to 16 keys
    matching them, and the tile counts come from that record.  Normal keys
    with 1,100 copies of one value near the k-th largest planted in a few
    tiles: most tiles take their recorded counts, the planted ones (more
    than 16 matching keys) are counted again from their keys, and the ties at
    the k-th value are taken in input order."""
    n = (1 << 20) + 333
    if dtype.is_floating_point:
        x = torch.randn(n, generator=gen(77)).to(dtype)
        v0 = 3.6 if largest else -3.6
    else:
        x = (torch.randn(n, generator=gen(77)) * 2000).round().clamp(-32768, 32767).to(dtype)
        v0 = 7200 if largest else -7200
    x[200_000:201_000] = v0
    x[700_000:700_500:5] = v0
    xd = x.to(dev)
    for k in (1, 100, 600, 1500, 3000):
        v, i = ops.run_topk(xd, k, largest)
        ev, ei = oracle.topk(to_np(x), k, largest)
        assert_exact(to_np(i), ei, f"k={k} indices")
        assert_exact(to_np(v).view(np.uint8), ev.view(np.uint8), f"k={k} values")


@pytest.mark.gpu
def test_topk_edges(ops, dev):
    x = torch.rand(1000, generator=gen(1)).half().to(dev)
    v, i = ops.run_topk(x, 0)
    assert v.numel() == 0 and i.numel() == 0
    with pytest.raises(RuntimeError):
        ops.run_topk(x, 1001)
    with pytest.raises(RuntimeError):
        ops.run_topk_fp16(x.float(), 3, 0.0, 1.0, 128)
    with pytest.raises(RuntimeError):
        ops.run_topk(torch.zeros(8, dtype=torch.float64, device=dev), 2)
    # all equal: the first k positions
    z = torch.full((50_000,), 3, dtype=torch.int16, device=dev)
    v, i = ops.run_topk(z, 100)
    assert torch.equal(i, torch.arange(100, dtype=torch.int32, device=dev))


@pytest.mark.gpu
def test_topk_at_scale(ops, dev):
    n = (1 << 26) + 11
    x = torch.randn(n, generator=gen(6)).half().to(dev)
    for k in (1, 4096, 1 << 20):
        v, i = ops.run_topk(x, k)
        tv, _ = torch.topk(x, k)
        assert torch.equal(v, tv), k
        assert torch.equal(x[i.long()], v)
        assert int(torch.unique(i).numel()) == k


@pytest.mark.gpu
def test_topk_more_tiles_than_one_scan_round(ops, dev):
    # 2^28 + 17 int8 keys: 16385 tiles of 16384, so the tile-count scan runs
    # two rounds, with heavy ties at the k-th value (256 distinct keys).
    # Expected: every key beyond the k-th value T, then the first `need`
    # positions holding T, ordered by value (ties by position).
    n = (1 << 28) + 17
    g = torch.Generator(device=dev).manual_seed(11)
    xd = torch.randint(-128, 128, (n,), generator=g, device=dev, dtype=torch.int16).to(torch.int8)
    x = xd.cpu().numpy()
    cnt = np.bincount(x.astype(np.int16) + 128, minlength=256)
    for k, largest in ((1000, True), (3_000_000, False)):
        v, i = ops.run_topk(xd, k, largest)
        acc = 0
        for b in (range(255, -1, -1) if largest else range(256)):
            if acc + cnt[b] >= k:
                T = b - 128
                break
            acc += cnt[b]
        beyond = np.flatnonzero(x > T) if largest else np.flatnonzero(x < T)
        need = k - beyond.size
        sel = np.concatenate([beyond, np.flatnonzero(x == T)[:need]])
        key = -x[sel].astype(np.int16) if largest else x[sel].astype(np.int16)
        ei = sel[np.lexsort((sel, key))].astype(np.int32)
        assert_exact(to_np(i), ei, f"k={k} largest={largest}")
        assert_exact(to_np(v), x[ei], f"k={k} largest={largest} values")
    del xd


@pytest.mark.gpu
@pytest.mark.parametrize("num_samples", [4, 16, 32, 64])
def test_topk_pivot_matches_oracle(ops, dev, num_samples):
    for n in (1000, 100_000, 3_000_017):
        x = (10 * torch.randn(n, generator=gen(n))).half()
        est = ops.run_topk_pivot_fp16(x.to(dev), 8, num_samples)
        assert est.dtype == torch.float16
        assert_exact(to_np(est), oracle.topk_pivot(x.numpy(), num_samples), f"n={n}")


@pytest.mark.gpu
@pytest.mark.parametrize("n", [10 * 1024 * 1024, 20 * 1024 * 1024, 30 * 1024 * 1024, 40 * 1024 * 1024])
def test_topk_pivot_reference_bound(ops, dev, n):
    # tests/test_topk_pivot.py:40-69: the largest estimate is below the
    # 2048-th largest value.
    x = (10 * torch.randn(n, generator=gen(n))).half().to(dev)
    est = torch.sort(ops.run_topk_pivot_fp16(x, 8)).values[-1]
    kth = torch.sort(torch.topk(x, 2048, largest=True).values).values[0]
    assert kth > est, (float(kth), float(est))
