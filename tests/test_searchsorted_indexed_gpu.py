"""Bucket-indexed searchsorted through the C ABI (tcs_searchsorted_indexed):
lower bounds of int32 needles against numpy's searchsorted(side='left') and
against tcs_searchsorted, bit for bit, on haystacks whose keys spread over
the int32 range, sit in a narrow range with long runs of equal keys, are all
equal, form two far-apart clusters, or follow a heavy-tailed distribution
(buckets holding thousands of keys, which take the sampled search), with
needles below, inside and above the haystack and at INT32_MIN / INT32_MAX,
for lengths from 1 to 2^24 (bucket tables of 2^10 to 2^20 entries)."""
import ctypes
import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(REPO, "pytorch-tcuscan_amd", "libtcuscan_hip.so")
I32_MIN, I32_MAX = -2**31, 2**31 - 1


@pytest.fixture(scope="module")
def lib():
    lib = ctypes.CDLL(LIB)
    lib.tcs_searchsorted_indexed.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                                             ctypes.c_int64, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
    lib.tcs_searchsorted.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.c_int64, ctypes.c_void_p]
    lib.tcs_searchsorted_workspace_bytes.restype = ctypes.c_size_t
    lib.tcs_searchsorted_workspace_bytes.argtypes = [ctypes.c_int64]
    return lib


def _search(lib, dev, hay, needles, indexed=True):
    h = torch.from_numpy(hay).to(dev)
    v = torch.from_numpy(needles).to(dev)
    out = torch.full((len(needles),), -7, dtype=torch.int32, device=dev)
    s = torch.cuda.current_stream().cuda_stream
    if indexed:
        wsb = lib.tcs_searchsorted_workspace_bytes(len(hay))
        ws = torch.full((max(wsb, 4) // 4,), -1, dtype=torch.int32, device=dev)  # stale contents
        rc = lib.tcs_searchsorted_indexed(h.data_ptr() if len(hay) else None, len(hay), v.data_ptr(),
                                          out.data_ptr(), len(needles), ws.data_ptr(), wsb, s)
    else:
        rc = lib.tcs_searchsorted(h.data_ptr() if len(hay) else None, len(hay), v.data_ptr(), out.data_ptr(),
                                  len(needles), s)
    assert rc == 0
    torch.cuda.synchronize()
    return out.cpu().numpy()


def _needles(rng, hay, m):
    lo = int(hay[0]) if len(hay) else 0
    hi = int(hay[-1]) if len(hay) else 0
    parts = [rng.integers(I32_MIN, I32_MAX, m // 2, dtype=np.int64, endpoint=True),
             rng.integers(lo, hi, m - m // 2, dtype=np.int64, endpoint=True) if hi >= lo else np.zeros(0, np.int64),
             np.array([I32_MIN, I32_MAX, lo, hi, lo - 1, hi + 1, 0], dtype=np.int64)]
    if len(hay):
        parts.append(hay[rng.integers(0, len(hay), min(m, 4096))].astype(np.int64))  # exact hits
    v = np.concatenate(parts)
    return np.clip(v, I32_MIN, I32_MAX).astype(np.int32)


def _haystack(kind, n, rng):
    if kind == "uniform":
        h = rng.integers(I32_MIN, I32_MAX, n, dtype=np.int64, endpoint=True)
    elif kind == "narrow":
        h = rng.integers(0, 100, n, endpoint=True)
    elif kind == "constant":
        h = np.full(n, 7)
    elif kind == "clusters":
        h = np.concatenate([rng.integers(I32_MIN, I32_MIN + 1000, n // 2, endpoint=True),
                            rng.integers(I32_MAX - 1000, I32_MAX, n - n // 2, endpoint=True)])
    elif kind == "heavy_tail":
        h = np.minimum(rng.pareto(0.7, n) * 100, I32_MAX).astype(np.int64)
    else:
        raise ValueError(kind)
    return np.sort(h).astype(np.int32)


@pytest.mark.parametrize("kind", ["uniform", "narrow", "constant", "clusters", "heavy_tail"])
@pytest.mark.parametrize("n", [1, 2, 3, 1000, 65537, (1 << 20) + 3])
def test_indexed_matches_numpy_and_sampled(lib, dev, kind, n):
    rng = np.random.default_rng(n * 7 + len(kind))
    hay = _haystack(kind, n, rng)
    needles = _needles(rng, hay, 200_000)
    exp = np.searchsorted(hay, needles, side="left").astype(np.int32)
    got = _search(lib, dev, hay, needles)
    assert np.array_equal(got, exp), f"{kind} n={n}: first mismatch at {np.flatnonzero(got != exp)[:5]}"
    assert np.array_equal(got, _search(lib, dev, hay, needles, indexed=False))


def test_indexed_full_table(lib, dev):
    """2^24 haystack: a bucket table of 2^20 + 1 entries."""
    rng = np.random.default_rng(24)
    hay = _haystack("uniform", 1 << 24, rng)
    needles = _needles(rng, hay, 1 << 20)
    assert lib.tcs_searchsorted_workspace_bytes(1 << 24) == ((1 << 20) + 1) * 4
    got = _search(lib, dev, hay, needles)
    assert np.array_equal(got, np.searchsorted(hay, needles, side="left").astype(np.int32))


def test_indexed_empty_and_rejects(lib, dev):
    assert lib.tcs_searchsorted_workspace_bytes(0) == 0
    assert lib.tcs_searchsorted_workspace_bytes(1000) == ((1 << 10) + 1) * 4
    needles = np.array([5, -5, 0], dtype=np.int32)
    assert np.array_equal(_search(lib, dev, np.zeros(0, np.int32), needles), np.zeros(3, np.int32))
    hay = torch.arange(100000, dtype=torch.int32, device=dev)
    v = torch.zeros(10, dtype=torch.int32, device=dev)
    out = torch.empty(10, dtype=torch.int32, device=dev)
    ws = torch.empty(64, dtype=torch.int32, device=dev)
    s = torch.cuda.current_stream().cuda_stream
    # a workspace below tcs_searchsorted_workspace_bytes
    assert lib.tcs_searchsorted_indexed(hay.data_ptr(), 100000, v.data_ptr(), out.data_ptr(), 10, ws.data_ptr(),
                                        256, s) != 0


def test_op_leaves_the_stream_workspace_intact(ops, dev):
    """run_searchsorted on a large haystack builds its bucket table in a
    buffer of its own: look-back ops on the same stream afterwards (the
    persistent scan's ticket counters and launch epoch live in the shared
    workspace's header) still give exact results, and so does a second
    search."""
    import oracle
    from _util import gen

    rng = np.random.default_rng(5)
    hay = _haystack("uniform", 1 << 18, rng)
    needles = _needles(rng, hay, 1 << 17)
    exp = np.searchsorted(hay, needles, side="left").astype(np.int32)
    h, v = torch.from_numpy(hay).to(dev), torch.from_numpy(needles).to(dev)
    x = (torch.rand((1 << 20) + 5, generator=gen(3)) * 2 - 1).half()
    xd = x.to(dev)
    for _ in range(3):
        assert np.array_equal(ops.run_searchsorted(h, v).cpu().numpy(), exp)
        y = ops.run_scan_multi_core(xd, 128)
        assert np.array_equal(y.cpu().numpy(), oracle.scan(x.numpy()))
        m = (torch.rand(xd.numel(), generator=gen(4)) < 0.3).to(torch.int8).to(dev)
        z = ops.run_compress(xd, m, 128)
        assert torch.equal(z.cpu(), x[m.cpu() == 1])
