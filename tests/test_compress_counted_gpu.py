"""Counted mask compaction through the C ABI: tcs_count_mask leaves chunk
counts and block offsets in the workspace, and tcs_compress_counted places
every 2048-element chunk from them (no look-back).  Checked against the
oracle for every element size, every index mode, ragged lengths, outputs
that start at any element-aligned address (each run's vector stores cover
whole 128-byte lines; the head before the first line boundary, up to 127
int8 elements, and the tail are stored element by element), a capacity
below the kept count, and against
tcs_compress (the look-back form) bit for bit."""
import ctypes
import os

import numpy as np
import pytest
import torch

import oracle
from _util import gen

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(REPO, "pytorch-tcuscan_amd", "libtcuscan_hip.so")


@pytest.fixture(scope="module")
def lib():
    lib = ctypes.CDLL(LIB)
    sig = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
           ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p, ctypes.c_size_t,
           ctypes.c_void_p]
    lib.tcs_compress_counted.argtypes = sig
    lib.tcs_compress.argtypes = sig
    lib.tcs_count_mask.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                                   ctypes.c_size_t, ctypes.c_void_p]
    lib.tcs_compress_workspace_bytes.restype = ctypes.c_size_t
    lib.tcs_compress_workspace_bytes.argtypes = [ctypes.c_int64]
    return lib


def _run(lib, dev, x, mask, idx_mode, idx=None, out_offset=0, cap=None, counted=True):
    """(values, indices) kept, with the outputs placed `out_offset` elements
    into larger buffers."""
    n = x.numel()
    stream = torch.cuda.current_stream().cuda_stream
    wsb = lib.tcs_compress_workspace_bytes(n)
    ws = torch.zeros(wsb, dtype=torch.uint8, device=dev)
    cnt = torch.zeros(1, dtype=torch.int64, device=dev)
    xd, md = x.to(dev), mask.to(dev)
    assert lib.tcs_count_mask(md.data_ptr(), n, cnt.data_ptr(), ws.data_ptr(), wsb, stream) == 0
    k = int(cnt.item())
    cap = k if cap is None else cap
    ox = torch.zeros((cap + out_offset + 8) * x.element_size(), dtype=torch.uint8, device=dev).view(x.dtype)
    ie = idx.element_size() if idx is not None else 4
    idt = idx.dtype if idx is not None else torch.int32
    oi = torch.zeros((cap + out_offset + 8) * ie, dtype=torch.uint8, device=dev).view(idt)
    idd = idx.to(dev) if idx is not None else None
    fn = lib.tcs_compress_counted if counted else lib.tcs_compress
    rc = fn(xd.data_ptr(), x.element_size(), idd.data_ptr() if idd is not None else None, idx_mode, ie,
            md.data_ptr(), ox[out_offset:].data_ptr(), oi[out_offset:].data_ptr(), n, cap, ws.data_ptr(), wsb,
            stream)
    assert rc == 0
    torch.cuda.synchronize()
    got_k = min(k, cap)
    return ox[out_offset:out_offset + got_k].cpu(), oi[out_offset:out_offset + got_k].cpu(), k, ox, oi


def _bits(t):
    return t.view(torch.uint8).numpy() if t.dtype != torch.uint8 else t.numpy()


@pytest.mark.parametrize("dtype", [torch.int8, torch.float16, torch.float32, torch.float64], ids=str)
@pytest.mark.parametrize("n", [1, 2047, 2048, 2049, 65536 * 3 + 4097, 1_000_003])
def test_counted_values_and_positions(lib, dev, dtype, n):
    if dtype == torch.int8:
        x = torch.randint(-128, 128, (n,), generator=gen(n), dtype=torch.int8)
    else:
        x = (torch.randn(n, generator=gen(n)) * 100).to(dtype)
    m = (torch.rand(n, generator=gen(n + 1)) < 0.37).to(torch.int8)
    m[torch.rand(n, generator=gen(n + 2)) < 0.05] = 3  # values other than 0/1 are not kept
    for mode in (0, 2, 3):
        z, zi, k, _, _ = _run(lib, dev, x, m, mode)
        keep = (m == 1).numpy()
        assert np.array_equal(_bits(z), _bits(x[torch.from_numpy(keep)])), f"mode {mode} values"
        if mode:
            pos = np.flatnonzero(keep).astype(np.int32) + (1 if mode == 3 else 0)
            assert np.array_equal(zi.numpy(), pos), f"mode {mode} positions"


@pytest.mark.parametrize("idx_dtype", [torch.int32, torch.int64], ids=str)
def test_counted_gathered_indices(lib, dev, idx_dtype):
    n = 3_000_017
    x = torch.rand(n, generator=gen(5))
    idx = torch.randint(-2**31, 2**31 - 1, (n,), generator=gen(6)).to(idx_dtype)
    m = (torch.rand(n, generator=gen(7)) < 0.3).to(torch.int8)
    z, zi, k, _, _ = _run(lib, dev, x, m, 1, idx)
    ez, ei = oracle.compress_ind(x.numpy(), idx.numpy(), m.numpy())
    assert np.array_equal(z.numpy(), ez) and np.array_equal(zi.numpy(), ei)


@pytest.mark.parametrize("offset", [1, 2, 3, 5, 7, 33, 100])
@pytest.mark.parametrize("dtype", [torch.int8, torch.float16, torch.float32, torch.float64], ids=str)
def test_counted_outputs_at_any_alignment(lib, dev, offset, dtype):
    """Outputs that start off a line boundary: the head before the first
    128-byte boundary (more elements than a wave has lanes for int8) is
    written element by element; nothing outside the range is touched."""
    n = 700_001
    if dtype == torch.int8:
        x = torch.randint(-128, 128, (n,), generator=gen(offset), dtype=torch.int8)
    else:
        x = torch.randn(n, generator=gen(offset)).to(dtype)
    m = (torch.rand(n, generator=gen(offset + 10)) < 0.5).to(torch.int8)
    z, zi, k, ox, oi = _run(lib, dev, x, m, 2, out_offset=offset)
    keep = (m == 1)
    assert np.array_equal(_bits(z), _bits(x[keep]))
    assert np.array_equal(zi.numpy(), np.flatnonzero(keep.numpy()).astype(np.int32))
    assert int(torch.count_nonzero(ox[:offset].view(torch.uint8))) == 0
    assert int(torch.count_nonzero(ox[offset + k:].view(torch.uint8))) == 0
    assert int(torch.count_nonzero(oi[offset + k:])) == 0


def test_counted_capacity_and_lookback_identical(lib, dev):
    n = 5_000_011
    x = torch.rand(n, generator=gen(21))
    m = (torch.rand(n, generator=gen(22)) < 0.3).to(torch.int8)
    z1, zi1, k, _, _ = _run(lib, dev, x, m, 2, counted=True)
    z2, zi2, _, _, _ = _run(lib, dev, x, m, 2, counted=False)
    assert torch.equal(z1, z2) and torch.equal(zi1, zi2)
    cap = k // 3
    z3, zi3, _, ox, _ = _run(lib, dev, x, m, 2, cap=cap)
    assert torch.equal(z3, z1[:cap]) and torch.equal(zi3, zi1[:cap])
    assert int(torch.count_nonzero(ox[cap:])) == 0
