"""Counted filters through the C ABI: tcs_count_filter leaves the kept count
of every 2048-element chunk in the workspace, and tcs_filter_counted
compacts each chunk on its own from them.  Checked against the oracle and
against tcs_filter (the look-back form) bit for bit, for every input type,
both comparisons, values / positions / both, ragged lengths, NaN, infinities
and ties at the pivot, a capacity below the kept count, and outputs that
start at any element-aligned address."""
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
F16, BF16, F32 = 0, 1, 2
CMP_LE, CMP_GE = 3, 4
CODE = {torch.float16: F16, torch.bfloat16: BF16, torch.float32: F32}


@pytest.fixture(scope="module")
def lib():
    lib = ctypes.CDLL(LIB)
    sig = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
           ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
    lib.tcs_filter_counted.argtypes = sig
    lib.tcs_filter.argtypes = sig
    lib.tcs_count_filter.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int64,
                                     ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
    lib.tcs_compress_workspace_bytes.restype = ctypes.c_size_t
    lib.tcs_compress_workspace_bytes.argtypes = [ctypes.c_int64]
    return lib


def mixed(n, dtype, seed, pivot):
    """Normals with exact ties at the pivot (as the type rounds it), +-inf,
    -0.0 and NaN."""
    x = torch.randn(n, generator=gen(seed)).to(dtype)
    if n >= 8:
        pick = torch.randint(0, n, (max(n // 50, 5),), generator=gen(seed + 1))
        x[pick] = torch.tensor(pivot, dtype=torch.float32).to(dtype)
        x[pick[:1]] = float("inf")
        x[pick[1:2]] = float("-inf")
        x[pick[2:3]] = float("nan")
        x[pick[3:4]] = -0.0
    return x


def _run(lib, dev, x, pivot, cmp, values=True, positions=False, counted=True, cap=None, out_offset=0):
    n = x.numel()
    stream = torch.cuda.current_stream().cuda_stream
    wsb = lib.tcs_compress_workspace_bytes(n)
    ws = torch.zeros(wsb, dtype=torch.uint8, device=dev)
    cnt = torch.zeros(1, dtype=torch.int64, device=dev)
    xd = x.to(dev)
    assert lib.tcs_count_filter(xd.data_ptr(), CODE[x.dtype], pivot, cmp, n, cnt.data_ptr(), ws.data_ptr(), wsb,
                                stream) == 0
    k = int(cnt.item())
    cap = k if cap is None else cap
    ov = torch.zeros((cap + out_offset + 8) * x.element_size(), dtype=torch.uint8, device=dev).view(x.dtype)
    op = torch.zeros(cap + out_offset + 8, dtype=torch.int32, device=dev)
    fn = lib.tcs_filter_counted if counted else lib.tcs_filter
    rc = fn(xd.data_ptr(), CODE[x.dtype], pivot, cmp, ov[out_offset:].data_ptr() if values else None,
            op[out_offset:].data_ptr() if positions else None, n, cap, ws.data_ptr(), wsb, stream)
    assert rc == 0
    torch.cuda.synchronize()
    got = min(k, cap)
    return k, ov, op, got


def _bits(t):
    return t.view(torch.int16).numpy() if t.element_size() == 2 else t.view(torch.int32).numpy()


@pytest.mark.parametrize("dtype", [torch.float16, torch.bfloat16, torch.float32], ids=str)
@pytest.mark.parametrize("n", [1, 2047, 2048, 2049, 65536 * 3 + 4097, 1_000_003])
@pytest.mark.parametrize("cmp", [CMP_GE, CMP_LE])
def test_counted_filter_matches_oracle_and_lookback(lib, dev, dtype, n, cmp):
    pivot = 0.3
    x = mixed(n, dtype, n + cmp, pivot)
    xf = x.float().numpy()  # float holds every bf16 value exactly
    for values, positions in ((True, False), (False, True), (True, True)):
        k, ov, op, got = _run(lib, dev, x, pivot, cmp, values, positions)
        k2, ov2, op2, _ = _run(lib, dev, x, pivot, cmp, values, positions, counted=False)
        assert k == k2
        if values:
            assert torch.equal(ov[:got].cpu().view(torch.uint8), ov2[:got].cpu().view(torch.uint8))
        if positions:
            assert torch.equal(op[:got].cpu(), op2[:got].cpu())
        if dtype != torch.bfloat16:
            if values:
                exp = oracle.filter_cmp(x.numpy(), pivot, cmp)
                assert np.array_equal(_bits(ov[:got].cpu()), exp.view(np.int16 if exp.itemsize == 2 else np.int32))
            if positions:
                assert np.array_equal(op[:got].cpu().numpy(), oracle.filter_cmp(x.numpy(), pivot, cmp, True))
        else:
            pv = float(torch.tensor(pivot, dtype=torch.float32).to(torch.bfloat16).float())
            m = (xf >= pv) if cmp == CMP_GE else (xf <= pv)
            if positions:
                assert np.array_equal(op[:got].cpu().numpy(), np.flatnonzero(m).astype(np.int32))
            if values:
                assert np.array_equal(_bits(ov[:got].cpu()), _bits(x[torch.from_numpy(m)]))


@pytest.mark.parametrize("offset", [1, 3, 7, 33, 63])
@pytest.mark.parametrize("dtype", [torch.float16, torch.float32], ids=str)
def test_counted_filter_outputs_at_any_alignment(lib, dev, dtype, offset):
    n = 500_001
    x = mixed(n, dtype, offset, -0.2)
    k, ov, op, got = _run(lib, dev, x, -0.2, CMP_GE, True, True, out_offset=offset)
    exp_v = oracle.filter_cmp(x.numpy(), -0.2, CMP_GE)
    exp_p = oracle.filter_cmp(x.numpy(), -0.2, CMP_GE, True)
    assert np.array_equal(_bits(ov[offset:offset + k].cpu()), exp_v.view(np.int16 if exp_v.itemsize == 2 else np.int32))
    assert np.array_equal(op[offset:offset + k].cpu().numpy(), exp_p)
    assert int(torch.count_nonzero(ov[:offset].view(torch.uint8) if dtype == torch.float16 else ov[:offset])) == 0
    assert int(torch.count_nonzero(op[:offset])) == 0
    assert int(torch.count_nonzero(op[offset + k:])) == 0


def test_counted_filter_capacity(lib, dev):
    n = 3_000_017
    x = mixed(n, torch.float32, 5, 0.0)
    k, ov, op, got = _run(lib, dev, x, 0.0, CMP_LE, True, True)
    cap = k // 3
    _, ov3, op3, got3 = _run(lib, dev, x, 0.0, CMP_LE, True, True, cap=cap)
    assert got3 == cap
    assert torch.equal(ov3[:cap].cpu(), ov[:cap].cpu()) and torch.equal(op3[:cap].cpu(), op[:cap].cpu())
    assert int(torch.count_nonzero(op3[cap:])) == 0


def test_counted_filter_rejects(lib, dev):
    x = torch.randn(4096, device=dev)
    ws = torch.zeros(lib.tcs_compress_workspace_bytes(4096), dtype=torch.uint8, device=dev)
    out = torch.empty(4096, device=dev)
    s = torch.cuda.current_stream().cuda_stream
    # comparisons other than >= and <=, no output, a short workspace
    assert lib.tcs_filter_counted(x.data_ptr(), F32, 0.0, 0, out.data_ptr(), None, 4096, 4096, ws.data_ptr(),
                                  ws.numel(), s) != 0
    assert lib.tcs_filter_counted(x.data_ptr(), F32, 0.0, CMP_GE, None, None, 4096, 4096, ws.data_ptr(),
                                  ws.numel(), s) != 0
    assert lib.tcs_filter_counted(x.data_ptr(), F32, 0.0, CMP_GE, out.data_ptr(), None, 4096, 4096, ws.data_ptr(),
                                  64, s) != 0
