"""tcs_scan_noalloc, the scan entry point that allocates nothing, through the
C ABI: its results equal tcs_scan's (the same kernels over a caller's
workspace) bit for bit on two calls in a row, the second with a smaller
input, for every accepted dtype; inputs beyond its fixed scratch are
rejected before anything is launched."""
import ctypes
import os

import numpy as np
import pytest
import torch

import oracle
from _util import gen, to_np, uniform_pm1

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(REPO, "pytorch-tcuscan_amd", "libtcuscan_hip.so")
TCS_OK, TCS_ERR_SCRATCH_LIMIT = 0, 5
SCRATCH_BYTES = 1 << 20
CODES = {torch.float16: 0, torch.bfloat16: 1, torch.float32: 2, torch.int8: 3, torch.int32: 5}


@pytest.fixture(scope="module")
def lib():
    lib = ctypes.CDLL(LIB)
    lib.tcs_scan_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int64,
                                     ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_void_p]
    lib.tcs_scan.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int,
                             ctypes.c_int, ctypes.c_double, ctypes.c_void_p, ctypes.c_size_t,
                             ctypes.c_void_p]
    lib.tcs_scan_workspace_bytes.restype = ctypes.c_size_t
    lib.tcs_scan_workspace_bytes.argtypes = [ctypes.c_int64, ctypes.c_int, ctypes.c_int]
    lib.tcs_workspace_reset.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
    return lib


def _input(n, dtype, seed):
    if dtype == torch.int8:
        return torch.randint(-4, 5, (n,), generator=gen(seed), dtype=torch.int8)
    if dtype == torch.int32:
        return torch.randint(-1000, 1001, (n,), generator=gen(seed), dtype=torch.int32)
    return uniform_pm1(n, dtype, seed)


def _out(x, dev):
    dt = torch.int32 if x.dtype in (torch.int8, torch.int32) else torch.float32
    return torch.full((x.numel(),), -7, dtype=dt, device=dev)


def _with_workspace(lib, xd, S, start, stream):
    n = xd.numel()
    need = lib.tcs_scan_workspace_bytes(n, S, 0)
    ws = torch.zeros(need, dtype=torch.uint8, device=xd.device)
    y = _out(xd, xd.device)
    rc = lib.tcs_scan(xd.data_ptr(), CODES[xd.dtype], y.data_ptr(), n, S, 0, start, ws.data_ptr(), need, stream)
    assert rc == TCS_OK
    return y


def _no_alloc(lib, xd, S, start, stream):
    y = _out(xd, xd.device)
    rc = lib.tcs_scan_noalloc(xd.data_ptr(), CODES[xd.dtype], y.data_ptr(), xd.numel(), S, 0, start, stream)
    assert rc == TCS_OK
    return y


@pytest.mark.parametrize("dtype", list(CODES), ids=str)
def test_noalloc_matches_tcs_scan_on_two_calls(lib, dev, dtype):
    stream = torch.cuda.current_stream().cuda_stream
    for n, start, seed in (((1 << 22) + 4099, 0.0, 11), (100_003, 5.0, 12)):
        x = _input(n, dtype, seed)
        xd = x.to(dev)
        got = _no_alloc(lib, xd, 128, start, stream)
        want = _with_workspace(lib, xd, 128, start, stream)
        torch.cuda.synchronize()
        g, w = got.cpu().numpy(), want.cpu().numpy()
        assert np.array_equal(g, w), f"n={n} {dtype}: no-alloc differs from tcs_scan"
        if n < 200_000:
            assert np.array_equal(g, oracle.scan(to_np(x), start)), f"n={n} {dtype}: oracle mismatch"


def test_noalloc_many_calls_and_sizes(lib, dev):
    """Calls of shrinking and growing sizes in a row leave nothing behind
    that a later call could take for its own descriptors."""
    stream = torch.cuda.current_stream().cuda_stream
    for i, n in enumerate((1 << 24, 5000, 1 << 20, 1, 3_000_017, 64 * 1024 + 1)):
        x = _input(n, torch.float16, 100 + i)
        xd = x.to(dev)
        got = _no_alloc(lib, xd, 32, 0.0, stream)
        want = _with_workspace(lib, xd, 32, 0.0, stream)
        torch.cuda.synchronize()
        assert torch.equal(got, want), f"call {i} n={n}"


def test_noalloc_rejects_inputs_beyond_scratch(lib, dev):
    stream = torch.cuda.current_stream().cuda_stream
    n = 1
    while lib.tcs_scan_workspace_bytes(n, 32, 1) <= SCRATCH_BYTES:
        n *= 2
    fake = 0x100000  # never dereferenced: the call is rejected before any launch
    assert lib.tcs_scan_noalloc(fake, 0, fake, n, 32, 1, 0.0, stream) == TCS_ERR_SCRATCH_LIMIT
    assert lib.tcs_scan_noalloc(fake, 0, fake, 0, 32, 1, 0.0, stream) == TCS_OK
