"""run_scan_cpu (the reference's CPU path, src/torch/torch_scan_cpu.h:23-107)
against torch.cumsum with the reference test's tolerance
(tests/test_scan_cpu.py:29-48), against the oracle's restatement of the
three-phase blocked algorithm, and against the C port of it."""
import ctypes
import os

import numpy as np
import pytest
import torch

import oracle
from _util import assert_exact

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("vec_len", [1024, 10 * 1024, 30 * 1024, 1_000_000])
@pytest.mark.parametrize("dtype", [torch.int8, torch.float16, torch.float32], ids=str)
def test_scan_cpu_matches_cumsum(ops, vec_len, dtype):
    g = torch.Generator().manual_seed(vec_len)
    if dtype == torch.int8:
        x = torch.randint(0, 10, (vec_len,), generator=g, dtype=dtype)
        out = torch.int32
    else:
        x = (0.1 * torch.randn(vec_len, generator=g)).to(dtype)
        out = torch.float32
    y = ops.run_scan_cpu(x)
    assert y.dtype == out
    assert torch.allclose(y, torch.cumsum(x, 0, dtype=out), atol=1e-3, rtol=1e-7)
    if dtype == torch.int8:
        assert torch.equal(y, torch.cumsum(x, 0, dtype=out))


@pytest.mark.parametrize("threads", [1, 3, 8])
def test_scan_cpu_blocked_rounding(ops, threads):
    """The product's blocked scan and the oracle restatement round alike for
    the same block count (set through torch's thread count)."""
    x = (torch.randn(100_003, generator=torch.Generator().manual_seed(1)) * 0.1).float()
    old = torch.get_num_threads()
    try:
        torch.set_num_threads(threads)
        y = ops.run_scan_cpu(x).numpy()
    finally:
        torch.set_num_threads(old)
    assert_exact(y, oracle.blocked_scan_cpu(x.numpy(), threads))


@pytest.fixture(scope="module")
def port():
    path = os.path.join(REPO, "oracle", "libscan_cpu_port.so")
    if not os.path.exists(path):
        import subprocess
        subprocess.run(["make", "-C", os.path.join(REPO, "oracle")], check=True)
    return ctypes.CDLL(path)


@pytest.mark.parametrize("threads", [1, 4, 7])
def test_c_port_matches_oracle_restatement(port, threads):
    rng = np.random.default_rng(threads)
    x = (rng.standard_normal(50_001) * 0.1).astype(np.float16)
    y = np.empty(x.size, np.float32)
    port.port_scan_cpu_f16(x.ctypes.data_as(ctypes.c_void_p), y.ctypes.data_as(ctypes.c_void_p),
                           ctypes.c_int64(x.size), threads)
    assert_exact(y, oracle.blocked_scan_cpu(x, threads))
    xi = rng.integers(-128, 128, 50_001).astype(np.int8)
    yi = np.empty(xi.size, np.int32)
    port.port_scan_cpu_i8(xi.ctypes.data_as(ctypes.c_void_p), yi.ctypes.data_as(ctypes.c_void_p),
                          ctypes.c_int64(xi.size), threads)
    assert_exact(yi, oracle.scan(xi))
