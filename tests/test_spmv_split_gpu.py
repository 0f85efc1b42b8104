"""The split SpMV gather (vectors above 2 MiB): the even workgroup of each
tile sums the products whose column lies in the lower half of x, the odd
one the upper half, on two look-back chains, and a combine pass adds the two
exact double partials of every row.  Checked against the float64 oracle,
and bit for bit against the single-chain path, which tcs_spmv_csr takes when
its workspace has no room for the split (called through the C ABI)."""
import ctypes
import os

import numpy as np
import pytest
import torch

import oracle
from _util import to_np

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(REPO, "pytorch-tcuscan_amd", "libtcuscan_hip.so")
TCS_F16, TCS_F32, TCS_I32 = 0, 2, 5


@pytest.fixture(scope="module")
def lib():
    lib = ctypes.CDLL(LIB)
    lib.tcs_spmv_csr.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                 ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64,
                                 ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
    lib.tcs_spmv_workspace_bytes.restype = ctypes.c_size_t
    lib.tcs_spmv_workspace_bytes.argtypes = [ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_int]
    lib.tcs_seg_sum_workspace_bytes.restype = ctypes.c_size_t
    lib.tcs_seg_sum_workspace_bytes.argtypes = [ctypes.c_int64]
    return lib


def _matrix(rows, x_len, mean, seed, dtype=np.float16, long_rows=()):
    rng = np.random.default_rng(seed)
    lens = oracle.lognormal_row_lengths(rows, mean, 1.2, 20000, rng)
    for r in long_rows:
        lens[r] = 20000
    lens[rows // 3: rows // 3 + 50] = 0  # a run of empty rows
    indptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    nnz = int(indptr[-1])
    row_id = np.repeat(np.arange(rows), lens)
    cols = rng.integers(0, x_len, nnz)
    cols = cols[np.lexsort((cols, row_id))].astype(np.int32)
    vals = (2 * rng.random(nnz) - 1).astype(dtype)
    x = (2 * rng.random(x_len) - 1).astype(dtype)
    return vals, indptr, cols, x


def _c_spmv(lib, dev, vals, indptr, cols, x, ws_bytes):
    """tcs_spmv_csr with a workspace of ws_bytes (zeroed, as the contract asks)."""
    nnz = int(indptr[-1])
    if nnz == 0:  # any valid device pointers: nothing is read
        vals, cols = np.zeros(8, vals.dtype), np.zeros(8, np.int32)
    v, ip, c, xv = (torch.from_numpy(a).to(dev) for a in (vals, indptr, cols, x))
    rows = indptr.size - 1
    y = torch.full((rows,), float("nan"), dtype=torch.float32, device=dev)
    ws = torch.zeros(ws_bytes, dtype=torch.uint8, device=dev)
    code = TCS_F16 if vals.dtype == np.float16 else TCS_F32
    stream = torch.cuda.current_stream().cuda_stream
    rc = lib.tcs_spmv_csr(v.data_ptr(), code, ip.data_ptr() + 4, TCS_I32, c.data_ptr(), xv.data_ptr(), x.size,
                          y.data_ptr(), rows, nnz, ws.data_ptr(), ws_bytes, stream)
    assert rc == 0
    torch.cuda.synchronize()
    return y.cpu().numpy(), ws


@pytest.mark.parametrize("dtype", [np.float16, np.float32], ids=["f16", "f32"])
@pytest.mark.parametrize("x_len", [(1 << 21) + 3, 3_000_001])
def test_split_matches_oracle_and_single_chain(lib, dev, dtype, x_len):
    vals, indptr, cols, x = _matrix(150_001, x_len, 16.0, x_len % 1000, dtype, long_rows=(7, 90_000))
    nnz, rows = vals.size, indptr.size - 1
    big = lib.tcs_spmv_workspace_bytes(nnz, rows, x_len, TCS_F16 if dtype == np.float16 else TCS_F32)
    small = lib.tcs_seg_sum_workspace_bytes(nnz)
    assert big > small, "a vector of this size takes the split"
    y_split, ws = _c_spmv(lib, dev, vals, indptr, cols, x, big)
    y_one, _ = _c_spmv(lib, dev, vals, indptr, cols, x, small)
    exp = oracle.spmv(vals, indptr, cols, x)
    np.testing.assert_allclose(y_split, exp, rtol=1e-6, atol=1e-6)
    assert np.array_equal(y_split.view(np.uint32), y_one.view(np.uint32)), "split and single chain differ"
    # The partial sums (the workspace's last rows * 16 bytes, rounded up to
    # 256) left no doubles behind.
    part_bytes = (rows * 16 + 255) // 256 * 256
    assert int(torch.count_nonzero(ws[big - part_bytes:])) == 0


def test_split_through_torch_op(ops, dev):
    vals, indptr, cols, x = _matrix(400_000, 1 << 22, 16.0, 5)
    y = ops.run_spmv_v2(*(torch.from_numpy(a).to(dev) for a in (vals, indptr, cols, x)), 128)
    np.testing.assert_allclose(to_np(y), oracle.spmv(vals, indptr, cols, x), rtol=1e-6, atol=1e-6)


def test_split_part_boundaries_and_specials(lib, dev):
    """Columns at the edges of both parts, infinities and NaN in vals and x:
    a product belongs to exactly one part, and the other part adds exactly
    0 (never 0 * inf)."""
    x_len = (1 << 21) + 7  # odd: parts [0, 1048580) and [1048580, x_len)
    w = (x_len + 1) // 2
    edge = np.array([0, 1, w - 2, w - 1, w, w + 1, x_len - 2, x_len - 1], dtype=np.int32)
    rows = 4096
    rng = np.random.default_rng(3)
    lens = rng.integers(0, 12, rows)
    indptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    nnz = int(indptr[-1])
    cols = edge[rng.integers(0, edge.size, nnz)]
    row_id = np.repeat(np.arange(rows), lens)
    cols = cols[np.lexsort((cols, row_id))]
    vals = (2 * rng.random(nnz) - 1).astype(np.float16)
    x = (2 * rng.random(x_len) - 1).astype(np.float16)
    vals[10] = np.inf
    vals[20] = -np.inf
    vals[30] = np.nan
    x[w] = np.inf
    x[w - 1] = np.float16(65504)
    big = lib.tcs_spmv_workspace_bytes(nnz, rows, x_len, TCS_F16)
    y_split, _ = _c_spmv(lib, dev, vals, indptr, cols, x, big)
    y_one, _ = _c_spmv(lib, dev, vals, indptr, cols, x, lib.tcs_seg_sum_workspace_bytes(nnz))
    with np.errstate(invalid="ignore", over="ignore"):
        exp = oracle.spmv(vals, indptr, cols, x)
    np.testing.assert_allclose(y_split, exp, rtol=1e-6, atol=1e-6, equal_nan=True)
    np.testing.assert_array_equal(np.isnan(y_split), np.isnan(y_one))
    ok = ~np.isnan(y_one)
    assert np.array_equal(y_split[ok].view(np.uint32), y_one[ok].view(np.uint32))


def test_split_empty_and_tiny(lib, dev):
    """All rows empty, and one nonzero, over a vector that takes the split."""
    x_len = 1 << 22
    x = np.ones(x_len, dtype=np.float16)
    for lens in (np.zeros(1000, np.int64), np.r_[np.zeros(500, np.int64), 1, np.zeros(499, np.int64)]):
        indptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
        nnz = int(indptr[-1])
        vals = np.full(nnz, 3.0, np.float16)
        cols = np.full(nnz, x_len - 1, np.int32)
        big = lib.tcs_spmv_workspace_bytes(nnz, lens.size, x_len, TCS_F16)
        y, _ = _c_spmv(lib, dev, vals, indptr, cols, x, big)
        np.testing.assert_array_equal(y, oracle.spmv(vals, indptr, cols, x))
