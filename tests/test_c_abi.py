"""CPU tests of the C ABI: the library loads, exports every declared entry
point, and rejects invalid arguments before anything reaches the GPU."""
import ctypes
import os
import re

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "tcuscan.h")
HEADERS = sorted(os.path.join(REPO, "include", f) for f in os.listdir(os.path.join(REPO, "include"))
                 if f.endswith(".h"))
LIB = os.path.join(REPO, "pytorch-tcuscan_amd", "libtcuscan_hip.so")

TCS_OK, TCS_ERR_INVALID_ARG, TCS_ERR_UNSUPPORTED_DTYPE, TCS_ERR_WORKSPACE = 0, 1, 2, 3


def declared_functions():
    """Every function declared in include/*.h."""
    names = set()
    for h in HEADERS:
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        names.update(re.findall(r"\b(tcs_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


@pytest.fixture(scope="module")
def lib():
    assert os.path.exists(LIB), "libtcuscan_hip.so is not built"
    return ctypes.CDLL(LIB)


def test_exports_every_declared_symbol(lib):
    names = declared_functions()
    assert len(names) >= 5
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"


def test_status_strings_and_version(lib):
    lib.tcs_status_string.restype = ctypes.c_char_p
    assert lib.tcs_status_string(0) == b"ok"
    assert b"dtype" in lib.tcs_status_string(2)
    assert lib.tcs_abi_version() >= 100


def _scan(lib, x, dtype, y, n, S, ws, ws_bytes):
    f = lib.tcs_scan
    f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int,
                  ctypes.c_int, ctypes.c_double, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
    return f(x, dtype, y, n, S, 0, 0.0, ws, ws_bytes, None)


def test_scan_argument_validation(lib):
    lib.tcs_scan_workspace_bytes.restype = ctypes.c_size_t
    lib.tcs_scan_workspace_bytes.argtypes = [ctypes.c_int64, ctypes.c_int, ctypes.c_int]
    need = lib.tcs_scan_workspace_bytes(1 << 20, 128, 0)
    assert need >= 256 + 16 * ((1 << 20) // 16384)
    fake = 0x100000  # aligned, never dereferenced: every call below fails validation
    assert _scan(lib, fake, 0, fake, 0, 128, None, 0) == TCS_OK  # empty input: no launch
    assert _scan(lib, fake, 0, fake, -1, 128, fake, need) == TCS_ERR_INVALID_ARG
    assert _scan(lib, fake, 0, fake, 1 << 20, 0, fake, need) == TCS_ERR_INVALID_ARG
    assert _scan(lib, None, 0, fake, 1 << 20, 128, fake, need) == TCS_ERR_INVALID_ARG
    assert _scan(lib, fake + 2, 0, fake, 1 << 20, 128, fake, need) == TCS_ERR_INVALID_ARG
    assert _scan(lib, fake, 0, fake, 1 << 20, 128, fake, 256) == TCS_ERR_WORKSPACE
    assert _scan(lib, fake, 0, fake, 1 << 20, 128, None, need) == TCS_ERR_WORKSPACE
    assert _scan(lib, fake, 99, fake, 1 << 20, 128, fake, need) == TCS_ERR_UNSUPPORTED_DTYPE
    assert _scan(lib, fake, 7, fake, 1 << 20, 128, fake, need) == TCS_ERR_UNSUPPORTED_DTYPE


def test_scan_method_argument_validation(lib):
    f = lib.tcs_scan_method
    f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int,
                  ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                  ctypes.c_size_t, ctypes.c_void_p]
    fake = 0x100000
    args = (fake, 0, fake, 1 << 20, 128, 0, 0.0)
    assert f(*args, 3, 0, fake, 1 << 24, None) == TCS_ERR_INVALID_ARG          # no such method
    assert f(*args, -1, 0, fake, 1 << 24, None) == TCS_ERR_INVALID_ARG
    assert f(*args, 1, 65, fake, 1 << 24, None) == TCS_ERR_INVALID_ARG         # window > 64
    assert f(*args, 1, -2, fake, 1 << 24, None) == TCS_ERR_INVALID_ARG
    for dt in (2, 5):  # f32, i32: no matrix-unit path
        assert f(fake, dt, fake, 1 << 20, 128, 0, 0.0, 2, 0, fake, 1 << 24, None) == TCS_ERR_UNSUPPORTED_DTYPE
    assert f(*args, 2, 0, fake, 256, None) == TCS_ERR_WORKSPACE
    assert f(fake, 0, fake, 0, 128, 0, 0.0, 2, 0, None, 0, None) == TCS_OK  # empty: no launch


def test_scan_rows_argument_validation(lib):
    f = lib.tcs_scan_rows
    f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64,
                  ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
    fake = 0x100000
    assert f(fake, 0, fake, 1000, 0, 64, 0, fake, 1 << 20, None) == TCS_ERR_INVALID_ARG
    assert f(fake, 42, fake, 1000, 10, 64, 0, fake, 1 << 20, None) == TCS_ERR_UNSUPPORTED_DTYPE


def test_extension_rejects_cpu_tensors():
    import torch
    import tcuscan_ops

    with pytest.raises(RuntimeError, match="GPU tensor"):
        tcuscan_ops.run_scan_multi_core(torch.ones(4, dtype=torch.float16), 64)


def test_integer_scan_rejects_unconvertible_start(lib):
    """An int8 / int32 scan's starting sum must be a finite int32; the check
    comes before the workspace check, so the null workspace below is never
    reached (and nothing is launched)."""
    f = lib.tcs_scan
    f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int,
                  ctypes.c_int, ctypes.c_double, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
    fake = 0x100000
    for dt in (3, 5):  # int8, int32
        for start in (2.0 ** 31, -(2.0 ** 31) - 1, float("inf"), float("-inf"), float("nan"), 1e300):
            assert f(fake, dt, fake, 1000, 128, 0, start, None, 0, None) == TCS_ERR_INVALID_ARG, (dt, start)
        for start in (2.0 ** 31 - 1, -(2.0 ** 31), 0.0, -5.0):
            assert f(fake, dt, fake, 1000, 128, 0, start, None, 0, None) == TCS_ERR_WORKSPACE, (dt, start)
    # Float scans round start to float: every double is accepted.
    assert f(fake, 0, fake, 1000, 128, 0, float("inf"), None, 0, None) == TCS_ERR_WORKSPACE


def test_workspace_maintenance_validation(lib):
    reset = lib.tcs_workspace_reset
    reset.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
    polls = lib.tcs_workspace_set_poll_limit
    polls.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_void_p]
    fake = 0x100000
    assert reset(None, 1 << 20, None) == TCS_ERR_WORKSPACE
    assert reset(fake, 64, None) == TCS_ERR_WORKSPACE          # smaller than the header
    assert polls(None, 1 << 20, 0, None) == TCS_ERR_WORKSPACE
    assert polls(fake, 1 << 20, 513, None) == TCS_ERR_INVALID_ARG  # above the default limit


def test_tuner_cache_entries_are_validated(lib):
    f = lib.tcs_batch_tuned_set
    f.argtypes = [ctypes.c_int] * 7 + [ctypes.c_float]
    assert f(1, 0, 20, 32, 16, 1, 8, 1.0) == TCS_OK
    assert f(0, 3, 28, 128, 4, 2, 32, 1.0) == TCS_OK          # int8 scan on the matrix units
    assert f(1, 0, 20, 48, 16, 1, 8, 1.0) == TCS_ERR_INVALID_ARG  # S not 32/64/128
    assert f(1, 0, 20, 32, 3, 1, 8, 1.0) == TCS_ERR_INVALID_ARG   # tiles_per_wg not a power of two
    assert f(1, 0, 20, 32, 16, 3, 8, 1.0) == TCS_ERR_INVALID_ARG  # no such method
    assert f(1, 0, 20, 32, 16, 1, 65, 1.0) == TCS_ERR_INVALID_ARG  # window past a wave
    assert f(2, 0, 20, 32, 16, 1, 8, 1.0) == TCS_ERR_INVALID_ARG   # seg-sums are not tuned
    assert f(1, 0, 20, 32, 16, 2, 8, 1.0) == TCS_ERR_UNSUPPORTED_DTYPE  # no matrix-unit seg-scan
    n = lib.tcs_batch_tuned_count()
    got = []
    for i in range(n):
        v = [ctypes.c_int() for _ in range(7)] + [ctypes.c_float()]
        assert lib.tcs_batch_tuned_entry(i, *[ctypes.byref(x) for x in v]) == TCS_OK
        got.append(tuple(x.value for x in v[:7]))
    assert (1, 0, 20, 32, 16, 1, 8) in got and (0, 3, 28, 128, 4, 2, 32) in got
