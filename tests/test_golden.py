"""Known-answer fixtures (tests/golden/small_cases.json, written by
tests/golden/make_fixtures.py with exact rational arithmetic and brute-force
loops): the oracle must reproduce every case bit for bit (CPU), and so must
the HIP operators (GPU; SpMV within rtol = atol = 1e-6, since it sums
products in double).  Covers the scans, segmented sums, SpMV, gather,
compaction and the consumers of scan (filters, count_if, split, sort,
top-k, searchsorted, the segmented-scan revert)."""
import json
import os

import numpy as np
import pytest
import torch

import oracle
from _util import assert_exact, to_np

HERE = os.path.dirname(os.path.abspath(__file__))
CASES = json.load(open(os.path.join(HERE, "golden", "small_cases.json")))


def f16(bits):
    return np.array(bits, dtype=np.uint16).view(np.float16)


def f32(bits):
    return np.array(bits, dtype=np.uint32).view(np.float32)


def i16(v):
    return np.array(v, dtype=np.int16)


# ---------------------------------------------------------------------------
# Oracle against the fixtures (CPU)
# ---------------------------------------------------------------------------

def test_oracle_scans():
    c = CASES["scan_f16"]
    assert_exact(oracle.scan(f16(c["x_f16"])), f32(c["y_f32"]), "scan_f16")
    c = CASES["scan_f16_start"]
    assert_exact(oracle.scan(f16(c["x_f16"]), c["start"]), f32(c["y_f32"]), "scan_f16_start")
    c = CASES["scan_i8"]
    assert_exact(oracle.scan(np.array(c["x_i8"], dtype=np.int8)), np.array(c["y_i32"], dtype=np.int32))
    c = CASES["seg_scan_f16"]
    assert_exact(oracle.seg_scan_flags(f16(c["x_f16"]), np.array(c["f"], dtype=np.int8)), f32(c["y_f32"]))
    c = CASES["scan_rows_f16"]
    assert_exact(oracle.scan_rows(f16(c["x_f16"]), c["row_len"]), f32(c["y_f32"]))


def test_oracle_segmented_sums():
    c = CASES["seg_sum_i8"]
    assert_exact(oracle.seg_sum_indptr(np.array(c["x_i8"], dtype=np.int8), np.array(c["indptr"])),
                 np.array(c["y_i32"], dtype=np.int32))
    c = CASES["seg_sum_f16"]
    assert_exact(oracle.seg_sum_indptr(f16(c["x_f16"]), np.array(c["indptr"])), f32(c["y_f32"]))
    c = CASES["seg_sum_single_cube_f16"]
    assert_exact(oracle.seg_sum_indptr(f16(c["x_f16"]), np.array(c["starts"]), ends_implicit=True),
                 f32(c["y_f32"]))
    c = CASES["seg_sum_flags_f16"]
    assert_exact(oracle.seg_sum_flags(f16(c["x_f16"]), np.array(c["f"], dtype=np.int8)), f32(c["y_f32"]))


def test_oracle_sparse_and_compaction():
    c = CASES["spmv_f16"]
    got = oracle.spmv(f16(c["vals_f16"]), np.array(c["indptr"]), np.array(c["cols"]), f16(c["x_f16"]))
    assert_exact(got.astype(np.float32), f32(c["y_f32"]), "spmv")
    c = CASES["csr_gather_f16"]
    got = oracle.csr_gather(f16(c["vals_f16"]), np.array(c["cols"]), f16(c["x_f16"]))
    assert_exact(got, f16(c["z_f16"]), "csr_gather_f16")
    c = CASES["csr_gather_i16"]
    assert_exact(oracle.csr_gather(i16(c["vals_i16"]), np.array(c["cols"]), i16(c["x_i16"])), i16(c["z_i16"]))
    c = CASES["compress_u32"]
    x = np.array(c["x_u32"], dtype=np.uint32).view(np.float32)
    m = np.array(c["mask"], dtype=np.int8)
    assert_exact(oracle.compress(x, m).view(np.uint32), np.array(c["z_u32"], dtype=np.uint32))
    zx, zi = oracle.compress_ind_no_arange(x, m)
    assert_exact(zi.astype(np.int64), np.array(c["pos"], dtype=np.int64))


def test_oracle_blocks():
    c = CASES["reduce_tiles_f16"]
    assert_exact(oracle.reduce_tiles(f16(c["x_f16"]), c["tile_len"], c["num_blocks"]), f32(c["y_f32"]))
    c = CASES["complete_blocks_f32"]
    got = oracle.complete_blocks(np.array(c["x_f32"], dtype=np.float32), np.array(c["sums_f32"], dtype=np.float32),
                                 c["tile_len"])
    assert_exact(got, f32(c["y_f32"]))
    c = CASES["complete_rows_f32"]
    got = oracle.complete_rows(np.array(c["x_f32"], dtype=np.float32), np.array(c["sums_f32"], dtype=np.float32),
                               c["tile_width"], c["tile_height"])
    assert_exact(got, f32(c["y_f32"]))


# ---------------------------------------------------------------------------
# HIP operators against the fixtures (GPU)
# ---------------------------------------------------------------------------

def _t(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


@pytest.mark.gpu
@pytest.mark.parametrize("S", [16, 32, 128])
def test_gpu_scans(ops, dev, S):
    c = CASES["scan_f16"]
    assert_exact(to_np(ops.run_scan_multi_core(_t(f16(c["x_f16"]), dev), S)), f32(c["y_f32"]))
    for m in ("dpp", "mfma"):
        assert_exact(to_np(ops.run_scan_ex(_t(f16(c["x_f16"]), dev), S, 0, m)), f32(c["y_f32"]), m)
    c = CASES["scan_f16_start"]
    assert_exact(to_np(ops.run_scan_single_core(_t(f16(c["x_f16"]), dev), S, c["start"])), f32(c["y_f32"]))
    c = CASES["scan_i8"]
    assert_exact(to_np(ops.run_scan_multi_core(_t(np.array(c["x_i8"], dtype=np.int8), dev), S)),
                 np.array(c["y_i32"], dtype=np.int32))
    c = CASES["seg_scan_f16"]
    assert_exact(to_np(ops.run_seg_scan(_t(f16(c["x_f16"]), dev), _t(np.array(c["f"], dtype=np.int8), dev), S)),
                 f32(c["y_f32"]))
    c = CASES["scan_rows_f16"]
    x = f16(c["x_f16"])[: (300 // 37) * 37].reshape(-1, 37)
    assert_exact(to_np(ops.run_scan_batch(_t(x, dev), S)).reshape(-1), f32(c["y_f32"])[: x.size])


@pytest.mark.gpu
def test_gpu_segmented_sums(ops, dev):
    c = CASES["seg_sum_i8"]
    x, ip = _t(np.array(c["x_i8"], dtype=np.int8), dev), _t(np.array(c["indptr"], dtype=np.int32), dev)
    for fn in (ops.run_seg_sum_single_core, ops.run_seg_sum_multi_core):
        assert_exact(to_np(fn(x, ip, 128)), np.array(c["y_i32"], dtype=np.int32))
    c = CASES["seg_sum_f16"]
    x, ip = _t(f16(c["x_f16"]), dev), _t(np.array(c["indptr"], dtype=np.int32), dev)
    assert_exact(to_np(ops.run_seg_sum_multi_core(x, ip, 128)), f32(c["y_f32"]))
    c = CASES["seg_sum_single_cube_f16"]
    up = torch.ones(128, 128, dtype=torch.float16, device=dev)
    st = _t(np.array(c["starts"], dtype=np.int32), dev)
    assert_exact(to_np(ops.run_seg_sum_single_cube(x, up, up, st, 128)), f32(c["y_f32"]))
    c = CASES["seg_sum_flags_f16"]
    got = ops.run_seg_sum(_t(f16(c["x_f16"]), dev), _t(np.array(c["f"], dtype=np.int8), dev), 128)
    assert_exact(to_np(got), f32(c["y_f32"]))


@pytest.mark.gpu
def test_gpu_sparse_and_compaction(ops, dev):
    c = CASES["spmv_f16"]
    y = ops.run_spmv_v2(_t(f16(c["vals_f16"]), dev), _t(np.array(c["indptr"], dtype=np.int32), dev),
                        _t(np.array(c["cols"], dtype=np.int32), dev), _t(f16(c["x_f16"]), dev), 128)
    np.testing.assert_allclose(to_np(y), f32(c["y_f32"]), rtol=1e-6, atol=1e-6)
    c = CASES["csr_gather_f16"]
    z = ops.run_csr_gather(_t(f16(c["vals_f16"]), dev), _t(np.array(c["cols"], dtype=np.int32), dev),
                           _t(f16(c["x_f16"]), dev))
    assert_exact(to_np(z), f16(c["z_f16"]))
    c = CASES["csr_gather_i16"]
    z = ops.run_csr_gather(_t(i16(c["vals_i16"]), dev), _t(np.array(c["cols"], dtype=np.int32), dev),
                           _t(i16(c["x_i16"]), dev))
    assert_exact(to_np(z), i16(c["z_i16"]))
    c = CASES["compress_u32"]
    x = _t(np.array(c["x_u32"], dtype=np.uint32).view(np.float32), dev)
    m = _t(np.array(c["mask"], dtype=np.int8), dev)
    assert_exact(to_np(ops.run_compress(x, m, 32)).view(np.uint32), np.array(c["z_u32"], dtype=np.uint32))
    zx, zi = ops.run_compress_ind_no_arange(x, m, 32)
    assert_exact(to_np(zi).astype(np.int64), np.array(c["pos"], dtype=np.int64))


@pytest.mark.gpu
def test_gpu_blocks(ops, dev):
    c = CASES["reduce_tiles_f16"]
    got = ops.run_reduce_tiles(_t(f16(c["x_f16"]), dev), c["tile_len"], c["num_blocks"])
    assert_exact(to_np(got), f32(c["y_f32"]))
    c = CASES["complete_blocks_f32"]
    got = ops.run_complete_blocks(_t(np.array(c["x_f32"], dtype=np.float32), dev),
                                  _t(np.array(c["sums_f32"], dtype=np.float32), dev), c["tile_len"])
    assert_exact(to_np(got), f32(c["y_f32"]))
    c = CASES["complete_rows_f32"]
    got = ops.run_complete_rows(_t(np.array(c["x_f32"], dtype=np.float32), dev),
                                _t(np.array(c["sums_f32"], dtype=np.float32), dev), c["tile_width"],
                                c["tile_height"])
    np.testing.assert_allclose(to_np(got), f32(c["y_f32"]), rtol=2e-6, atol=2e-6)


# ---------------------------------------------------------------------------
# Consumers of scan: oracle (CPU) and HIP operators (GPU) against fixtures
# ---------------------------------------------------------------------------

def test_oracle_consumers():
    c = CASES["filters_f16"]
    x = f16(c["x_f16"])
    assert_exact(oracle.filter_cmp(x, c["pivot"], oracle.CMP_GE), f16(c["ge_f16"]))
    assert_exact(oracle.filter_cmp(x, c["pivot"], oracle.CMP_LE), f16(c["le_f16"]))
    assert_exact(oracle.filter_cmp(x, c["pivot"], oracle.CMP_GE, positions=True), np.array(c["where"], np.int32))
    assert [oracle.count_if(x, c["pivot"], m) for m in range(6)] == c["count_if"]
    c = CASES["split_i16"]
    z, zi = oracle.split(i16(c["x_i16"]), np.array(c["mask"], np.int8), np.arange(1, len(c["mask"]) + 1))
    assert_exact(z, i16(c["z_i16"]))
    assert zi.tolist() == c["idx_out"]
    c = CASES["sort_f16"]
    assert oracle.sort(f16(c["x_f16"]))[1].tolist() == c["asc_idx"]
    assert oracle.sort(f16(c["x_f16"]), descending=True)[1].tolist() == c["desc_idx"]
    c = CASES["sort_i16"]
    assert oracle.sort(i16(c["x_i16"]))[1].tolist() == c["asc_idx"]
    assert oracle.sort(i16(c["x_i16"]), descending=True)[1].tolist() == c["desc_idx"]
    c = CASES["topk_f16"]
    assert oracle.topk(f16(c["x_f16"]), c["k"])[1].tolist() == c["largest_idx"]
    assert oracle.topk(f16(c["x_f16"]), c["k"], largest=False)[1].tolist() == c["smallest_idx"]
    c = CASES["searchsorted_i32"]
    assert oracle.searchsorted(np.array(c["sorted"], np.int32), np.array(c["values"], np.int32)).tolist() == c["out"]
    c = CASES["seg_scan_revert_f32"]
    got = oracle.seg_scan_revert(np.array(c["scan_x"], np.float32), np.array(c["scan_f"]), np.array(c["seg_ends"]))
    assert_exact(got, np.array(c["z"], np.float32))


@pytest.mark.gpu
def test_gpu_consumers(ops, dev):
    c = CASES["filters_f16"]
    x = _t(f16(c["x_f16"]), dev)
    assert_exact(to_np(ops.run_filter_greater_eq(x, c["pivot"], 32)), f16(c["ge_f16"]))
    assert_exact(to_np(ops.run_filter_less_eq(x, c["pivot"], 32)), f16(c["le_f16"]))
    assert to_np(ops.run_where(x, c["pivot"], 32)).tolist() == c["where"]
    assert [int(ops.run_count_if(x, c["pivot"], 128, m).item()) for m in range(6)] == c["count_if"]
    c = CASES["split_i16"]
    n = len(c["mask"])
    z, zi = ops.run_split_ind(_t(i16(c["x_i16"]), dev), _t(np.array(c["mask"], np.int8), dev),
                              torch.arange(1, n + 1, dtype=torch.int32, device=dev), 32)
    assert_exact(to_np(z), i16(c["z_i16"]))
    assert to_np(zi).tolist() == c["idx_out"]
    c = CASES["sort_f16"]
    xs = _t(f16(c["x_f16"]), dev)
    v, i = ops.run_radix_sort(xs, 32)
    assert to_np(i).tolist() == c["asc_idx"]
    assert_exact(to_np(v).view(np.uint16), f16(c["x_f16"])[c["asc_idx"]].view(np.uint16))
    assert to_np(ops.run_radix_sort_ex(xs, True)[1]).tolist() == c["desc_idx"]
    c = CASES["sort_i16"]
    assert to_np(ops.run_radix_sort(_t(i16(c["x_i16"]), dev), 32)[1]).tolist() == c["asc_idx"]
    c = CASES["topk_f16"]
    v, i = ops.run_topk(_t(f16(c["x_f16"]), dev), c["k"])
    assert to_np(i).tolist() == c["largest_idx"]
    assert to_np(ops.run_topk(_t(f16(c["x_f16"]), dev), c["k"], False)[1]).tolist() == c["smallest_idx"]
    c = CASES["searchsorted_i32"]
    got = ops.run_searchsorted(_t(np.array(c["sorted"], np.int32), dev), _t(np.array(c["values"], np.int32), dev))
    assert to_np(got).tolist() == c["out"]
    c = CASES["seg_scan_revert_f32"]
    got = ops.run_seg_scan_mc_revert(_t(np.array(c["scan_x"], np.float32), dev),
                                     _t(np.array(c["scan_f"], np.int32), dev),
                                     _t(np.array(c["seg_ends"], np.float32), dev))
    assert_exact(to_np(got), np.array(c["z"], np.float32))
