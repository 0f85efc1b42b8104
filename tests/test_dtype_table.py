"""Accepted input dtypes of every hot-path binding, against the reference
wrappers (CPU only: no kernel runs).

Each binding checks its arguments' dtypes before it checks that they live on
the GPU, so a call with CPU tensors tells the two apart: an accepted dtype
ends in the device error, a rejected one in a dtype error.  REFERENCE lists,
per binding, the dtypes that the reference wrapper dispatches to a kernel
(file:line in src/torch/); SUPERSET lists what this build accepts beyond
them on purpose.  Every reference dtype must be accepted, and everything
outside REFERENCE | SUPERSET must be rejected.
"""
import pytest
import torch

import tcuscan_ops as ops

F16, BF16, F32, F64 = torch.float16, torch.bfloat16, torch.float32, torch.float64
I8, I16, I32, I64, U8 = torch.int8, torch.int16, torch.int32, torch.int64, torch.uint8

ALL = [F16, BF16, F32, F64, I8, I16, I32, I64, U8]

# binding -> (reference-accepted x dtypes, source, deliberate superset)
REFERENCE = {
    # torch_scan.h:75-101 (int8, fp16, fp32 kernels)
    "run_scan_single_core": ({F16, F32, I8}, {BF16}),
    # torch_scan.h:216-235: fp16, every other dtype runs the int8 kernel;
    # only fp16 and int8 are valid inputs.
    "run_scan_multi_core": ({F16, I8}, {BF16}),
    "run_scan_multi_core_no_l2": ({F16, I8}, {BF16}),
    # torch_scan.h:149-167
    "run_scan_batch": ({F16, F32}, {BF16, I8}),
    # torch_scan.h:337-345 (fp16 only; other dtypes return uninitialised memory)
    "run_row_scan": ({F16}, {BF16, F32, I8}),
    # torch_scan.h:389-398
    "run_block_scan": ({F16}, {BF16, F32, I8}),
    # torch_scan.h:447-456
    "run_scan_multi_cube": ({F16}, {BF16}),
    # torch_reduce.h:59-69
    "run_reduce_tiles": ({F16, I8}, {BF16, F32, I32}),
    # torch_reduce.h:165-175 (int64 runs the int32 kernel, output int32)
    "run_complete_rows": ({F32, I64}, {I32}),
    # torch_reduce.h:214-224
    "run_complete_blocks": ({F32, I32}, set()),
    # torch_seg_ops.h:239-256
    "run_seg_sum_single_core": ({F16, I8}, {BF16, F32}),
    # torch_seg_ops.h:388-414
    "run_seg_sum_multi_core": ({F16, I8}, {BF16, F32}),
    # torch_seg_ops.h:300-310
    "run_seg_sum_single_cube": ({F16}, {BF16, F32, I8}),
    # torch_seg_ops.h:452-453 (TORCH_CHECK fp16)
    "run_seg_sum_multi_cube": ({F16}, set()),
    # torch_seg_ops.h:85-107: run_scan_multi_core(x) then compaction
    "run_seg_sum": ({F16, I8}, {BF16, F32}),
    # torch_gather.h:103-105
    "run_csr_gather": ({F16, I16, F32}, set()),
    # torch_spmv.h:117-120 (vals and x both fp16 or both fp32)
    "run_spmv_v2": ({F16, F32}, set()),
    # torch_compress.h:78-92: 16-bit types run the 16-bit kernel, every other
    # dtype the 32-bit one (so 2- and 4-byte types are meaningful); this build
    # copies 1-, 2-, 4- and 8-byte elements bit for bit.
    "run_compress": ({F16, I16, F32, I32}, {BF16, F64, I8, I64, U8}),
    "run_compress_pos": ({F16, I16, F32, I32}, {BF16, F64, I8, I64, U8}),
    "run_compress_ind": ({F16, I16, F32, I32}, {BF16, F64, I8, I64, U8}),
    "run_compress_ind_no_arange": ({F16, I16, F32, I32}, {BF16, F64, I8, I64, U8}),
}


def _args(name, dt):
    """CPU arguments of one call of `name` with x (or vals) of dtype dt."""
    n = 256
    x = torch.zeros(n, dtype=dt)
    i8 = torch.zeros(n, dtype=I8)
    indptr = torch.tensor([0, 100, n], dtype=I32)
    tri = torch.ones(16, 16, dtype=F16)
    if name in ("run_scan_single_core", "run_scan_multi_core", "run_scan_multi_core_no_l2",
                "run_scan_batch", "run_row_scan"):
        return (x.reshape(16, 16) if name != "run_scan_single_core" else x, 16)
    if name in ("run_block_scan", "run_scan_multi_cube"):
        return (x, tri, tri)
    if name == "run_reduce_tiles":
        return (x, 16, 4)
    if name == "run_complete_rows":
        return (x, torch.zeros(4, dtype=dt), 16, 4)
    if name == "run_complete_blocks":
        return (x, torch.zeros(4, dtype=dt), 16)
    if name in ("run_seg_sum_single_core", "run_seg_sum_multi_core"):
        return (x, indptr, 16)
    if name == "run_seg_sum_single_cube":
        return (x, tri, tri, indptr[:-1], 16)
    if name == "run_seg_sum_multi_cube":
        return (x, tri, tri, indptr)
    if name == "run_seg_sum":
        return (x, i8, 16)
    if name == "run_csr_gather":
        return (x, torch.zeros(n, dtype=I32), torch.zeros(n, dtype=dt))
    if name == "run_spmv_v2":
        return (x, indptr, torch.zeros(n, dtype=I32), torch.zeros(n, dtype=dt), 16)
    if name == "run_compress":
        return (x, i8, 16)
    if name == "run_compress_pos":
        return (x, i8, 0, 16)
    if name == "run_compress_ind":
        return (x, torch.arange(n, dtype=I32), i8, 16)
    if name == "run_compress_ind_no_arange":
        return (x, i8, 16)
    raise KeyError(name)


def _outcome(name, dt):
    with pytest.raises(RuntimeError) as e:
        getattr(ops, name)(*_args(name, dt))
    msg = str(e.value)
    if "must be a GPU tensor" in msg:
        return "accepted"
    return "rejected"


@pytest.mark.parametrize("name", sorted(REFERENCE))
def test_accepted_dtypes(name):
    ref, extra = REFERENCE[name]
    got = {dt for dt in ALL if _outcome(name, dt) == "accepted"}
    missing = ref - got
    assert not missing, f"{name}: reference dtypes rejected: {sorted(map(str, missing))}"
    unexpected = got - ref - extra
    assert not unexpected, f"{name}: accepted beyond the stated superset: {sorted(map(str, unexpected))}"
    absent = extra - got
    assert not absent, f"{name}: stated superset not accepted: {sorted(map(str, absent))}"


def test_scan_cpu_dtypes():
    """run_scan_cpu runs on the CPU: fp16, fp32, int8 (torch_scan_cpu.h:98-102)."""
    for dt, out in ((F16, F32), (F32, F32), (I8, I32)):
        assert ops.run_scan_cpu(torch.ones(8, dtype=dt)).dtype == out
    for dt in (BF16, F64, I16, I32, I64):
        with pytest.raises(RuntimeError, match="must be float16, float32 or int8"):
            ops.run_scan_cpu(torch.ones(8, dtype=dt))
