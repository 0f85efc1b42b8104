"""CPU oracle for the tcuscan hot path -- TEST INFRASTRUCTURE ONLY.

This module restates, on the CPU and in numpy, what each reference operator
computes.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg may import it, and only as the checker; the product path never calls it.

Floating-point results are computed in float64 and rounded once to the
output dtype; integer results are exact (int64 internally, wrapped to int32
where the reference emits int32).  Each function cites the reference file
it follows (paths relative to the huawei-csl/pytorch-tcuscan root).

Pinning (see tests/test_oracle.py): every function is checked against a
closed form, an invariant, a library routine (numpy/scipy/torch CPU) or a
brute-force loop chosen so that a dropped term, a wrong index or an
off-by-one segment boundary fails.  There are no stored golden vectors in
the reference (its tests draw seeded random data on the device), and the one
docstring known-answer example (kernel_seg_sum_vec_revert.h:22-24) is
internally inconsistent, so it is checked in its corrected form.
"""
from __future__ import annotations

import math

import numpy as np

__all__ = [
    "scan", "scan_rows", "seg_scan_flags", "blocked_scan_cpu", "reduce_tiles",
    "work_distribution", "complete_rows", "complete_blocks", "seg_sum_indptr",
    "seg_sum_flags", "csr_gather", "spmv", "compress", "compress_ind",
    "compress_ind_no_arange", "out_dtype_for", "compare_mask", "count_if", "filter_cmp",
    "searchsorted", "split", "sort", "topk", "topk_pivot", "top_p_filter", "seg_scan_revert",
]

_FLOAT_IN = ("float16", "bfloat16", "float32")


def _dtype_name(a) -> str:
    return str(a.dtype)


def out_dtype_for(in_dtype_name: str) -> np.dtype:
    """fp16/bf16/fp32 -> float32, int8/int32 -> int32.

    Reference: src/host_utils.h:108-134 (CubeOutType), src/torch/torch_scan.h:52-54.
    """
    if in_dtype_name in _FLOAT_IN:
        return np.dtype(np.float32)
    return np.dtype(np.int32)


def _as_acc(x: np.ndarray) -> np.ndarray:
    """Widen to the exact accumulation type: float64 for floats, int64 for ints."""
    if x.dtype.kind == "f":
        return x.astype(np.float64)
    return x.astype(np.int64)


def _finish(acc: np.ndarray, is_float: bool) -> np.ndarray:
    if is_float:
        return acc.astype(np.float32)
    # int32 two's-complement wrap, as torch.cumsum(dtype=int32).
    return acc.astype(np.int64).astype(np.uint64).astype(np.uint32).view(np.int32)


def scan(x: np.ndarray, start: float = 0.0) -> np.ndarray:
    """Inclusive prefix sum plus `start` added to every output.

    Reference: run_scan_multi_core (src/torch/torch_scan.h:184-241) and
    run_scan_single_core with starting_sum (src/torch/torch_scan.h:47-110,
    src/kernels/kernel_scan_single_core.h:156-244); tests compare against
    torch.cumsum(x, dtype=fp32/int32) (tests/test_scan_multi_core.py:45).
    `start` is rounded to float32 (int32 for integer input) first, as the
    wrapper stores it in the tiling struct (torch_scan.h:65-71).
    """
    x = np.asarray(x)
    is_float = x.dtype.kind == "f"
    acc = np.cumsum(_as_acc(x))
    if is_float:
        acc = acc + float(np.float32(start))
    else:
        acc = acc + int(np.int32(int(start)))
    return _finish(acc, is_float)


def scan_rows(x: np.ndarray, row_len: int) -> np.ndarray:
    """Independent inclusive scans of consecutive rows of `row_len` elements.

    Reference: run_scan_batch (src/torch/torch_scan.h:119-175,
    kernel_scan_batch.h:319-390), run_row_scan (torch_scan.h:312-351,
    kernel_row_scan.h:19-46: C = A @ U with U upper-triangular ones) and
    run_block_scan (torch_scan.h:362-405, kernel_block_scan.h:16-45:
    scan of each S*S block).  The last row may be shorter.
    """
    x = np.asarray(x).reshape(-1)
    is_float = x.dtype.kind == "f"
    n = x.size
    acc = _as_acc(x)
    out = np.empty(n, dtype=acc.dtype)
    full = (n // row_len) * row_len
    if full:
        out[:full] = np.cumsum(acc[:full].reshape(-1, row_len), axis=1).reshape(-1)
    if full < n:
        out[full:] = np.cumsum(acc[full:])
    return _finish(out, is_float)


def seg_scan_flags(x: np.ndarray, f: np.ndarray) -> np.ndarray:
    """Segmented inclusive scan, f[i] != 0 starts a segment at i.

    Reference: tests/test_seg_scan_single_core.py:40-63 (ref_segscan) and
    src/torch/torch_seg_ops.h:164-193 (run_seg_scan).  Computed as the
    global float64/int64 prefix minus the prefix before each segment's head,
    which is exact whenever the global prefix is (fp16 and integer inputs).
    """
    x = np.asarray(x).reshape(-1)
    f = np.asarray(f).reshape(-1) != 0
    is_float = x.dtype.kind == "f"
    acc = _as_acc(x)
    c = np.cumsum(acc)
    seg_id = np.cumsum(f)  # 0 for elements before the first head
    starts = np.flatnonzero(f)
    before = np.zeros(starts.size + 1, dtype=acc.dtype)
    before[1:] = c[starts] - acc[starts]
    return _finish(c - before[seg_id], is_float)


def blocked_scan_cpu(x: np.ndarray, threads: int) -> np.ndarray:
    """Three-phase blocked scan in the output precision (float32 / int32).

    Restates parallel_inclusive_scan_blocked (src/torch/torch_scan_cpu.h:23-80)
    and run_scan_cpu (:88-107): blocks = min(max(1, threads), n), block size
    ceil(n / blocks); phase 1 scans each block sequentially in the output
    type, phase 2 scans block totals serially, phase 3 adds each block's
    offset.  The rounding therefore depends on `threads`.
    """
    x = np.asarray(x).reshape(-1)
    is_float = x.dtype.kind == "f"
    t = np.float32 if is_float else np.int32
    z = x.astype(t).copy()
    n = z.size
    if n == 0:
        return z
    blocks = min(max(1, threads), n)
    bs = (n + blocks - 1) // blocks
    totals = np.zeros(blocks, dtype=t)
    with np.errstate(over="ignore"):
        for b in range(blocks):
            s, e = b * bs, min(b * bs + bs, n)
            if s >= e:
                totals[b] = t(0)
                continue
            # Sequential float32 accumulation (np.cumsum on float32 is sequential).
            z[s:e] = np.cumsum(z[s:e], dtype=t)
            totals[b] = z[e - 1]
        running = t(0)
        prefix = np.zeros(blocks, dtype=t)
        for b in range(blocks):
            prefix[b] = running
            running = t(running + totals[b])
        for b in range(blocks):
            if prefix[b] == 0:
                continue
            s, e = b * bs, min(b * bs + bs, n)
            z[s:e] = (z[s:e] + prefix[b]).astype(t)
    return z


def work_distribution(vec_len: int, tile_len: int, num_blocks: int):
    """Contiguous block ranges used by the reference's multi-core kernels.

    Reference: scalar::GetWorkDistribution (src/kernels/tcuscan_utils.h:1389-1408):
    max_tiles = ceil(ceil(vec_len / tile_len) / num_blocks); block b covers
    [b * max_tiles * tile_len, min((b+1) * max_tiles * tile_len, vec_len)).
    Returns a list of (start, end) with empty ranges for idle blocks.
    """
    num_tiles = -(-vec_len // tile_len)
    per = -(-num_tiles // num_blocks) if num_blocks else 0
    blen = per * tile_len
    out = []
    for b in range(num_blocks):
        s = min(b * blen, vec_len)
        e = min(s + blen, vec_len)
        out.append((s, e))
    return out


def reduce_tiles(x: np.ndarray, tile_len: int, num_blocks: int) -> np.ndarray:
    """Sum of each contiguous block of ceil(tiles/num_blocks)*tile_len elements.

    Reference: KernelReduceTiles (src/kernels/kernel_reduce_tiles.h:79-135),
    tests/test_reduce_tiles.py:37-47 (ref_reduce_tiles).  fp16 -> float32,
    int8 -> int32 (exact).
    """
    x = np.asarray(x).reshape(-1)
    is_float = x.dtype.kind == "f"
    acc = _as_acc(x)
    out = np.zeros(num_blocks, dtype=acc.dtype)
    for b, (s, e) in enumerate(work_distribution(x.size, tile_len, num_blocks)):
        if e > s:
            out[b] = acc[s:e].sum()
    return _finish(out, is_float)


def _complete_rows_height(tile_width: int, tile_height: int) -> int:
    # tile_height is capped so that a tile holds at most 8192 elements
    # (kernel_complete_rows.h:35,49-50).
    return min(tile_height, -(-8192 // tile_width))


def complete_rows(x: np.ndarray, sums: np.ndarray, tile_width: int, tile_height: int) -> np.ndarray:
    """Down-sweep of MCSCAN: block b gets the exclusive prefix of `sums`, then
    each row of `tile_width` elements gets the running carry and the carry
    becomes the row's last output.

    Reference: KernelCompleteRows (src/kernels/kernel_complete_rows.h:99-161)
    with blocks from GetWorkDistribution over tiles of tile_width*height
    elements, height capped at ceil(8192/tile_width); tests/test_complete_rows.py:27-55.
    Accumulated in float64 / int64 and rounded once.
    """
    x = np.asarray(x).reshape(-1)
    sums = np.asarray(sums).reshape(-1)
    is_float = x.dtype.kind == "f"
    acc = _as_acc(x)
    nb = sums.size
    h = _complete_rows_height(tile_width, tile_height)
    tile = tile_width * h
    prefix = np.concatenate([[0], np.cumsum(_as_acc(sums))[:-1]]).astype(acc.dtype)
    out = np.empty_like(acc)
    w = tile_width
    for b, (s, e) in enumerate(work_distribution(x.size, tile, nb)):
        if e <= s:
            continue
        seg = acc[s:e]
        m = seg.size
        rows = -(-m // w)
        ends = np.minimum(np.arange(1, rows + 1) * w, m) - 1
        carry = np.empty(rows, dtype=acc.dtype)
        carry[0] = prefix[b]
        if rows > 1:
            carry[1:] = prefix[b] + np.cumsum(seg[ends[:-1]])
        out[s:e] = seg + np.repeat(carry, w)[:m]
    return _finish(out, is_float)


def complete_blocks(x: np.ndarray, sums: np.ndarray, tile_len: int) -> np.ndarray:
    """Adds sum(sums[:b]) to every element of block b (len(x)/len(sums) elements).

    Reference: KernelCompleteBlocks (src/kernels/kernel_complete_blocks.h:109-179),
    run_complete_blocks (src/complete_blocks.cpp:36-40: block_len = vec_len /
    num_blocks), tests/test_complete_blocks.py:27-49.  `tile_len` only tiles
    the work and does not change the result.
    """
    x = np.asarray(x).reshape(-1)
    sums = np.asarray(sums).reshape(-1)
    is_float = x.dtype.kind == "f"
    acc = _as_acc(x)
    nb = sums.size
    bl = x.size // nb
    prefix = np.concatenate([[0], np.cumsum(_as_acc(sums))[:-1]]).astype(acc.dtype)
    out = acc.copy()
    for b in range(nb):
        out[b * bl:(b + 1) * bl] += prefix[b]
    return _finish(out, is_float)


def seg_sum_indptr(x: np.ndarray, indptr: np.ndarray, ends_implicit: bool = False) -> np.ndarray:
    """y[r] = sum(x[indptr[r]:indptr[r+1]]).

    Reference: run_seg_sum_single_core / multi_core / multi_cube take the full
    CSR indptr (n+1 entries) and return n sums (src/torch/torch_seg_ops.h:211-266,
    333-426, 443-522); run_seg_sum_single_cube takes indptr[:-1] and returns
    len(indptr) sums, the last segment ending at len(x)
    (tests/test_seg_sum_single_cube.py:68-69,82-84).  Tests compare against
    scipy `csr @ ones` (tests/test_seg_sum_multi_core.py:96).
    """
    x = np.asarray(x).reshape(-1)
    ip = np.asarray(indptr).astype(np.int64).reshape(-1)
    if ends_implicit:
        ip = np.concatenate([ip, [x.size]])
    return _row_sums(_as_acc(x), ip, x.dtype.kind == "f")


def _row_sums(acc: np.ndarray, ip: np.ndarray, is_float: bool) -> np.ndarray:
    """Per-row sums of `acc` over CSR row pointers `ip`; empty rows give 0."""
    out = np.zeros(ip.size - 1, dtype=acc.dtype)
    nz = np.flatnonzero(ip[1:] > ip[:-1])
    if nz.size:
        # Consecutive non-empty rows tile [ip[nz[0]], ip[-1]) without gaps, so
        # reduceat over their starts yields each row's own sum.
        out[nz] = np.add.reduceat(acc[: ip[-1]], ip[:-1][nz])
    return _finish(out, is_float)


def seg_sum_flags(x: np.ndarray, f: np.ndarray) -> np.ndarray:
    """Segment sums where f[i] == 1 marks the LAST element of a segment.

    Reference: run_seg_sum (src/torch/torch_seg_ops.h:85-107) computes
    diff([0] ++ scan(x)[f == 1]); tests/test_segmented_sum.py:70-75 shifts
    start flags into end flags.  Output length is the number of ones; elements
    after the last flag belong to no segment.  Summed exactly per segment.
    """
    x = np.asarray(x).reshape(-1)
    ends = np.flatnonzero(np.asarray(f).reshape(-1) == 1)
    ip = np.concatenate([[0], ends + 1])
    return seg_sum_indptr(x, ip)


def csr_gather(vals: np.ndarray, cols: np.ndarray, x: np.ndarray) -> np.ndarray:
    """z[i] = vals[i] * x[cols[i]] rounded to the input dtype.

    Reference: KernelCSRGather (src/kernels/kernel_csr_gather.h:122-137, the
    product is computed and stored in the input dtype), checked bit-exactly
    against `values * x[cols]` (tests/test_csr_gather.py:60-61,77).
    """
    vals = np.asarray(vals)
    g = np.asarray(x)[np.asarray(cols).astype(np.int64)]
    if vals.dtype.kind == "f":
        return (vals.astype(np.float64) * g.astype(np.float64)).astype(vals.dtype)
    # Integer product wraps in the input width (int16 * int16 -> int16).
    prod = vals.astype(np.int64) * g.astype(np.int64)
    bits = vals.dtype.itemsize * 8
    prod = ((prod + (1 << (bits - 1))) % (1 << bits)) - (1 << (bits - 1))
    return prod.astype(vals.dtype)


def spmv(vals: np.ndarray, indptr: np.ndarray, cols: np.ndarray, x: np.ndarray) -> np.ndarray:
    """y = A @ x for CSR A, accumulated in float64, rounded to float32.

    Reference: run_spmv_v2 (src/torch/torch_spmv.h:112-209), tested against
    scipy `B @ vector` (tests/test_spmv_v2.py:73).  Products are taken exactly
    (the reference rounds them to fp16 first, kernel_csr_gather.h:133, which
    this build does not copy).
    """
    vals = np.asarray(vals).astype(np.float64)
    ip = np.asarray(indptr).astype(np.int64)
    prod = vals * np.asarray(x).astype(np.float64)[np.asarray(cols).astype(np.int64)]
    return _row_sums(prod, ip, True)


def compress(x: np.ndarray, mask: np.ndarray) -> np.ndarray:
    """Stable compaction x[mask == 1].

    Reference: KernelCompress keeps an element iff its int8 mask equals 1
    (src/kernels/kernel_compress.h:167-168); run_compress sizes the output as
    the number of ones (src/torch/torch_compress.h:62-67); tests compare with
    torch.masked_select (tests/test_compress.py:44).
    """
    return np.asarray(x).reshape(-1)[np.asarray(mask).reshape(-1) == 1]


def compress_ind(x: np.ndarray, idx_in: np.ndarray, mask: np.ndarray):
    """(x[mask == 1], idx_in[mask == 1]).

    Reference: run_compress_ind (src/torch/torch_compress.h:172-239,
    kernel_compress_ind.h:38-58), tests/test_compress_ind.py:46-47.
    """
    m = np.asarray(mask).reshape(-1) == 1
    return np.asarray(x).reshape(-1)[m], np.asarray(idx_in).reshape(-1)[m]


def compress_ind_no_arange(x: np.ndarray, mask: np.ndarray):
    """(x[mask == 1], positions of the kept elements as int32).

    Reference: run_compress_ind_no_arange (src/torch/torch_compress.h:249-309,
    kernel_compress_ind.h:75-93, KernelWhere kernel_where.h:20-238).
    """
    m = np.asarray(mask).reshape(-1) == 1
    return np.asarray(x).reshape(-1)[m], np.flatnonzero(m).astype(np.int32)


def lognormal_row_lengths(rows: int, mean: float, sigma: float, cap: int, rng) -> np.ndarray:
    """Row lengths round(lognormal) with the given arithmetic mean, clipped to [0, cap]."""
    mu = math.log(mean) - sigma * sigma / 2.0
    return np.clip(np.rint(rng.lognormal(mu, sigma, size=rows)), 0, cap).astype(np.int64)


# Comparison modes of run_count_if (AscendC CMPMODE, tests/test_count_if.py:44-50).
CMP_LT, CMP_GT, CMP_EQ, CMP_LE, CMP_GE, CMP_NE = range(6)


def compare_mask(x: np.ndarray, pivot: float, cmp: int) -> np.ndarray:
    """Boolean (x cmp pivot), with the pivot rounded to float32 (the
    reference's `float pivot` argument) and then to x's type, and the
    comparison made in that type (src/compare.cpp:25-39 casts the pivot to
    half; torch's `x <= pivot` on a half tensor does the same)."""
    x = np.asarray(x).reshape(-1)
    p = np.array(np.float32(pivot)).astype(x.dtype)
    return [np.less, np.greater, np.equal, np.less_equal, np.greater_equal, np.not_equal][cmp](x, p)


def count_if(x: np.ndarray, pivot: float, cmp: int) -> int:
    """Number of i with (x[i] cmp pivot).  Reference: run_count_if
    (src/torch/torch_compare.h:33-67, src/kernels/kernel_count_if.h)."""
    return int(np.count_nonzero(compare_mask(x, pivot, cmp)))


def filter_cmp(x: np.ndarray, pivot: float, cmp: int, positions: bool = False) -> np.ndarray:
    """x[x >= pivot] (cmp CMP_GE) or x[x <= pivot] (cmp CMP_LE), compared as
    in compare_mask; with positions=True the int32 indices of the kept
    elements instead.

    Reference: run_filter_greater_equal / run_filter_less_equal / run_where
    (src/torch/torch_compress.h:319-398): a mask from greater_equal_fp16 or
    torch's `x <= pivot`, then run_compress; its tests compare with
    torch.masked_select and torch.argwhere (tests/test_filter_greater_equal.py:40-54,
    tests/test_where.py:36-50).
    """
    x = np.asarray(x).reshape(-1)
    keep = compare_mask(x, pivot, cmp)
    if positions:
        return np.flatnonzero(keep).astype(np.int32)
    return x[keep]

def searchsorted(sorted_arr: np.ndarray, values: np.ndarray) -> np.ndarray:
    """First j with sorted[j] >= v for each needle v (lower bound), int32:
    a bisection over [0, len) run for all needles at once.

    Reference: run_searchsorted (src/torch/torch_searchsorted.h:21-33), whose
    tests compare with torch.searchsorted(side='left', out_int32=True)
    (tests/test_searchsorted.py:28-31).
    """
    a = np.asarray(sorted_arr).reshape(-1).astype(np.int64)
    v = np.asarray(values).reshape(-1).astype(np.int64)
    lo = np.zeros(v.size, dtype=np.int64)
    hi = np.full(v.size, a.size, dtype=np.int64)
    while True:
        active = lo < hi
        if not active.any():
            break
        mid = (lo + hi) // 2
        right = active & (a[np.minimum(mid, max(a.size - 1, 0))] < v) if a.size else active & False
        lo = np.where(right, mid + 1, lo)
        hi = np.where(active & ~right, mid, hi)
    return lo.astype(np.int32)


def split(x: np.ndarray, mask: np.ndarray, idx: np.ndarray | None = None):
    """Stable partition: the x[i] with mask[i] == 1, in order, then every
    other x[i], in order; with idx, the same permutation applied to idx.

    Reference: run_split / run_split_ind (src/torch/torch_split.h:25-141);
    tests/test_split.py:52-66 and tests/test_split_ind.py compare the parts
    with torch.masked_select(x, mask == 1) and (x, mask == 0).
    """
    x = np.asarray(x).reshape(-1)
    keep = np.asarray(mask).reshape(-1) == 1
    order = np.concatenate([np.flatnonzero(keep), np.flatnonzero(~keep)])
    if idx is None:
        return x[order]
    return x[order], np.asarray(idx).reshape(-1)[order]


def sort(x: np.ndarray, descending: bool = False):
    """Stable sort with the input positions: (values, int32 positions).

    Order as torch.sort(stable=True): ascending or descending by value, ties
    in input order, -0.0 equal to +0.0, NaN after everything (before
    everything when descending).  Reference: run_radix_sort
    (src/torch/torch_sort.h:21-81, ascending); tests/test_radix_sort.py
    compares with torch.sort(x, descending=False).
    """
    x = np.asarray(x).reshape(-1)
    idx = np.arange(x.size)
    if x.dtype.kind == "f":
        nan = np.isnan(x)
        v = np.where(nan, 0.0, x.astype(np.float64))
        if descending:
            order = np.lexsort((idx, -v, ~nan))
        else:
            order = np.lexsort((idx, v, nan))
    else:
        v = x.astype(np.int64)
        order = np.lexsort((idx, -v if descending else v))
    return x[order], order.astype(np.int32)


def topk(x: np.ndarray, k: int, largest: bool = True):
    """The k largest (smallest) elements with their input positions, largest
    (smallest) first; ties by input position, so among equal elements the
    lowest positions are taken.  The first k entries of the stable sort.

    Reference: run_topk_fp16 / run_topk_int16 (src/torch/torch_topk.h:69-207);
    tests/test_topk.py compares the sorted values and sorted indices with
    torch.topk's.
    """
    v, i = sort(x, descending=largest)
    return v[:k], i[:k]


def topk_pivot(x: np.ndarray, num_samples: int = 32, blocks: int | None = None) -> np.ndarray:
    """Per block: over its first eight tiles of 32 * num_samples elements,
    the largest of (per tile) the 4th largest of the 8th largests of each 32
    consecutive elements (positions past the end count as -inf).

    Reference: run_topk_pivot_fp16 (src/torch/torch_topk.h:23-60) and
    KernelPivotTopkEstimator (src/kernels/kernel_topk_pivot.h, k_inner 8,
    k_outer 4, eight tiles per block).  Blocks: min(64, tiles), each taking
    ceil(tiles / blocks) consecutive tiles.  Non-NaN inputs.
    """
    x = np.asarray(x, dtype=np.float16).reshape(-1)
    tile = 32 * num_samples
    tiles = -(-x.size // tile)
    if blocks is None:
        blocks = min(64, tiles)
    per = -(-tiles // blocks)
    out = np.empty(blocks, dtype=np.float16)
    for b in range(blocks):
        best = None
        for it in range(min(8, per)):
            base = (b * per + it) * tile
            if base >= x.size:
                break
            t = np.full(tile, -np.inf, dtype=np.float16)
            seg = x[base:base + tile]
            t[:seg.size] = seg
            g8 = np.sort(t.reshape(num_samples, 32), axis=1)[:, ::-1][:, 7]
            est = np.sort(g8)[::-1][3]
            best = est if best is None else max(best, est)
        out[b] = best if best is not None else np.float16(-np.inf)
    return out


def top_p_filter(probs: np.ndarray, p: float):
    """(kept, order) of top-p filtering: probabilities sorted largest first
    (stable), entries whose preceding mass (float32 prefix sums rounded
    once) exceeds p set to 0, the rest renormalised (float32).

    Reference: sample_top_p in scripts/profiling/profile_tcuscan_ops.py:372-392
    (torch.sort descending, torch.cumsum, mask `probs_sum - probs_sort > p`).
    """
    v, order = sort(np.asarray(probs, dtype=np.float16), descending=True)
    mass = scan(v)
    kept = v.astype(np.float32)
    kept = np.where(mass - kept > np.float32(p), np.float32(0), kept).astype(np.float32)
    return kept / kept.sum(dtype=np.float64), order


def seg_scan_revert(scan_x: np.ndarray, scan_f: np.ndarray, seg_ends: np.ndarray) -> np.ndarray:
    """z[i] = scan_x[i] - seg_ends[scan_f[i] - 1] in float32 (scan_x[i] in
    segment 0).

    Reference: run_seg_scan_mc_revert (src/torch/torch_seg_ops.h:108-153);
    tests/test_seg_scan_mc_revert.py builds scan_x = cumsum(x), scan_f =
    cumsum(f) and seg_ends = scan_x at every segment's last element, and
    compares with a per-segment cumsum.
    """
    sx = np.asarray(scan_x, dtype=np.float32).reshape(-1)
    sf = np.asarray(scan_f).reshape(-1).astype(np.int64)
    d = np.asarray(seg_ends, dtype=np.float32).reshape(-1)
    j = sf - 1
    ok = (j >= 0) & (j < d.size)
    sub = np.where(ok, d[np.clip(j, 0, max(d.size - 1, 0))] if d.size else 0, 0).astype(np.float32)
    return (sx - sub).astype(np.float32)
