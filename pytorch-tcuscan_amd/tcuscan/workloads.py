"""Synthetic workloads of the batched and served forms (seeded, built on the
GPU where they are large).

uneven_scans      many independent fp16 scans with log-uniform lengths,
                  concatenated, with int8 segment-start flags;
geometric_segsums many independent indptr segmented sums over int8 data
                  with geometric segment lengths, concatenated, with one
                  global CSR indptr;
lognormal_csr     a CSR matrix with lognormal row lengths and sorted,
                  distinct columns per row (the SpMV workload's shape).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch


@dataclass
class UnevenScans:
    x: torch.Tensor          # fp16 [total], the scans back to back
    flags: torch.Tensor      # int8 [total], 1 at the first element of every scan
    offsets: np.ndarray      # int64 [num + 1], scan i is x[offsets[i]:offsets[i+1]]

    def pieces(self):
        """The scans as separate, freshly allocated tensors."""
        o = self.offsets
        return [self.x[o[i]:o[i + 1]].clone() for i in range(len(o) - 1)]


def uneven_scans(num: int, lo_log2: float, hi_log2: float, seed: int, device,
                 dtype=torch.float16) -> UnevenScans:
    """`num` scans, lengths floor(2^U(lo_log2, hi_log2)), values U[-1, 1]
    rounded to `dtype`."""
    rng = np.random.default_rng(seed)
    lens = np.floor(np.exp2(rng.uniform(lo_log2, hi_log2, num))).astype(np.int64)
    offsets = np.concatenate([[0], np.cumsum(lens)])
    total = int(offsets[-1])
    g = torch.Generator(device=device).manual_seed(seed)
    x = (torch.rand(total, generator=g, device=device) * 2 - 1).to(dtype)
    flags = torch.zeros(total, dtype=torch.int8, device=device)
    flags[torch.from_numpy(offsets[:-1]).to(device)] = 1
    return UnevenScans(x, flags, offsets)


@dataclass
class SegSums:
    x: torch.Tensor          # int8 [nnz], the problems' data back to back
    indptr: torch.Tensor     # int32 [rows + 1], global CSR row pointers
    data_off: np.ndarray     # int64 [num + 1], problem i's data range
    row_off: np.ndarray      # int64 [num + 1], problem i's rows range
    indptr_host: np.ndarray  # int64 copy of indptr

    def pieces(self):
        """(x_i, indptr_i) per problem as separate tensors (indptr_i int32,
        starting at 0)."""
        out = []
        for i in range(len(self.data_off) - 1):
            d0, d1 = self.data_off[i], self.data_off[i + 1]
            r0, r1 = self.row_off[i], self.row_off[i + 1]
            ip = self.indptr_host[r0:r1 + 1] - d0
            out.append((self.x[d0:d1].clone(), torch.from_numpy(ip.astype(np.int32)).to(self.x.device)))
        return out


def geometric_segsums(num: int, mean_len: float, seg_lo_log2: float, seg_hi_log2: float,
                      seed: int, device) -> SegSums:
    """`num` seg-sum problems over int8 data U{-128..127}; problem i has
    floor(2^U(seg_lo_log2, seg_hi_log2)) segments whose lengths are
    geometric with mean `mean_len` (at least 1)."""
    rng = np.random.default_rng(seed)
    nseg = np.floor(np.exp2(rng.uniform(seg_lo_log2, seg_hi_log2, num))).astype(np.int64)
    seg_len = rng.geometric(1.0 / mean_len, int(nseg.sum())).astype(np.int64)
    indptr = np.concatenate([[0], np.cumsum(seg_len)])
    row_off = np.concatenate([[0], np.cumsum(nseg)])
    data_off = indptr[row_off]
    nnz = int(indptr[-1])
    g = torch.Generator(device=device).manual_seed(seed)
    x = torch.randint(-128, 128, (nnz,), generator=g, device=device, dtype=torch.int8)
    return SegSums(x, torch.from_numpy(indptr.astype(np.int32)).to(device), data_off, row_off, indptr)


def lognormal_csr(rows: int, cols: int, mean: float, sigma: float, cap: int, seed: int, device,
                  dtype=torch.float16):
    """CSR (vals, indptr int32, cols int32, x) with lognormal row lengths of
    the given mean (mu = ln(mean) - sigma^2 / 2), clipped to [0, cap], and
    distinct sorted columns per row (random keys sorted within each row)."""
    rng = np.random.default_rng(seed)
    mu = np.log(mean) - sigma * sigma / 2
    lens = np.clip(np.rint(rng.lognormal(mu, sigma, rows)), 0, min(cap, cols)).astype(np.int64)
    indptr = np.concatenate([[0], np.cumsum(lens)])
    nnz = int(indptr[-1])
    g = torch.Generator(device=device).manual_seed(seed)
    row_of = torch.repeat_interleave(torch.arange(rows, device=device),
                                     torch.from_numpy(lens).to(device))
    # Distinct columns per row: a random permutation prefix per row is costly
    # on the host, so draw with replacement, then resolve duplicates by
    # sorting (row, col) keys and nudging repeats (rows are far shorter than
    # `cols`, so few repeats occur).
    c = torch.randint(0, cols, (nnz,), generator=g, device=device, dtype=torch.int64)
    for _ in range(16):
        key, _ = torch.sort(row_of * cols + c)
        c = key - row_of * cols
        dup = torch.zeros(nnz, dtype=torch.bool, device=device)
        if nnz > 1:
            dup[1:] = key[1:] == key[:-1]
        if not bool(dup.any()):
            break
        c = torch.where(dup, torch.randint(0, cols, (nnz,), generator=g, device=device), c)
    else:
        raise RuntimeError("lognormal_csr: could not make the columns of every row distinct")
    c = c.to(torch.int32)
    vals = (torch.rand(nnz, generator=g, device=device) * 2 - 1).to(dtype)
    x = (torch.rand(cols, generator=g, device=device) * 2 - 1).to(dtype)
    return vals, torch.from_numpy(indptr.astype(np.int32)).to(device), c, x
