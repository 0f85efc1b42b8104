"""Writes tests/golden/small_cases.json: known-answer cases for the scan
family, segmented sums, SpMV, gather, compaction and the consumers of scan
(filters, count_if, split, sort, top-k, searchsorted, segmented-scan revert).

The reference ships no golden vectors (its tests draw seeded random data on
the device), so these fixtures are built here from first principles:
inputs are small seeded values, and every expected output is computed with
exact rational arithmetic (fractions.Fraction) in plain Python loops that
follow the operators' definitions in the reference's tests
(tests/test_scan_multi_core.py:45, tests/test_seg_scan_single_core.py:40-63,
tests/test_seg_sum_multi_core.py:96, tests/test_seg_sum_single_cube.py:68-84,
tests/test_segmented_sum.py:70-75, tests/test_csr_gather.py:77,
tests/test_compress.py:44, tests/test_reduce_tiles.py:37-47,
tests/test_complete_rows.py:27-55, tests/test_complete_blocks.py:27-49),
then rounded once (round-to-nearest-even) to the output type.  They are
independent of oracle/oracle.py, which they pin.

Usage: python tests/golden/make_fixtures.py  (rewrites small_cases.json)
"""
import json
import os
import random
import struct
from fractions import Fraction

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def f16_bits_to_frac(b: int) -> Fraction:
    return Fraction(*float(np.array([b], dtype=np.uint16).view(np.float16)[0]).as_integer_ratio())


def round_frac(fr: Fraction, mant_bits: int, emin: int) -> Fraction:
    """Round-half-even of fr to a binary float with `mant_bits` significand
    bits and minimum normal exponent emin (subnormals included)."""
    if fr == 0:
        return Fraction(0)
    sign = -1 if fr < 0 else 1
    a = abs(fr)
    e = a.numerator.bit_length() - a.denominator.bit_length()
    if Fraction(2) ** e > a:
        e -= 1
    e = max(e, emin)
    ulp = Fraction(2) ** (e - (mant_bits - 1))
    q, r = divmod(a, ulp)
    q = int(q)
    if r * 2 > ulp or (r * 2 == ulp and q % 2 == 1):
        q += 1
    return sign * q * ulp


def f32_bits(fr: Fraction) -> int:
    v = round_frac(fr, 24, -126)
    return struct.unpack("<I", struct.pack("<f", float(v)))[0]


def f16_bits(fr: Fraction) -> int:
    v = round_frac(fr, 11, -14)
    return int(np.array([float(v)], dtype=np.float16).view(np.uint16)[0])


def rand_f16(rng, n, lo=-2.0, hi=2.0):
    vals = np.array([rng.uniform(lo, hi) for _ in range(n)], dtype=np.float16)
    return [int(b) for b in vals.view(np.uint16)]


def wrap32(v: int) -> int:
    return (v + 2**31) % 2**32 - 2**31


def scan_frac(xs, start=Fraction(0)):
    out, run = [], start
    for v in xs:
        run += v
        out.append(run)
    return out


def main():
    rng = random.Random(20260101)
    cases = {}

    # Plain scans (run_scan_multi_core / run_scan_single_core semantics).
    x = rand_f16(rng, 1000)
    fx = [f16_bits_to_frac(b) for b in x]
    cases["scan_f16"] = {"x_f16": x, "y_f32": [f32_bits(v) for v in scan_frac(fx)]}
    cases["scan_f16_start"] = {"x_f16": x, "start": 3.25,
                               "y_f32": [f32_bits(v) for v in scan_frac(fx, Fraction(13, 4))]}
    xi = [rng.randint(-128, 127) for _ in range(777)]
    run, yi = 0, []
    for v in xi:
        run += v
        yi.append(wrap32(run))
    cases["scan_i8"] = {"x_i8": xi, "y_i32": yi}

    # Flag-segmented scan: f = 1 starts a segment.
    x = rand_f16(rng, 600)
    f = [1 if rng.random() < 0.05 else 0 for _ in range(600)]
    f[0] = 0
    run, y = Fraction(0), []
    for b, fl in zip(x, f):
        run = f16_bits_to_frac(b) if fl else run + f16_bits_to_frac(b)
        y.append(f32_bits(run))
    cases["seg_scan_f16"] = {"x_f16": x, "f": f, "y_f32": y}

    # Row scans (run_scan_batch / run_row_scan): rows of 37, ragged last row.
    x = rand_f16(rng, 300)
    y = []
    for r0 in range(0, 300, 37):
        y += [f32_bits(v) for v in scan_frac([f16_bits_to_frac(b) for b in x[r0:r0 + 37]])]
    cases["scan_rows_f16"] = {"x_f16": x, "row_len": 37, "y_f32": y}

    # Segmented sums over a full CSR indptr, with empty rows (leading,
    # inner, trailing).
    lens = [0, 3, 0, 0, 7, 1, 0, 12, 5, 0, 2, 9, 0]
    indptr = [0]
    for L in lens:
        indptr.append(indptr[-1] + L)
    nnz = indptr[-1]
    xi = [rng.randint(-128, 127) for _ in range(nnz)]
    cases["seg_sum_i8"] = {"x_i8": xi, "indptr": indptr,
                           "y_i32": [sum(xi[indptr[r]:indptr[r + 1]]) for r in range(len(lens))]}
    x = rand_f16(rng, nnz)
    fx = [f16_bits_to_frac(b) for b in x]
    cases["seg_sum_f16"] = {"x_f16": x, "indptr": indptr,
                            "y_f32": [f32_bits(sum(fx[indptr[r]:indptr[r + 1]], Fraction(0)))
                                      for r in range(len(lens))]}
    # Starts-only convention of run_seg_sum_single_cube: indptr[:-1], last
    # segment ends at len(x).
    starts = indptr[:-1]
    cases["seg_sum_single_cube_f16"] = {
        "x_f16": x, "starts": starts,
        "y_f32": [f32_bits(sum(fx[starts[r]:(starts[r + 1] if r + 1 < len(starts) else nnz)], Fraction(0)))
                  for r in range(len(starts))]}
    # Flag segmented sum (run_seg_sum): f = 1 marks the last element of a
    # segment; output one sum per flag.
    x = rand_f16(rng, 200)
    f = [1 if rng.random() < 0.1 else 0 for _ in range(200)]
    f[-1] = 1
    out, run = [], Fraction(0)
    for b, fl in zip(x, f):
        run += f16_bits_to_frac(b)
        if fl:
            out.append(f32_bits(run))
            run = Fraction(0)
    cases["seg_sum_flags_f16"] = {"x_f16": x, "f": f, "y_f32": out}

    # CSR SpMV: 40 x 50 fp16, per-row distinct sorted columns, some empty rows.
    rows, cols = 40, 50
    indptr, colv = [0], []
    for r in range(rows):
        k = 0 if r % 7 == 3 else rng.randint(1, 12)
        colv += sorted(rng.sample(range(cols), k))
        indptr.append(len(colv))
    vals = rand_f16(rng, len(colv))
    xv = rand_f16(rng, cols)
    y = []
    for r in range(rows):
        acc = Fraction(0)
        for p in range(indptr[r], indptr[r + 1]):
            acc += f16_bits_to_frac(vals[p]) * f16_bits_to_frac(xv[colv[p]])
        y.append(f32_bits(acc))
    cases["spmv_f16"] = {"vals_f16": vals, "indptr": indptr, "cols": colv, "x_f16": xv, "y_f32": y}

    # CSR gather: z = vals * x[cols] rounded to fp16; int16 wraps.
    z = [f16_bits(f16_bits_to_frac(vals[p]) * f16_bits_to_frac(xv[colv[p]])) for p in range(len(colv))]
    cases["csr_gather_f16"] = {"vals_f16": vals, "cols": colv, "x_f16": xv, "z_f16": z}
    vi = [rng.randint(-300, 300) for _ in range(len(colv))]
    xi = [rng.randint(-300, 300) for _ in range(cols)]
    zi = [((vi[p] * xi[colv[p]] + 2**15) % 2**16) - 2**15 for p in range(len(colv))]
    cases["csr_gather_i16"] = {"vals_i16": vi, "cols": colv, "x_i16": xi, "z_i16": zi}

    # Compaction: keep mask == 1 exactly (other nonzero values are dropped).
    n = 500
    xb = [rng.getrandbits(32) for _ in range(n)]
    m = [rng.choice([0, 0, 1, 1, 1, 2, -1]) for _ in range(n)]
    keep = [i for i in range(n) if m[i] == 1]
    cases["compress_u32"] = {"x_u32": xb, "mask": m, "z_u32": [xb[i] for i in keep], "pos": keep}

    # reduce_tiles: blocks of ceil(ceil(n/T)/B)*T elements.
    x = rand_f16(rng, 1000)
    fx = [f16_bits_to_frac(b) for b in x]
    T, B = 64, 7
    per = -(-(-(-1000 // T)) // B) * T
    cases["reduce_tiles_f16"] = {"x_f16": x, "tile_len": T, "num_blocks": B,
                                 "y_f32": [f32_bits(sum(fx[b * per:min(1000, (b + 1) * per)], Fraction(0)))
                                           if b * per < 1000 else 0 for b in range(B)]}

    # complete_blocks: x + sum(sums[:b]) over len(x)/len(sums) blocks (float32
    # data, exact dyadic values).
    xs = [rng.randint(-1000, 1000) / 8 for _ in range(480)]
    sums = [rng.randint(-1000, 1000) / 4 for _ in range(6)]
    bl = 480 // 6
    cases["complete_blocks_f32"] = {
        "x_f32": xs, "sums_f32": sums, "tile_len": 16,
        "y_f32": [f32_bits(Fraction(xs[i]) + sum((Fraction(s) for s in sums[:i // bl]), Fraction(0)))
                  for i in range(480)]}

    # complete_rows: block b (tiles of w * min(h, ceil(8192/w))) starts from
    # sum(sums[:b]); inside a block, each row of w elements adds the running
    # carry, which becomes the row's last output.
    w, h, nb = 16, 4, 5
    tile = w * min(h, -(-8192 // w))
    n = tile * nb
    xs = [rng.randint(-1000, 1000) / 8 for _ in range(n)]
    sums = [rng.randint(-1000, 1000) / 4 for _ in range(nb)]
    y = []
    for b in range(nb):
        carry = sum((Fraction(s) for s in sums[:b]), Fraction(0))
        for r0 in range(b * tile, (b + 1) * tile, w):
            row = [Fraction(v) + carry for v in xs[r0:r0 + w]]
            y += row
            carry = row[-1]
    cases["complete_rows_f32"] = {"x_f32": xs, "sums_f32": sums, "tile_width": w, "tile_height": h,
                                  "y_f32": [f32_bits(v) for v in y]}

    # --- Consumers of scan: filters, count_if, split, sort, top-k, search,
    # the segmented-scan revert.  Values compared as exact fractions, NaN
    # first tested separately; ties broken by input position.
    def f16_is_nan(b):
        return (b & 0x7FFF) > 0x7C00

    x = rand_f16(rng, 300)
    p16 = int(np.array([np.float32(0.1)]).astype(np.float16).view(np.uint16)[0])  # pivot 0.1 as half
    for k in range(0, 300, 23):
        x[k] = p16  # ties at the pivot
    x[5], x[17], x[29], x[41] = 0x7E00, 0x7C00, 0xFC00, 0x8000  # NaN, +inf, -inf, -0
    fp = f16_bits_to_frac(p16)

    def fv(b):  # exact value (inf as +-2^20, beyond every finite half)
        if (b & 0x7FFF) == 0x7C00:
            return Fraction(2 ** 20) * (-1 if b & 0x8000 else 1)
        return f16_bits_to_frac(b)
    ge = [b for b in x if not f16_is_nan(b) and fv(b) >= fp]
    le = [b for b in x if not f16_is_nan(b) and fv(b) <= fp]
    pos = [i for i, b in enumerate(x) if not f16_is_nan(b) and fv(b) >= fp]
    cmps = [lambda a: a < fp, lambda a: a > fp, lambda a: a == fp, lambda a: a <= fp, lambda a: a >= fp,
            lambda a: a != fp]
    counts = [sum(1 for b in x if (f16_is_nan(b) and m == 5) or (not f16_is_nan(b) and cmps[m](fv(b))))
              for m in range(6)]
    cases["filters_f16"] = {"x_f16": x, "pivot": 0.1, "ge_f16": ge, "le_f16": le, "where": pos,
                            "count_if": counts}

    xi = [rng.randint(-300, 300) for _ in range(257)]
    m = [rng.choice([0, 1, 1, 2, -1]) for _ in range(257)]
    order = [i for i in range(257) if m[i] == 1] + [i for i in range(257) if m[i] != 1]
    cases["split_i16"] = {"x_i16": xi, "mask": m, "z_i16": [xi[i] for i in order],
                          "idx_out": [i + 1 for i in order]}  # indices_in = 1..n

    def sort_key(b, i):  # NaN last, -0 == +0, ties by position
        return (1, Fraction(0), i) if f16_is_nan(b) else (0, fv(b), i)
    asc = sorted(range(300), key=lambda i: sort_key(x[i], i))
    desc = sorted(range(300), key=lambda i: (0 if f16_is_nan(x[i]) else 1,
                                              Fraction(0) if f16_is_nan(x[i]) else -fv(x[i]), i))
    cases["sort_f16"] = {"x_f16": x, "asc_idx": asc, "desc_idx": desc}
    ys = [rng.randint(-20, 20) for _ in range(400)]
    cases["sort_i16"] = {"x_i16": ys, "asc_idx": sorted(range(400), key=lambda i: (ys[i], i)),
                         "desc_idx": sorted(range(400), key=lambda i: (-ys[i], i))}
    cases["topk_f16"] = {"x_f16": x, "k": 7, "largest_idx": desc[:7], "smallest_idx": asc[:7]}

    hay = sorted(rng.randint(-50, 50) for _ in range(200))
    needles = [rng.randint(-60, 60) for _ in range(100)] + [-2 ** 31, 2 ** 31 - 1]
    cases["searchsorted_i32"] = {"sorted": hay, "values": needles,
                                 "out": [sum(1 for h in hay if h < v) for v in needles]}

    xr = [rng.randint(0, 9) for _ in range(400)]
    fr = [0] + [1 if rng.random() < 0.05 else 0 for _ in range(399)]
    sx = [sum(xr[:i + 1]) for i in range(400)]
    sf = [sum(fr[:i + 1]) for i in range(400)]
    ends = [sx[i] for i in range(400) if i == 399 or fr[i + 1] == 1]
    seg = []
    for i in range(400):
        if i == 0 or fr[i]:
            run = 0
        run += xr[i]
        seg.append(run)
    cases["seg_scan_revert_f32"] = {"scan_x": sx, "scan_f": sf, "seg_ends": ends, "z": seg}

    with open(os.path.join(HERE, "small_cases.json"), "w") as fh:
        json.dump(cases, fh, separators=(",", ":"))
    print("wrote", len(cases), "cases")


if __name__ == "__main__":
    main()
