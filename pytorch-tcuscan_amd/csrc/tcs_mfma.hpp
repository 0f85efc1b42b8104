// Matrix-unit (MFMA) scan of 1024-element units on gfx950.
//
// A unit is 1024 consecutive elements viewed as 32 rows of 32: row n holds
// elements [32n, 32n + 32).  One product C = L * X^T, with L the 32 x 32
// lower-triangular ones matrix (L[m][k] = 1 for k <= m) and X the unit,
// gives C[m][n] = x[32n] + ... + x[32n + m]: the inclusive scan of all 32
// rows in one instruction pair.  X^T is the B operand: lane l (row n =
// l % 32, half h = l / 32) supplies row n's elements in the k slots the
// instruction gives that lane, and L is built in the A operand with the
// same slot assignment, so every x is paired with its own column of L
// whatever order the hardware uses among a lane's k slots.  The result
// comes back with lane l holding column n at rows m = 8q + 4h + s
// (accumulator 4q + s): four runs of four consecutive outputs of row n.
//
// fp16.  v_mfma_f32_32x32x16_f16 accumulates in fp32, which is exact only
// while every partial sum fits in 24 bits.  Each value is split into
// hi = trunc(64 x) / 64 and lo = x - hi (both exact in fp16).  For |x| < 256
// the row sums of the hi parts are multiples of 2^-6 below 2^13 (19 bits)
// and those of the lo parts multiples of 2^-24 below 2^-1 (23 bits), so both
// products are exact and hi + lo in double is the exact row prefix.  A unit
// holding any |x| >= 256, infinity or NaN is scanned sequentially by its
// lanes from LDS instead: also exact, and it keeps inf/NaN semantics, which
// products with the zero entries of L would not (inf * 0 = NaN).
//
// bf16.  v_mfma_f32_32x32x16_bf16 accumulates in fp32 too, and bf16 values
// span fp32's exponent range, so a fixed split cannot work.  Each row n
// takes E = the largest exponent among its 32 values and splits every value
// into three bands of weights: bits of weight >= 2^(E-11), in [2^(E-23),
// 2^(E-12)] and in [2^(E-35), 2^(E-24)].  A band piece of a value is a subset
// of its 8 significant bits (exact in bf16); the row sums of one band are
// integer multiples of its quantum below 2^17 quanta (exact in fp32), and
// the three sums together span fewer than 53 bits (exact in double).  A unit
// holding a value with bits below 2^(E-35) of its row, a row reaching 2^122
// (whose fp32 row sums could overflow), or an infinity or a NaN, takes the
// sequential path.
//
// int8.  v_mfma_i32_32x32x32_i8 covers a whole row (K = 32) in one
// instruction and accumulates exactly in int32; a lane's 16 bytes of a unit
// are one lane group (CodecI8U).
//
// This replaces the reference's Cube-unit row scan C = A * U
// (src/kernels/kernel_row_scan.h:97-131) in transposed form.
#pragma once

#include "tcs_common.hpp"

namespace tcs {

typedef _Float16 h8v __attribute__((ext_vector_type(8)));
typedef __bf16 b8v __attribute__((ext_vector_type(8)));
typedef _Float16 h2v __attribute__((ext_vector_type(2)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef int i32x16 __attribute__((ext_vector_type(16)));

constexpr int kUnit = 1024;

// Output position, within its unit, of the first of the four consecutive
// outputs lane `lane` holds in accumulators 4q .. 4q + 3.
__device__ __forceinline__ int unit_out_pos(int q, int lane) {
  return 32 * (lane & 31) + 8 * q + 4 * (lane >> 5);
}

// Specialised for the codecs the matrix-unit scan covers (fp16, bf16, int8).
template <class C>
struct UnitScan {
  static constexpr int kStageHalfs = 1;
  static constexpr int kGroups = 2;  // lane groups per unit
};

template <>
struct UnitScan<CodecF16> {
  static constexpr int kStageHalfs = kUnit;  // LDS staging per wave (sequential path)
  static constexpr int kGroups = 2;

  // Element offset, within a unit, of the 8 elements lane `lane` loads as
  // piece p (p = 0: k slots 0-15 of row n, p = 1: slots 16-31).
  __device__ static int piece(int p, int lane) { return 32 * (lane & 31) + 16 * p + 8 * (lane >> 5); }

  __device__ static void split(const Packed<uint16_t, 8>& p, h8v& hi, h8v& lo) {
    const h2v k64 = {(_Float16)64.0f, (_Float16)64.0f};
    const h2v kInv = {(_Float16)0.015625f, (_Float16)0.015625f};
#pragma unroll
    for (int r = 0; r < 8; r += 2) {
      const h2v x = __builtin_bit_cast(h2v, p.w[r >> 1]);
      const h2v t = __builtin_elementwise_trunc(x * k64) * kInv;
      const h2v d = x - t;
      hi[r] = t[0];
      hi[r + 1] = t[1];
      lo[r] = d[0];
      lo[r + 1] = d[1];
    }
  }

  // v[i] = inclusive prefix, within its row, of the output the lane holds
  // in accumulator i.
  __device__ static void scan(const Packed<uint16_t, 8>& p0, const Packed<uint16_t, 8>& p1, double (&v)[16],
                              uint16_t* __restrict__ stage, int lane) {
    const int n = lane & 31, h = lane >> 5;
    uint32_t big = 0;
#pragma unroll
    for (int r = 0; r < 8; ++r)
      big |= (uint32_t)((p0[r] & 0x7fffu) >= 0x5c00u) | (uint32_t)((p1[r] & 0x7fffu) >= 0x5c00u);
    if (__ballot(big) == 0) {
      h8v a0, a1, hi0, lo0, hi1, lo1;
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        a0[r] = (8 * h + r <= n) ? (_Float16)1.0f : (_Float16)0.0f;
        a1[r] = (16 + 8 * h + r <= n) ? (_Float16)1.0f : (_Float16)0.0f;
      }
      split(p0, hi0, lo0);
      split(p1, hi1, lo1);
      f32x16 ch = __builtin_amdgcn_mfma_f32_32x32x16_f16(a0, hi0, f32x16{}, 0, 0, 0);
      ch = __builtin_amdgcn_mfma_f32_32x32x16_f16(a1, hi1, ch, 0, 0, 0);
      f32x16 cl = __builtin_amdgcn_mfma_f32_32x32x16_f16(a0, lo0, f32x16{}, 0, 0, 0);
      cl = __builtin_amdgcn_mfma_f32_32x32x16_f16(a1, lo1, cl, 0, 0, 0);
#pragma unroll
      for (int i = 0; i < 16; ++i) v[i] = (double)ch[i] + (double)cl[i];
    } else {
      // Sequential path: stage the unit in element order, then lane l
      // walks row n and keeps the prefixes of the rows m it owns.
      u32x4 w0, w1;
      __builtin_memcpy(&w0, p0.w, 16);
      __builtin_memcpy(&w1, p1.w, 16);
      *reinterpret_cast<u32x4*>(stage + piece(0, lane)) = w0;
      *reinterpret_cast<u32x4*>(stage + piece(1, lane)) = w1;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      double run = 0.0;
#pragma unroll
      for (int m = 0; m < 32; ++m) {
        run += to_acc_f(stage[32 * n + m], F16{});
        const int i = 4 * (m >> 3) + (m & 3);
        if (((m >> 2) & 1) == h) v[i] = run;
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
  }
};

template <>
struct UnitScan<CodecBF16> {
  static constexpr int kStageHalfs = kUnit;  // LDS staging per wave (sequential path)
  static constexpr int kGroups = 2;

  __device__ static int piece(int p, int lane) { return UnitScan<CodecF16>::piece(p, lane); }

  // 2^e as a float, e in [-126, 127].
  __device__ static float pow2(int e) { return __uint_as_float((uint32_t)(e + 127) << 23); }

  // The three band pieces of the 8 values of one lane group (as bf16 bit
  // patterns) and whether any value keeps bits below the lowest band.
  __device__ static bool split(const Packed<uint16_t, 8>& p, int E, b8v& a, b8v& b, b8v& c) {
    const float qa = pow2(E - 11), qb = pow2(E - 23), qc = pow2(E - 35);
    const float ia = pow2(11 - E), ib = pow2(23 - E), ic = pow2(35 - E);
    bool rest = false;
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const float x = __uint_as_float((uint32_t)p[r] << 16);
      const float pa = __builtin_truncf(x * ia) * qa;
      float t = x - pa;
      const float pb = __builtin_truncf(t * ib) * qb;
      t = t - pb;
      const float pc = __builtin_truncf(t * ic) * qc;
      rest |= (t - pc) != 0.0f;
      a[r] = __builtin_bit_cast(__bf16, (uint16_t)(__float_as_uint(pa) >> 16));
      b[r] = __builtin_bit_cast(__bf16, (uint16_t)(__float_as_uint(pb) >> 16));
      c[r] = __builtin_bit_cast(__bf16, (uint16_t)(__float_as_uint(pc) >> 16));
    }
    return rest;
  }

  __device__ static void scan(const Packed<uint16_t, 8>& p0, const Packed<uint16_t, 8>& p1, double (&v)[16],
                              uint16_t* __restrict__ stage, int lane) {
    const int n = lane & 31, h = lane >> 5;
    // Largest exponent field of row n (this lane's 16 values and its
    // partner lane's); an exponent field of 255 is an infinity or a NaN.
    int ef = 0;
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      ef = max(ef, (int)((p0[r] >> 7) & 0xffu));
      ef = max(ef, (int)((p1[r] >> 7) & 0xffu));
    }
    ef = max(ef, __shfl_xor(ef, 32));
    // Rows of tiny values (largest below 2^-91) use quanta of E = -91; their
    // bits below 2^-126 then send the unit to the sequential path.
    const int E = (ef < 36 ? 36 : ef) - 127;
    b8v pa0, pb0, pc0, pa1, pb1, pc1;
    // A row whose values reach 2^122 could overflow fp32 in a row sum (32
    // terms below 2^(E+1)); it takes the sequential path, as do rows with an
    // infinity or a NaN (exponent field 255).
    bool seq = ef > 248;
    seq |= split(p0, E, pa0, pb0, pc0);
    seq |= split(p1, E, pa1, pb1, pc1);
    if (__ballot(seq) == 0) {
      b8v a0, a1;
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        a0[r] = (8 * h + r <= n) ? (__bf16)1.0f : (__bf16)0.0f;
        a1[r] = (16 + 8 * h + r <= n) ? (__bf16)1.0f : (__bf16)0.0f;
      }
      f32x16 ca = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, pa0, f32x16{}, 0, 0, 0);
      ca = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, pa1, ca, 0, 0, 0);
      f32x16 cb = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, pb0, f32x16{}, 0, 0, 0);
      cb = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, pb1, cb, 0, 0, 0);
      f32x16 cc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, pc0, f32x16{}, 0, 0, 0);
      cc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, pc1, cc, 0, 0, 0);
#pragma unroll
      for (int i = 0; i < 16; ++i) v[i] = (double)ca[i] + (double)cb[i] + (double)cc[i];
    } else {
      u32x4 w0, w1;
      __builtin_memcpy(&w0, p0.w, 16);
      __builtin_memcpy(&w1, p1.w, 16);
      *reinterpret_cast<u32x4*>(stage + piece(0, lane)) = w0;
      *reinterpret_cast<u32x4*>(stage + piece(1, lane)) = w1;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      double run = 0.0;
#pragma unroll
      for (int m = 0; m < 32; ++m) {
        run += to_acc_f(stage[32 * n + m], BF16{});
        const int i = 4 * (m >> 3) + (m & 3);
        if (((m >> 2) & 1) == h) v[i] = run;
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
  }
};

// The int8 product: A bytes L[n][k] for the lane's k slots 16h .. 16h + 15.
__device__ __forceinline__ i32x4 lower_ones_i8(int lane) {
  const int n = lane & 31, h = lane >> 5;
  int ones = n - 16 * h + 1;
  ones = ones < 0 ? 0 : (ones > 16 ? 16 : ones);
  i32x4 a;
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    int c = ones - 4 * w;
    c = c < 0 ? 0 : (c > 4 ? 4 : c);
    a[w] = (int)(0x01010101u & (uint32_t)((1ull << (8 * c)) - 1ull));
  }
  return a;
}

template <>
struct UnitScan<CodecI8U> {
  static constexpr int kStageHalfs = 1;
  static constexpr int kGroups = 1;

  // Lane l holds the 16 bytes of row n at k = 16h .. 16h + 15.
  __device__ static int piece(int, int lane) { return 32 * (lane & 31) + 16 * (lane >> 5); }

  __device__ static void scan(const Packed<int8_t, 16>& p, int (&v)[16], uint16_t* __restrict__, int lane) {
    i32x4 b;
    __builtin_memcpy(&b, p.w, 16);
    const i32x16 c = __builtin_amdgcn_mfma_i32_32x32x32_i8(lower_ones_i8(lane), b, i32x16{}, 0, 0, 0);
#pragma unroll
    for (int i = 0; i < 16; ++i) v[i] = c[i];
  }
};

}  // namespace tcs
