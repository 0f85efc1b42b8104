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

// Specialised for the codecs the matrix-unit scan covers (fp16, int8).
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
