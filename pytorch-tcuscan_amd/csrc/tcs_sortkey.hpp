// Order-preserving unsigned keys of the element types that sort and top-k
// accept: key(a) < key(b) exactly when a sorts before b ascending, with
// -0.0 equal to +0.0 and every NaN after +inf (torch.sort's order).
#pragma once

#include "tcs_common.hpp"

namespace tcs {

template <int DT> struct SortKey;
template <> struct SortKey<TCS_F16> {
  using Raw = uint16_t;
  static constexpr int kBits = 16;
  __device__ static uint32_t key(uint16_t r) {
    if ((r & 0x7fffu) > 0x7c00u) return 0xffffu;  // NaN after +inf
    if (r == 0x8000u) r = 0;                      // -0.0 ties with +0.0
    return (r & 0x8000u) ? (uint16_t)~r : (uint32_t)(r | 0x8000u);
  }
};
template <> struct SortKey<TCS_BF16> {
  using Raw = uint16_t;
  static constexpr int kBits = 16;
  __device__ static uint32_t key(uint16_t r) {
    if ((r & 0x7fffu) > 0x7f80u) return 0xffffu;
    if (r == 0x8000u) r = 0;
    return (r & 0x8000u) ? (uint16_t)~r : (uint32_t)(r | 0x8000u);
  }
};
template <> struct SortKey<TCS_F32> {
  using Raw = uint32_t;
  static constexpr int kBits = 32;
  __device__ static uint32_t key(uint32_t r) {
    if ((r & 0x7fffffffu) > 0x7f800000u) return 0xffffffffu;
    if (r == 0x80000000u) r = 0;
    return (r & 0x80000000u) ? ~r : (r | 0x80000000u);
  }
};
template <> struct SortKey<TCS_I8> {
  using Raw = uint8_t;
  static constexpr int kBits = 8;
  __device__ static uint32_t key(uint8_t r) { return (uint32_t)(r ^ 0x80u); }
};
template <> struct SortKey<TCS_I16> {
  using Raw = uint16_t;
  static constexpr int kBits = 16;
  __device__ static uint32_t key(uint16_t r) { return (uint32_t)(r ^ 0x8000u); }
};
template <> struct SortKey<TCS_I32> {
  using Raw = uint32_t;
  static constexpr int kBits = 32;
  __device__ static uint32_t key(uint32_t r) { return r ^ 0x80000000u; }
};
template <> struct SortKey<TCS_U32> {
  using Raw = uint32_t;
  static constexpr int kBits = 32;
  __device__ static uint32_t key(uint32_t r) { return r; }
};

}  // namespace tcs
