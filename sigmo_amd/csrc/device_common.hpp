// Device helpers shared by the plain-form and batched-form kernels (wave64, gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace sigmo_k {

constexpr int kWave = 64;

__device__ __forceinline__ int lane_id() { return static_cast<int>(threadIdx.x & (kWave - 1)); }

__device__ __forceinline__ uint32_t rl32(uint32_t v, uint32_t lane) {
  return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), static_cast<int>(lane)));
}
__device__ __forceinline__ uint64_t rl64(uint64_t v, uint32_t lane) {
  return static_cast<uint64_t>(rl32(static_cast<uint32_t>(v), lane)) | (static_cast<uint64_t>(rl32(static_cast<uint32_t>(v >> 32), lane)) << 32);
}
__device__ __forceinline__ uint32_t uniform32(uint32_t v) {
  return static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(v)));
}
__device__ __forceinline__ uint64_t uniform64(uint64_t v) {
  return static_cast<uint64_t>(uniform32(static_cast<uint32_t>(v))) | (static_cast<uint64_t>(uniform32(static_cast<uint32_t>(v >> 32))) << 32);
}
// Bits at positions >= pos of a 64-bit word (0 when pos >= 64).
__device__ __forceinline__ uint64_t bits_from(uint32_t pos) { return pos >= 64 ? 0ull : (~0ull << pos); }

// bit b of c spread over 32 bits (0 or ~0): one signed bit-field extract
__device__ __forceinline__ uint32_t bit_mask(uint32_t c, uint32_t b) {
  return static_cast<uint32_t>(static_cast<int32_t>(c << (31u - b)) >> 31);
}

// The hardware's three-input bit operation, v_bitop3_b32, weighs its operands 0xF0, 0xCC and
// 0xAA in order; the truth tables below follow from that.
// Majority of a, b, c on 32 positions (table 0xE8, the same for any operand order).
__device__ __forceinline__ uint32_t maj3(uint32_t a, uint32_t b, uint32_t c) { return __builtin_amdgcn_bitop3_b32(a, b, c, 0xE8); }
// m & (l ^ k) on 32 positions (table 0x60): with k all ones where a label bit is 0, the
// positions whose label plane matches that bit.
__device__ __forceinline__ uint32_t and_xor(uint32_t m, uint32_t l, uint32_t k) { return __builtin_amdgcn_bitop3_b32(m, l, k, 0x60); }

// x >= c for 64 four-bit values held as bit planes x[0..3] and a wave-uniform c in 1..15:
// x >= c exactly when the 4-bit sum x + (16 - c) carries out.  The carry planes are built from
// the lowest bit up, carry' = maj(x_b, carry, k_b) with k_b a uniform all-ones or all-zero mask:
// one maj3 per plane and 32 positions (the compiler's form of (x & carry) | (k & (x | carry))
// took an AND, an XOR and a bit-select).
__device__ __forceinline__ uint64_t ge4(const uint64_t x[4], uint32_t c) {
  const uint32_t k = 16u - c;
  const uint32_t k0 = bit_mask(k, 0);
  uint32_t lo = static_cast<uint32_t>(x[0]) & k0, hi = static_cast<uint32_t>(x[0] >> 32) & k0;
#pragma unroll
  for (int b = 1; b < 4; ++b) {
    const uint32_t km = bit_mask(k, static_cast<uint32_t>(b));
    lo = maj3(static_cast<uint32_t>(x[b]), lo, km);
    hi = maj3(static_cast<uint32_t>(x[b] >> 32), hi, km);
  }
  return lo | (static_cast<uint64_t>(hi) << 32);
}

// The positions of a 64-bit word whose 8-bit label (bit planes L[0..7]) equals a wave-uniform
// ql, within `valid`: one and_xor per plane and 32 positions.
__device__ __forceinline__ uint64_t label_match(const uint64_t L[8], uint32_t ql, uint64_t valid) {
  uint32_t lo = static_cast<uint32_t>(valid), hi = static_cast<uint32_t>(valid >> 32);
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    const uint32_t k = bit_mask(~ql, static_cast<uint32_t>(b));
    lo = and_xor(lo, static_cast<uint32_t>(L[b]), k);
    hi = and_xor(hi, static_cast<uint32_t>(L[b] >> 32), k);
  }
  return lo | (static_cast<uint64_t>(hi) << 32);
}

// 64 x 64 bit-matrix transpose across the wavefront: lane p holds row p on entry, lane i
// holds column i (bit p = bit i of lane p's row) on exit.  Six butterfly stages, each
// exchanging the off-diagonal j x j blocks between lanes p and p ^ j.
__device__ __forceinline__ uint64_t transpose64(uint64_t x) {
  const uint32_t lane = static_cast<uint32_t>(lane_id());
  constexpr uint64_t keep[6] = {0x00000000FFFFFFFFull, 0x0000FFFF0000FFFFull, 0x00FF00FF00FF00FFull,
                                0x0F0F0F0F0F0F0F0Full, 0x3333333333333333ull, 0x5555555555555555ull};
#pragma unroll
  for (int t = 0; t < 6; ++t) {
    const uint32_t j = 32u >> t;
    const uint64_t L = keep[t];
    const uint64_t y = static_cast<uint64_t>(__shfl_xor(static_cast<unsigned long long>(x), static_cast<int>(j)));
    x = (lane & j) ? ((x & ~L) | ((y & ~L) >> j)) : ((x & L) | ((y & L) << j));
  }
  return x;
}

// One stage of ge4_bits on 32 positions: m ? (x & ge) : (x | ge) (the majority of x, ge and ~m)
// as one three-input bit operation, v_bitop3_b32.  Its truth table 0xB2 puts the select m in
// the middle operand, whose table weight (0xCC) is the same whichever of the outer operands is
// weighted 0xF0; x and ge enter the function symmetrically, so the table does not depend on
// that order either.
__device__ __forceinline__ uint32_t ge_stage(uint32_t x, uint32_t m, uint32_t ge) { return __builtin_amdgcn_bitop3_b32(x, m, ge, 0xB2); }

// x_0 >= c_0 on 32 positions: x | ~m, written (x & x) | ~m so that the table (0xB3) is
// symmetric in the outer operands as well.
__device__ __forceinline__ uint32_t ge_first(uint32_t x, uint32_t m) { return __builtin_amdgcn_bitop3_b32(x, m, x, 0xB3); }

// x >= c for four-bit values held as bit planes x0..x3 and a per-lane c, from the lowest bit
// up: ge = c_b ? (x_b & ge) : (x_b | ge), starting from x_0 >= c_0.  Each stage is one
// ge_stage per 32 positions (the compiler's form was an AND, an OR, a compare and a select).
template <typename T>
__device__ __forceinline__ T ge4_bits(T x0, T x1, T x2, T x3, uint32_t c) {
  static_assert(sizeof(T) == 4 || sizeof(T) == 8, "32 or 64 positions");
  const uint32_t m0 = bit_mask(c, 0), m1 = bit_mask(c, 1), m2 = bit_mask(c, 2), m3 = bit_mask(c, 3);
  if constexpr (sizeof(T) == 4) {
    uint32_t ge = ge_first(x0, m0);
    ge = ge_stage(x1, m1, ge);
    ge = ge_stage(x2, m2, ge);
    return ge_stage(x3, m3, ge);
  } else {
    uint32_t lo = ge_first(static_cast<uint32_t>(x0), m0), hi = ge_first(static_cast<uint32_t>(x0 >> 32), m0);
    lo = ge_stage(static_cast<uint32_t>(x1), m1, lo);
    hi = ge_stage(static_cast<uint32_t>(x1 >> 32), m1, hi);
    lo = ge_stage(static_cast<uint32_t>(x2), m2, lo);
    hi = ge_stage(static_cast<uint32_t>(x2 >> 32), m2, hi);
    lo = ge_stage(static_cast<uint32_t>(x3), m3, lo);
    hi = ge_stage(static_cast<uint32_t>(x3 >> 32), m3, hi);
    return static_cast<T>(lo) | (static_cast<T>(hi) << 32);
  }
}

}  // namespace sigmo_k
