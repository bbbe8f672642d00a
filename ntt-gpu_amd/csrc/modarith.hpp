// Device-side modular arithmetic for 64-bit RNS towers on gfx950.
//
// CDNA4 has no 64x64 multiplier: every 64-bit product below is lowered by
// the compiler to v_mad_u64_u32 / v_mul_lo_u32 / v_mul_hi_u32 chains, which
// are the scarce resource of this path. The forms here minimise those:
//   * Shoup multiply by a fixed constant w with w' = floor(w * 2^64 / q):
//     one 64x64 high product and two low products, result in [0, 2q).
//   * Montgomery reduction of a 128-bit accumulator S < q * 2^64 against
//     constants stored pre-multiplied by 2^64 (so the result is the plain
//     residue): one low and one high product.
// Ranges are documented per function; all results leaving a kernel are
// canonical (in [0, q)).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nttamd {

struct U128 {
  uint64_t lo, hi;
};

__device__ __forceinline__ uint64_t mulhi64(uint64_t a, uint64_t b) { return __umul64hi(a, b); }

__device__ __forceinline__ U128 mul_wide(uint64_t a, uint64_t b) { return {a * b, __umul64hi(a, b)}; }

__device__ __forceinline__ void add_wide(U128 &acc, U128 v) {
  const uint64_t lo = acc.lo + v.lo;
  acc.hi += v.hi + (lo < acc.lo ? 1u : 0u);
  acc.lo = lo;
}

// x * w mod q in [0, 2q) for any x < 2^64, q < 2^63, wp = floor(w * 2^64 / q).
__device__ __forceinline__ uint64_t shoup_lazy(uint64_t x, uint64_t w, uint64_t wp, uint64_t q) {
  return x * w - mulhi64(x, wp) * q;
}

__device__ __forceinline__ uint64_t reduce_once(uint64_t x, uint64_t q) { return x >= q ? x - q : x; }

// Canonical x * w mod q.
__device__ __forceinline__ uint64_t shoup(uint64_t x, uint64_t w, uint64_t wp, uint64_t q) {
  return reduce_once(shoup_lazy(x, w, wp, q), q);
}

__device__ __forceinline__ uint64_t add_mod(uint64_t a, uint64_t b, uint64_t q) {
  const uint64_t s = a + b;  // a, b < q < 2^63: no overflow
  return s >= q ? s - q : s;
}

__device__ __forceinline__ uint64_t sub_mod(uint64_t a, uint64_t b, uint64_t q) {
  return a >= b ? a - b : a + (q - b);
}

// Montgomery reduction: S = hi:lo with S < q * 2^64 -> S * 2^-64 mod q, canonical.
// qinv_neg = -q^-1 mod 2^64.
__device__ __forceinline__ uint64_t mont_reduce(U128 s, uint64_t q, uint64_t qinv_neg) {
  const uint64_t m = s.lo * qinv_neg;
  const uint64_t t = s.hi + mulhi64(m, q) + (s.lo != 0 ? 1u : 0u);
  return t >= q ? t - q : t;
}

// Montgomery product of canonical a and b_mont = b * 2^64 mod q -> a * b mod q.
__device__ __forceinline__ uint64_t mont_mul(uint64_t a, uint64_t b_mont, uint64_t q, uint64_t qinv_neg) {
  return mont_reduce(mul_wide(a, b_mont), q, qinv_neg);
}

}  // namespace nttamd
