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

__host__ __device__ __forceinline__ void add_wide(U128 &acc, U128 v) {
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

// a * b for any a < 2^64 and b < 2^63 from four 32x32 products chained
// through v_mad_u64_u32: with b1 = b >> 32 < 2^31 none of the partial sums
// can overflow 64 bits, so no carries are tracked (the generic a * b plus
// __umul64hi recomputes partial products and propagates carries).
//   t = a0 b1 + hi32(a0 b0) < 2^63 + 2^32,  u = a1 b0 + lo32(t) < 2^64,
//   a b = lo32(a0 b0) + 2^32 lo32(u) + 2^64 (a1 b1 + hi32(t) + hi32(u)).
__host__ __device__ __forceinline__ U128 mul_wide63(uint64_t a, uint64_t b) {
  const uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32), b0 = (uint32_t)b, b1 = (uint32_t)(b >> 32);
  const uint64_t p00 = (uint64_t)a0 * b0;
  const uint64_t t = (uint64_t)a0 * b1 + (p00 >> 32);
  const uint64_t u = (uint64_t)a1 * b0 + (uint32_t)t;
  const uint64_t hi = (uint64_t)a1 * b1 + ((t >> 32) + (u >> 32));
  return {(u << 32) | (uint32_t)p00, hi};
}

// (S + m q) / 2^64 with m = lo * qinv_neg mod 2^64, which makes the sum a
// multiple of 2^64 (its low words carry exactly when lo != 0): congruent to
// S * 2^-64 mod q and below S / 2^64 + q, so below 2q for S < q * 2^64.
__host__ __device__ __forceinline__ uint64_t mont_half(U128 s, uint64_t q, uint64_t qinv_neg) {
  const uint64_t m = s.lo * qinv_neg;
  return s.hi + mul_wide63(m, q).hi + (s.lo != 0 ? 1u : 0u);
}

// Montgomery reduction: S = hi:lo with S < q * 2^64 -> S * 2^-64 mod q, canonical.
// qinv_neg = -q^-1 mod 2^64; q < 2^63 (every prime of this library).
__host__ __device__ __forceinline__ uint64_t mont_reduce(U128 s, uint64_t q, uint64_t qinv_neg) {
  const uint64_t t = mont_half(s, q, qinv_neg);
  return t >= q ? t - q : t;
}

// Montgomery product of canonical a and b_mont = b * 2^64 mod q -> a * b mod q.
__device__ __forceinline__ uint64_t mont_mul(uint64_t a, uint64_t b_mont, uint64_t q, uint64_t qinv_neg) {
  return mont_reduce(mul_wide(a, b_mont), q, qinv_neg);
}

// ---------------------------------------------------------------- wide-lazy butterflies (q < 2^61)
// For q < 2^61 the NTT keeps residues in [0, 8q) and multiplies with a Shoup
// quotient that may be short by up to 2: floor(y * w' / 2^64) is estimated as
// y1 p1 + hi32(y1 p0) + hi32(y0 p1), dropping y0 p0 and the carry out of the
// low halves (the dropped part is below 3 * 2^64, so the estimate is low by
// 0, 1 or 2). Then T = y w - qh q lies in [0, 4q) instead of [0, 2q), and the
// butterfly folds its add into the multiply-add chain:
//   x' = x + T,  y' = x - T + 4q,  x first brought below 4q.
// Per butterfly that is 3 v_mad_u64_u32 + 1 add pair for the quotient,
// 2 v_mad_u64_u32 + 4 v_mul_lo_u32 + 2 adds for x', and a few 64-bit adds:
// about 20 VALU instructions against ~31 for the exact form.
// The negated constants come precomputed (PrimeConst on the device): were
// they derived in the kernel, the compiler would fold x + (2^64 - c) back into
// a subtract with a carry chain, which is what the selects below avoid.
struct Lazy8 {
  uint64_t q, q4, nq, nq2, nq4;  // q, 4q, 2^64 - q, 2^64 - 2q, 2^64 - 4q
};

__host__ __device__ __forceinline__ Lazy8 lazy8_of(uint64_t q) { return {q, q << 2, 0 - q, 0 - 2 * q, 0 - 4 * q}; }
__host__ __device__ __forceinline__ Lazy8 lazy8_of(uint64_t q, uint64_t nq, uint64_t nq2, uint64_t nq4) {
  return {q, q << 2, nq, nq2, nq4};
}

// floor(y * p / 2^64) - e with e in {0, 1, 2}, any y, p < 2^64.
__host__ __device__ __forceinline__ uint64_t mulhi_est(uint64_t y, uint64_t p) {
  const uint32_t y0 = (uint32_t)y, y1 = (uint32_t)(y >> 32), p0 = (uint32_t)p, p1 = (uint32_t)(p >> 32);
  const uint64_t a = (uint64_t)y1 * p0, b = (uint64_t)y0 * p1;
  return (uint64_t)y1 * p1 + ((a >> 32) + (b >> 32));
}

// (x + y * w + qh * nq) mod 2^64 from 32-bit limbs (nq = 2^64 - q, so the
// last term is -qh * q): two multiply-adds carry the low words, four low
// products land in the high word.
__host__ __device__ __forceinline__ uint64_t mad_lo(uint64_t x, uint64_t y, uint64_t w, uint64_t qh, uint64_t nq) {
  const uint32_t y0 = (uint32_t)y, y1 = (uint32_t)(y >> 32), w0 = (uint32_t)w, w1 = (uint32_t)(w >> 32);
  const uint32_t h0 = (uint32_t)qh, h1 = (uint32_t)(qh >> 32), n0 = (uint32_t)nq, n1 = (uint32_t)(nq >> 32);
  uint64_t r = (uint64_t)y0 * w0 + x;
  r = (uint64_t)h0 * n0 + r;
  uint32_t hi = (uint32_t)(r >> 32) + y0 * w1 + y1 * w0 + h0 * n1 + h1 * n0;
#if defined(__HIP_DEVICE_COMPILE__)
  // Opaque to the optimizer: otherwise it folds the cross products into a
  // 64-bit add of (sum << 32) onto r (a v_mov of zero and a v_lshl_add_u64)
  // instead of one 32-bit add into the high word. 2% fewer instructions in
  // the forward row pass, forward 59-bit transform 1.8% faster.
  asm("" : "+v"(hi));
#endif
  return ((uint64_t)hi << 32) | (uint32_t)r;
}

// y * w mod q in [0, 4q) for any y < 2^64 (w < q, wp = floor(w 2^64 / q), q < 2^61).
__host__ __device__ __forceinline__ uint64_t shoup4(uint64_t y, uint64_t w, uint64_t wp, const Lazy8 &m) {
  return mad_lo(0, y, w, mulhi_est(y, wp), m.nq);
}

// x >= c ? x - c : x, selected on the borrow of the subtraction (sub, subb,
// two selects) rather than a separate 64-bit compare.
__host__ __device__ __forceinline__ uint64_t sub_if_ge(uint64_t x, uint64_t c) {
  uint64_t d;
  const bool borrow = __builtin_sub_overflow(x, c, &d);
  return borrow ? x : d;
}

// mask ? a : b bitwise: one v_bfi_b32 on the device (written as asm so the
// compiler cannot turn the select back into a compare and condition code).
__host__ __device__ __forceinline__ uint32_t bit_select(uint32_t mask, uint32_t a, uint32_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
  uint32_t r;
  asm("v_bfi_b32 %0, %1, %2, %3" : "=v"(r) : "v"(mask), "v"(a), "v"(b));
  return r;
#else
  return (mask & a) | (~mask & b);
#endif
}

// x >= c ? x - c : x for c <= 2^63 and x < 2c, with neg_c = 2^64 - c: t = x + neg_c
// is x - c (top bit clear, x >= c) or wraps above 2^63 (x < c), so the sign of
// t picks x or t. One 64-bit add, one shift, two bitfield selects; no carry
// chain or condition code, hence no hazard wait states between them.
__host__ __device__ __forceinline__ uint64_t sub_if_ge_neg(uint64_t x, uint64_t neg_c) {
  const uint64_t t = x + neg_c;
  const uint32_t mask = (uint32_t)((int32_t)(uint32_t)(t >> 32) >> 31);  // all ones iff x < c
  return ((uint64_t)bit_select(mask, (uint32_t)(x >> 32), (uint32_t)(t >> 32)) << 32) |
         bit_select(mask, (uint32_t)x, (uint32_t)t);
}

// Forward (Cooley-Tukey) butterfly: x, y in [0, 8q) -> x + yw, x - yw, both in [0, 8q).
__host__ __device__ __forceinline__ void ct8(uint64_t &x, uint64_t &y, uint64_t w, uint64_t wp, const Lazy8 &m) {
  const uint64_t xr = sub_if_ge_neg(x, m.nq4);  // [0, 4q)
  const uint64_t xn = mad_lo(xr, y, w, mulhi_est(y, wp), m.nq);  // xr + T, T in [0, 4q)
  y = 2 * xr + m.q4 - xn;  // xr - T + 4q in (0, 8q); the wrap of 2xr + 4q cancels mod 2^64
  x = xn;
}

// Inverse (Gentleman-Sande) butterfly: x, y in [0, 4q) -> x + y, (x - y) w, both in [0, 4q).
__host__ __device__ __forceinline__ void gs8(uint64_t &x, uint64_t &y, uint64_t w, uint64_t wp, const Lazy8 &m) {
  const uint64_t s = x + y;  // [0, 8q)
  const uint64_t d = x + m.q4 - y;  // (0, 8q)
  x = sub_if_ge_neg(s, m.nq4);
  y = shoup4(d, w, wp, m);
}

// x < 2^7 q -> [0, 2q), congruent, for q of bit length b >= 24 (every
// residue of a planned wide-lazy forward transform is below 2^7 q): x - e q
// with e = floor(x' r / 2^55), x' = floor(x / 2^(b - 24)) < 2^31 and
// r = floor(2^(b + 31) / q) in [2^31, 2^32) (shift = b - 24, mul = r).
// x' r / 2^55 <= x / q, and it exceeds x / q - 2^-24 - 2^-23 (the truncations
// of x' and r), so e is floor(x / q) or one less. One 64-bit shift, one
// v_mul_hi_u32 and a shift for e, then x + e (2^64 - q): about 6 VALU
// instructions against 11 for a 64-bit quotient estimate or 4 per
// conditional subtract.
__host__ __device__ __forceinline__ uint64_t reduce2_est(uint64_t x, uint32_t shift, uint32_t mul, uint64_t nq) {
  const uint32_t xs = (uint32_t)(x >> shift);
  const uint32_t e = (uint32_t)(((uint64_t)xs * mul) >> 55);
  return x + (uint64_t)e * nq;
}

// [0, 8q) -> [0, q).
__host__ __device__ __forceinline__ uint64_t canon8(uint64_t x, const Lazy8 &m) {
  return sub_if_ge_neg(sub_if_ge_neg(sub_if_ge_neg(x, m.nq4), m.nq2), m.nq);
}

// ---------------------------------------------------------------- forward reduction schedule (N = 2^16)
// A forward (Cooley-Tukey) butterfly only needs its x below 2^64 - 4q: it
// computes x + T and x - T + 4q with T = y w mod q in [0, 4q) for any y, so
// both outputs grow by 4q over the bound of x. With every q < 2^(61 - H) a
// residue may reach 2^(3 + H) q, so a tower can run several stages before x
// must be brought down. make_fwd_plan(H) is the greedy schedule over the 16 stages
// of a 2^16-point transform (pass A: stages 0-7, pass B: 8-15), starting from
// canonical input: stage g reduces x below c = 2^red[g] q (red[g] = -1: no
// reduction) only when skipping it would pass the bound; `bound_log` bounds
// the output by 2^bound_log q. H = 0 keeps one reduction per stage (below
// 4q, outputs below 8q): the schedule of every other launch shape.
// Reductions per transform: 16 for H = 0, 7 for q < 2^60, 3 for q < 2^59,
// 1 for q < 2^58 and none from q < 2^57 down.
struct FwdPlanData {
  int red[16];
  int bound_log;
};
__host__ __device__ constexpr FwdPlanData make_fwd_plan(int H) {
  FwdPlanData p{};
  const int cap = 8 << H;  // residues below cap * q <= 2^64
  int bound = 1;           // canonical input
  for (int g = 0; g < 16; ++g) {
    if (H > 0 && bound + 4 <= cap) {
      p.red[g] = -1;
      bound += 4;
    } else {
      int k = 2;
      while ((2 << k) < bound) ++k;  // x < 2c is what one conditional subtract needs
      p.red[g] = k;
      bound = (1 << k) + 4;
    }
  }
  int l = 0;
  while ((1 << l) < bound) ++l;
  p.bound_log = l;
  return p;
}

// ---------------------------------------------------------------- inverse reduction schedule (N = 2^16)
// An inverse (Gentleman-Sande) butterfly whose inputs are below b q computes
// s = x + y < 2b q and d = x + b q - y in (0, 2b q); its product output
// d w mod q lies in [0, 4q) for any 64-bit d (shoup4), so only the sums grow,
// doubling per stage. With every q < 2^(61 - H) a sum below 2^(3 + H) q still
// fits 64 bits: make_inv_plan(H) lets inputs grow to cap q (cap = 2^(2 + H),
// at most 64) and then brings the stage's sums back below 2q with the
// one-multiply quotient estimate (reduce2_est, valid below 2^7 q), 4, 3 or 2
// estimates over the 16 stages for H = 2, 3 and 4-5 against a conditional
// subtract at every stage. H = 0 (cap 4) keeps one conditional subtract of
// 4q per stage. c_log[g]: stage g's inputs are below 2^c_log[g] q, the
// multiple of q its d adds; red[g]: 0 none, 1 subtract 4q, 2 estimate. The
// last stage (n^-1 folded in) reduces its outputs itself.
struct InvPlanData {
  int red[16];
  int c_log[16];
};
__host__ __device__ constexpr InvPlanData make_inv_plan(int H) {
  InvPlanData p{};
  const int cap = H == 0 ? 4 : (4 << H) < 64 ? (4 << H) : 64;
  int b = 1;  // canonical input
  for (int g = 0; g < 16; ++g) {
    int l = 0;
    while ((1 << l) < b) ++l;
    p.c_log[g] = l;
    const int next = 2 * b > 4 ? 2 * b : 4;
    if (g == 15 || next <= cap) {
      p.red[g] = 0;
      b = next;
    } else {
      p.red[g] = H == 0 ? 1 : 2;
      b = 4;
    }
  }
  return p;
}

// ---------------------------------------------------------------- Barrett
// Variable x variable products a * b mod q (the pointwise product of two
// residue arrays; the reference's candidate form is include/ntt/barrett.hpp:6-49)
// with mu = floor(2^128 / q): one reduction instead of two
// Montgomery steps (the product would otherwise need a conversion into the
// Montgomery domain). With x = ab = xh 2^64 + xl and mu = mh 2^64 + ml,
//   floor(x mu / 2^128) = xh mh + hi(xh ml) + hi(xl mh) + c,
// c the carry out of lo(xh ml) + lo(xl mh) + hi(xl ml), in {0, 1, 2}.
// floor(x mu / 2^128) is floor(x / q) or one less (x < 2^128), so with the
// carry the remainder x - e q is in [0, 2q); without it (e short by up to 2
// more) in [0, 4q), which fits a word only when q < 2^62.
__host__ __device__ __forceinline__ uint64_t mulhi64_hd(uint64_t a, uint64_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __umul64hi(a, b);
#else
  return (uint64_t)(((unsigned __int128)a * b) >> 64);
#endif
}

// q < 2^62: carries dropped, two conditional subtracts.
__host__ __device__ __forceinline__ uint64_t barrett_mul_lazy(uint64_t a, uint64_t b, uint64_t q, uint64_t mu_lo,
                                                              uint64_t mu_hi) {
  const uint64_t xl = a * b, xh = mulhi64_hd(a, b);
  const uint64_t e = xh * mu_hi + mulhi64_hd(xh, mu_lo) + mulhi64_hd(xl, mu_hi);
  const uint64_t r = xl - e * q;  // [0, 4q)
  const uint64_t r2 = r >= 2 * q ? r - 2 * q : r;
  return r2 >= q ? r2 - q : r2;
}

// q < 2^63: every carry kept, one conditional subtract.
__host__ __device__ __forceinline__ uint64_t barrett_mul_exact(uint64_t a, uint64_t b, uint64_t q, uint64_t mu_lo,
                                                               uint64_t mu_hi) {
  const uint64_t xl = a * b, xh = mulhi64_hd(a, b);
  const uint64_t s1 = xh * mu_lo, s2 = xl * mu_hi, s3 = mulhi64_hd(xl, mu_lo);
  const uint64_t t = s1 + s2;
  const uint64_t u = t + s3;
  const uint64_t c = (uint64_t)(t < s1) + (uint64_t)(u < t);
  const uint64_t e = xh * mu_hi + mulhi64_hd(xh, mu_lo) + mulhi64_hd(xl, mu_hi) + c;
  const uint64_t r = xl - e * q;  // [0, 2q)
  return r >= q ? r - q : r;
}

__host__ __device__ __forceinline__ uint64_t barrett_mul(uint64_t a, uint64_t b, uint64_t q, uint64_t mu_lo,
                                                         uint64_t mu_hi) {
  return q < (uint64_t{1} << 62) ? barrett_mul_lazy(a, b, q, mu_lo, mu_hi) : barrett_mul_exact(a, b, q, mu_lo, mu_hi);
}

// ---------------------------------------------------------------- narrow towers
// q < 2^30: residues fit 32-bit registers with the Harvey lazy ranges
// ([0, 4q) forward, [0, 2q) inverse, 4q < 2^32), so a butterfly is one
// 32-bit high product and two 32-bit low products instead of 64-bit chains.
// The Shoup companion is floor(w 2^32 / q) = the high word of the 64-bit
// companion floor(w 2^64 / q) (a floor of a floor by 2^32), so the 64-bit
// twiddle tables serve both widths.
struct Narrow {
  uint32_t q, q2;
};

__host__ __device__ __forceinline__ Narrow narrow_of(uint64_t q) { return {(uint32_t)q, (uint32_t)(2 * q)}; }

// x < 2c, c <= 2^31 -> x mod c: x - c wraps above x exactly when x < c.
__host__ __device__ __forceinline__ uint32_t min_sub(uint32_t x, uint32_t c) {
  const uint32_t d = x - c;
  return d < x ? d : x;
}

// y * w mod q in [0, 2q) for any y < 2^32, w < q < 2^31, wp = floor(w 2^32 / q).
__host__ __device__ __forceinline__ uint32_t shoup32(uint32_t y, uint32_t w, uint32_t wp, uint32_t q) {
  const uint32_t qh = (uint32_t)(((uint64_t)y * wp) >> 32);
  return y * w - qh * q;
}

// Forward (Cooley-Tukey): x, y in [0, 4q) -> x + yW, x - yW in [0, 4q).
__host__ __device__ __forceinline__ void ct32(uint32_t &x, uint32_t &y, uint32_t w, uint32_t wp, const Narrow &m) {
  const uint32_t a = min_sub(x, m.q2);      // [0, 2q)
  const uint32_t t = shoup32(y, w, wp, m.q);  // [0, 2q)
  x = a + t;
  y = a - t + m.q2;
}

// Inverse (Gentleman-Sande): x, y in [0, 2q) -> x + y, (x - y) W in [0, 2q).
__host__ __device__ __forceinline__ void gs32(uint32_t &x, uint32_t &y, uint32_t w, uint32_t wp, const Narrow &m) {
  const uint32_t s = x + y;          // [0, 4q)
  const uint32_t d = x - y + m.q2;   // (0, 4q)
  x = min_sub(s, m.q2);
  y = shoup32(d, w, wp, m.q);
}

// [0, 4q) -> [0, q).
__host__ __device__ __forceinline__ uint32_t canon4_32(uint32_t x, const Narrow &m) {
  return min_sub(min_sub(x, m.q2), m.q);
}

}  // namespace nttamd
