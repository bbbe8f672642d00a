// Launch descriptors of the conversion / element-wise kernels.
#pragma once
#include "ntt_device.hpp"

namespace nttamd {

struct SourceConst {
  uint64_t q, inv, inv_shoup;  // b_i, [Bhat_i^-1]_{b_i} and its Shoup companion
};
struct TargetConst {
  uint64_t q, qinv_neg;
};
// Device-resident plan of one approximate CRT switch (s sources -> t targets).
struct BasisConst {
  uint32_t s, t;
  const SourceConst *src;    // [s]
  const TargetConst *dst;    // [t]
  const uint64_t *hat_mont;  // [s * t]: (Bhat_i mod m) * 2^64 mod m, source-major
  // [t][s][3][2], target-major: for limb j, hat_mont * 2^(21 j) mod m as
  // {low, high} 32-bit halves (the register kernel's form)
  const uint32_t *hat_limbs;
};

struct BasisLaunch {
  const BasisConst *plan;  // device pointer
  const uint64_t *in;      // s towers, tower-major, n each
  uint64_t *out;           // t towers
  uint64_t in_poly_stride, out_poly_stride;
  uint32_t n, targets_per_block;  // 0: chosen from the grid size
  uint32_t plan_host_s, plan_host_t;  // host copies for the launch geometry
  // target m is written to output tower (m < skip_at ? m : m + skip_len), which
  // lets a Hybrid digit's complement land around the digit's own towers.
  uint32_t skip_at, skip_len;
};

// Digit d's own Q towers (t in [d*alpha, d*alpha + alpha) and t < active_q)
// are read from `own` (the evaluation input), every other tower from `digits`.
struct MacLaunch {
  const uint64_t *own;     // (poly, tower, i)
  uint64_t own_poly_stride;
  const uint64_t *const *own_table;  // when set: polynomial p's own towers at own_table[p]
  uint32_t alpha, active_q;
  const uint64_t *digits;  // (poly, digit, tower, i)
  const uint64_t *key_b, *key_a;  // (digit, tower, i), Montgomery form
  uint64_t *out_b, *out_a;        // (poly, tower, i)
  const PrimeConst *primes;
  const uint32_t *tower_prime;    // [towers]
  uint64_t digit_poly_stride, out_poly_stride;
  uint32_t n, towers, digits_count;
  // towers [first_tower, first_tower + launch_towers) only (0 = all towers)
  uint32_t first_tower, launch_towers;
};

struct CorrectLaunch {
  uint64_t *out;
  const uint64_t *left, *right;
  uint64_t out_poly_stride, left_poly_stride, right_poly_stride;
  const PrimeConst *primes;
  const ulonglong2 *scale;  // [towers] {P^-1 mod q, Shoup}
  uint32_t first_prime, n, towers;
};

// sigma_k(B + P c0), sigma_k(A) for `polys` polynomials: polynomial j at
// out + j * out_poly_stride, ext + j * ext_poly_stride, c0 + j * c0_poly_stride.
struct RotateLaunch {
  uint64_t *out_b, *out_a;
  const uint64_t *ext_b, *ext_a, *c0;
  const PrimeConst *primes;
  const ulonglong2 *scale;  // [active_q] {P mod q, Shoup}
  uint32_t first_prime, n, log_n, active_q, automorphism;
  uint64_t ext_poly_stride, out_poly_stride, c0_poly_stride;
};

hipError_t launch_basis_switch(const BasisLaunch &a, bool per_term, uint32_t polys, hipStream_t stream);
hipError_t launch_mac_digit(uint64_t *b, uint64_t *a, const uint64_t *digit, const uint64_t *kb,
                            const uint64_t *ka, const PrimeConst *primes, uint32_t first_prime, uint32_t n,
                            uint32_t towers, bool first_digit, hipStream_t stream);
hipError_t launch_mac_fused(const MacLaunch &m, bool per_term, uint32_t polys, hipStream_t stream);

// Several per-digit MAC steps of the resident driver in one pass: digit j at
// digits[j], its canonical key rows at key_b[j], key_a[j] (same extent as b, a).
constexpr uint32_t kMacDigitsPerLaunch = 8;
struct MacDigitsLaunch {
  uint64_t *b, *a;
  const uint64_t *digits[kMacDigitsPerLaunch], *key_b[kMacDigitsPerLaunch], *key_a[kMacDigitsPerLaunch];
  const PrimeConst *primes;
  uint32_t first_prime, n, towers, count;
  int first_digit;  // 1: b, a = sum; 0: b, a += sum
};
hipError_t launch_mac_digits(const MacDigitsLaunch &m, hipStream_t stream);
hipError_t launch_to_mont(uint64_t *data, const PrimeConst *primes, const uint32_t *tower_prime, uint32_t n,
                          uint32_t towers, hipStream_t stream);
hipError_t launch_correct(const CorrectLaunch &c, uint32_t polys, hipStream_t stream);
hipError_t launch_pointwise(uint64_t *out, const uint64_t *right, const PrimeConst *primes, uint32_t first_prime,
                            uint32_t n, uint32_t towers, hipStream_t stream);
hipError_t launch_rotate(const RotateLaunch &r, uint32_t towers, uint32_t polys, hipStream_t stream);

}  // namespace nttamd

namespace nttamd {
hipError_t launch_copy_words(uint64_t *dst, const uint64_t *src, uint64_t words, hipStream_t stream);
hipError_t launch_fill_uniform(uint64_t *dst, uint64_t words, uint64_t seed, uint64_t bound, hipStream_t stream);
}
