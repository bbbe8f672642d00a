/*
 * CPU ORACLE — TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference's CPU algorithms for the CKKS hot
 * path (negacyclic NTT over 64-bit RNS towers, approximate-CRT basis
 * conversion, ModUp/ModDown, Hybrid key switching). Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load this library,
 * and only as the checker: the product path never calls it.
 *
 * Every function cites the reference file:line it restates. Arithmetic is
 * exact: unsigned __int128 products and remainders, canonical residues.
 *
 * Pinning (see tests/test_oracle.py): each function is checked against the
 * reference's known-answer tests and closed forms (N=4/q=17 worked example,
 * q=1e9+7 KATs, bit-reversal KAT, x^(N-1)*x = -1, monomial/inverse-basis
 * closed forms, naive O(N^2) evaluation, the eight pinned OpenFHE (q, psi)
 * pairs as minimum roots, an arbitrary-precision CRT sum for basis
 * conversion, independent Python big-integer restatements of ModUp, ModDown
 * and the key-switch MAC, the automorphism map and rotation against
 * a(X) -> a(X^k) mod X^N + 1 applied to coefficients and evaluated by the
 * naive transform, and a decryption check of a genuine key-switching key).
 * No function here is "parity unpinned".
 */
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

#define OR_EXPORT __attribute__((visibility("default")))

/* ---- modular arithmetic: contract docs/modular-arithmetic-contract.md:23-29 ---- */

/* canonical (a+b) mod q, 0<q<2^63 (docs/modular-arithmetic-contract.md:23) */
OR_EXPORT uint64_t or_add_mod(uint64_t a, uint64_t b, uint64_t q) {
  u128 s = (u128)a + b;
  return (uint64_t)(s >= q ? s - q : s);
}
/* canonical (a-b) mod q (docs/modular-arithmetic-contract.md:24,59-61) */
OR_EXPORT uint64_t or_sub_mod(uint64_t a, uint64_t b, uint64_t q) {
  return a >= b ? a - b : a + (q - b);
}
/* ab mod q via a 128-bit remainder (docs/modular-arithmetic-contract.md:25,62-64) */
OR_EXPORT uint64_t or_mul_mod(uint64_t a, uint64_t b, uint64_t q) {
  return (uint64_t)(((u128)a * b) % q);
}
/* square-and-multiply; q=1 returns 0 (docs/modular-arithmetic-contract.md:26,67-69;
 * tests/unit_tests.cpp:37-38) */
OR_EXPORT uint64_t or_pow_mod(uint64_t a, uint64_t e, uint64_t q) {
  uint64_t result = 1 % q, base = a % q;
  while (e) {
    if (e & 1) result = or_mul_mod(result, base, q);
    base = or_mul_mod(base, base, q);
    e >>= 1;
  }
  return result;
}
/* Fermat inverse x^(q-2); returns 0 and *ok=0 when q | x
 * (docs/modular-arithmetic-contract.md:27; tests/oracle/test_arithmetic.py:143-144) */
OR_EXPORT uint64_t or_inv_mod(uint64_t x, uint64_t q, int *ok) {
  if (q < 2 || x % q == 0) {
    if (ok) *ok = 0;
    return 0;
  }
  if (ok) *ok = 1;
  return or_pow_mod(x % q, q - 2, q);
}

/* ---- Shoup: include/ntt/shoup.hpp:7-26 ---- */
OR_EXPORT uint64_t or_shoup_precompute(uint64_t w, uint64_t q) {
  return (uint64_t)(((u128)w << 64) / q);
}
OR_EXPORT uint64_t or_shoup_multiply(uint64_t x, uint64_t w, uint64_t wp, uint64_t q) {
  uint64_t est = (uint64_t)(((u128)x * wp) >> 64);
  uint64_t r = x * w - est * q;
  return r >= q ? r - q : r;
}

/* ---- Barrett: include/ntt/barrett.hpp:6-49 ---- */
OR_EXPORT void or_barrett_precompute(uint64_t q, uint64_t out[2]) {
  u128 maximum = ~(u128)0;
  u128 mu = maximum / q + (maximum % q == q - 1);
  out[0] = (uint64_t)mu;
  out[1] = (uint64_t)(mu >> 64);
}
static uint64_t hi64(uint64_t a, uint64_t b) { return (uint64_t)(((u128)a * b) >> 64); }
OR_EXPORT uint64_t or_barrett_quotient(uint64_t low, uint64_t high, uint64_t mlo, uint64_t mhi) {
  /* low limb of floor((high:low) * mu / 2^128) (include/ntt/barrett.hpp:28-37) */
  uint64_t first = low * mhi, second = high * mlo, sum = first + second;
  uint64_t middle = sum + hi64(low, mlo);
  uint64_t carry = (uint64_t)(sum < first) + (uint64_t)(middle < sum);
  return high * mhi + hi64(low, mhi) + hi64(high, mlo) + carry;
}
OR_EXPORT uint64_t or_barrett_multiply(uint64_t a, uint64_t b, uint64_t q, uint64_t mlo, uint64_t mhi) {
  uint64_t low = a * b;
  uint64_t est = or_barrett_quotient(low, hi64(a, b), mlo, mhi);
  uint64_t r = low - est * q;
  return r >= q ? r - q : r;
}

/* ---- bit reversal: include/ntt/bitrev.hpp:23-37 ---- */
OR_EXPORT uint64_t or_log2(uint64_t n) {
  uint64_t lg = 0;
  while (((uint64_t)1 << lg) < n) ++lg;
  return lg;
}
OR_EXPORT uint64_t or_bit_reverse(uint64_t x, uint64_t lg) {
  uint64_t r = 0;
  for (uint64_t i = 0; i < lg; ++i) {
    r = (r << 1) | (x & 1);
    x >>= 1;
  }
  return r;
}

/* ---- primes and roots: include/ntt/param_gen.hpp:29-96 ---- */
static const uint64_t kBases[12] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
OR_EXPORT int or_is_prime(uint64_t n) {
  if (n < 2) return 0;
  for (int i = 0; i < 12; ++i)
    if (n % kBases[i] == 0) return n == kBases[i];
  uint64_t d = n - 1, s = 0;
  while ((d & 1) == 0) { d >>= 1; ++s; }
  for (int i = 0; i < 12; ++i) {
    uint64_t x = or_pow_mod(kBases[i] % n, d, n);
    if (x == 1 || x == n - 1) continue;
    int witness = 1;
    for (uint64_t r = 0; r + 1 < s; ++r) {
      x = or_mul_mod(x, x, n);
      if (x == n - 1) { witness = 0; break; }
    }
    if (witness) return 0;
  }
  return 1;
}
/* ascending walk from the first multiple of 2n above 2^(bits-1), +1
 * (include/ntt/param_gen.hpp:64-79); returns the count found (< count if exhausted) */
OR_EXPORT int or_find_primes(uint64_t n, uint64_t count, uint64_t bits, uint64_t *out) {
  uint64_t step = 2 * n;
  uint64_t q = (((uint64_t)1 << (bits - 1)) / step + 1) * step + 1;
  uint64_t found = 0;
  while (found < count) {
    if (q >= ((uint64_t)1 << bits)) return (int)found;
    if (or_is_prime(q)) out[found++] = q;
    q += step;
  }
  return (int)found;
}
/* xorshift candidate walk (include/ntt/param_gen.hpp:82-96) */
OR_EXPORT uint64_t or_find_psi(uint64_t q, uint64_t n) {
  uint64_t s = q ^ n;
  if (s == 0) s = 1;
  for (;;) {
    s ^= s << 13;
    s ^= s >> 7;
    s ^= s << 17;
    uint64_t g = 2 + s % (q - 3);
    uint64_t zeta = or_pow_mod(g, (q - 1) / (2 * n), q);
    if (or_pow_mod(zeta, n, q) == q - 1) return zeta;
  }
}
/* OpenFHE's parameter convention (tests/openfhe_representation_tests.cpp:55-62):
 * the root is the minimum primitive 2n-th root among psi^(2j+1). */
OR_EXPORT uint64_t or_min_root(uint64_t q, uint64_t n) {
  uint64_t psi = or_find_psi(q, n);
  uint64_t square = or_mul_mod(psi, psi, q), power = psi, best = psi;
  for (uint64_t j = 0; j < n; ++j) {
    if (power < best) best = power;
    power = or_mul_mod(power, square, q);
  }
  return best;
}
/* Largest prime below 2^bits congruent to 1 mod m, walking down by m
 * (OpenFHE LastPrime convention used by tests/cuda_openfhe_tests.cpp:133) */
OR_EXPORT uint64_t or_last_prime(uint64_t bits, uint64_t m) {
  uint64_t q = ((uint64_t)1 << bits) - m + 1;
  while (!or_is_prime(q)) q -= m;
  return q;
}
OR_EXPORT uint64_t or_previous_prime(uint64_t q, uint64_t m) {
  q -= m;
  while (!or_is_prime(q)) q -= m;
  return q;
}

/* ---- reference transform: include/ntt/ntt.hpp:96-158 ---- */
static int check_shape(uint64_t n, uint64_t q, uint64_t psi) {
  if (n < 2 || (n & (n - 1))) return -1;
  if (q < 2 || (q - 1) % (2 * n) != 0) return -2;
  if (or_pow_mod(psi, n, q) != q - 1) return -3;
  if (q <= n) return -4;
  return 0;
}
/* DIF forward, twist folded into ladder constants psi^d (include/ntt/ntt.hpp:96-121) */
OR_EXPORT int or_ntt_forward(uint64_t *a, uint64_t n, uint64_t q, uint64_t psi) {
  int err = check_shape(n, q, psi);
  if (err) return err;
  for (uint64_t d = n / 2; d >= 1; d >>= 1) {
    uint64_t ladder = or_pow_mod(psi, d, q), step = or_pow_mod(psi, n / d, q);
    for (uint64_t start = 0; start < n; start += 2 * d) {
      uint64_t w = 1;
      for (uint64_t j = 0; j < d; ++j) {
        uint64_t u = a[start + j], v = or_mul_mod(ladder, a[start + j + d], q);
        a[start + j] = or_add_mod(u, v, q);
        a[start + j + d] = or_mul_mod(or_sub_mod(u, v, q), w, q);
        w = or_mul_mod(w, step, q);
      }
    }
  }
  return 0;
}
/* DIT inverse then n^-1 (include/ntt/ntt.hpp:130-158) */
OR_EXPORT int or_ntt_inverse(uint64_t *a, uint64_t n, uint64_t q, uint64_t psi) {
  int err = check_shape(n, q, psi);
  if (err) return err;
  uint64_t psi_inv = or_inv_mod(psi, q, 0), n_inv = or_inv_mod(n % q, q, 0);
  for (uint64_t d = 1; d <= n / 2; d <<= 1) {
    uint64_t ladder = or_pow_mod(psi_inv, d, q), step = or_pow_mod(psi_inv, n / d, q);
    for (uint64_t start = 0; start < n; start += 2 * d) {
      uint64_t w = 1;
      for (uint64_t j = 0; j < d; ++j) {
        uint64_t u = a[start + j], t = or_mul_mod(a[start + j + d], w, q);
        a[start + j] = or_add_mod(u, t, q);
        a[start + j + d] = or_mul_mod(or_sub_mod(u, t, q), ladder, q);
        w = or_mul_mod(w, step, q);
      }
    }
  }
  for (uint64_t i = 0; i < n; ++i) a[i] = or_mul_mod(a[i], n_inv, q);
  return 0;
}
/* Bit-reversed power table T[bitrev(e)] = root^e, T[n] = root^n
 * (include/ntt/transform_stage.hpp:19-30) */
OR_EXPORT void or_twiddles(uint64_t n, uint64_t root, uint64_t q, uint64_t *table) {
  uint64_t width = or_log2(n), power = 1;
  for (uint64_t e = 0; e < n; ++e) {
    table[or_bit_reverse(e, width)] = power;
    power = or_mul_mod(power, root, q);
  }
  table[n] = power;
}
/* Group-twiddle CT forward / GS inverse with the bit-reversed tables
 * (include/ntt/transform_stage.hpp:40-80, launch order src/cuda/ntt_kernels.cu:286-326).
 * Same outputs as or_ntt_forward/or_ntt_inverse, one multiply per butterfly. */
static void table_transform(uint64_t *a, uint64_t n, uint64_t q, const uint64_t *t, int inverse) {
  if (!inverse) {
    for (uint64_t m = 1; m < n; m <<= 1) {
      uint64_t d = n / (2 * m);
      for (uint64_t g = 0; g < m; ++g) {
        uint64_t w = t[m + g];
        for (uint64_t j = 0; j < d; ++j) {
          uint64_t i = 2 * g * d + j;
          uint64_t x = a[i], y = or_mul_mod(a[i + d], w, q);
          a[i] = or_add_mod(x, y, q);
          a[i + d] = or_sub_mod(x, y, q);
        }
      }
    }
  } else {
    for (uint64_t m = n / 2; m >= 1; m >>= 1) {
      uint64_t d = n / (2 * m);
      for (uint64_t g = 0; g < m; ++g) {
        uint64_t w = t[m + g];
        for (uint64_t j = 0; j < d; ++j) {
          uint64_t i = 2 * g * d + j;
          uint64_t x = a[i], y = a[i + d];
          a[i] = or_add_mod(x, y, q);
          a[i + d] = or_mul_mod(or_sub_mod(x, y, q), w, q);
        }
      }
    }
    uint64_t n_inv = or_inv_mod(n % q, q, 0);
    for (uint64_t i = 0; i < n; ++i) a[i] = or_mul_mod(a[i], n_inv, q);
  }
}
OR_EXPORT int or_ntt_table(uint64_t *a, uint64_t n, uint64_t q, uint64_t psi, int inverse) {
  int err = check_shape(n, q, psi);
  if (err) return err;
  uint64_t *t = (uint64_t *)malloc((n + 1) * sizeof(uint64_t));
  if (!t) return -5;
  or_twiddles(n, inverse ? or_inv_mod(psi, q, 0) : psi, q, t);
  table_transform(a, n, q, t, inverse);
  free(t);
  return 0;
}
/* Many towers at once (tower-major, t*n + i), each with its own (q, psi);
 * towers run in parallel when built with OpenMP. */
OR_EXPORT int or_ntt_towers(uint64_t *a, uint64_t n, uint64_t towers, const uint64_t *q,
                            const uint64_t *psi, int inverse) {
  int err = 0;
#pragma omp parallel for schedule(dynamic) reduction(| : err)
  for (long k = 0; k < (long)towers; ++k) err |= or_ntt_table(a + k * n, n, q[k], psi[k], inverse) != 0;
  return err ? -1 : 0;
}
/* O(N^2) definition A_k = sum_i a_i psi^((2 rev(k)+1) i), bit-reversed output order
 * (docs/openfhe-representation-contract.md:81-180; tests/oracle/ntt_oracle.py:64-77) */
OR_EXPORT void or_ntt_naive(const uint64_t *a, uint64_t *out, uint64_t n, uint64_t q, uint64_t psi) {
  uint64_t lg = or_log2(n);
  for (uint64_t k = 0; k < n; ++k) {
    uint64_t point = or_pow_mod(psi, 2 * or_bit_reverse(k, lg) + 1, q);
    uint64_t acc = 0, power = 1;
    for (uint64_t i = 0; i < n; ++i) {
      acc = or_add_mod(acc, or_mul_mod(a[i], power, q), q);
      power = or_mul_mod(power, point, q);
    }
    out[k] = acc;
  }
}

/* ---- approximate CRT basis switch: include/ntt/basis_conversion.hpp:13-114 ---- */
/* Returns 0, or a negative code for the plan's rejections (:41-53, :91-100). */
OR_EXPORT int or_basis_switch(const uint64_t *in, uint64_t n, const uint64_t *src, uint64_t s,
                              const uint64_t *dst, uint64_t t, uint64_t *out) {
  if (s == 0 || s > 40 || t == 0 || t > 40) return -1;
  uint64_t all[80];
  for (uint64_t i = 0; i < s; ++i) all[i] = src[i];
  for (uint64_t i = 0; i < t; ++i) all[s + i] = dst[i];
  for (uint64_t i = 0; i < s + t; ++i)
    for (uint64_t j = i + 1; j < s + t; ++j)
      if (all[i] == all[j]) return -2;
  for (uint64_t i = 0; i < s + t; ++i)
    if (all[i] < 3 || all[i] >= ((uint64_t)1 << 63) || !or_is_prime(all[i])) return -3;
  if (n < 2 || n > 65536 || (n & (n - 1))) return -4;
  for (uint64_t k = 0; k < s; ++k)
    for (uint64_t i = 0; i < n; ++i)
      if (in[k * n + i] >= src[k]) return -5;
  uint64_t inv[40], invq[40];
  uint64_t *hat = (uint64_t *)malloc(s * t * sizeof(uint64_t));
  if (!hat) return -6;
  for (uint64_t k = 0; k < s; ++k) {
    uint64_t h = 1;
    for (uint64_t o = 0; o < s; ++o)
      if (o != k) h = or_mul_mod(h, src[o] % src[k], src[k]);
    inv[k] = or_inv_mod(h, src[k], 0);
    invq[k] = or_shoup_precompute(inv[k], src[k]);
    for (uint64_t m = 0; m < t; ++m) {
      uint64_t r = 1;
      for (uint64_t o = 0; o < s; ++o)
        if (o != k) r = or_mul_mod(r, src[o] % dst[m], dst[m]);
      hat[k * t + m] = r;
    }
  }
#pragma omp parallel for schedule(static)
  for (long m = 0; m < (long)t; ++m)
    for (uint64_t i = 0; i < n; ++i) {
      uint64_t sum = 0;
      for (uint64_t k = 0; k < s; ++k) {
        uint64_t scaled = or_shoup_multiply(in[k * n + i], inv[k], invq[k], src[k]);
        sum = or_add_mod(sum, or_mul_mod(scaled % dst[m], hat[k * t + m], dst[m]), dst[m]);
      }
      out[m * n + i] = sum;
    }
  free(hat);
  return 0;
}

/* ---- resident chain formulas: src/resident.cpp:404-570 ---- */
/* P mod q (include/ntt/resident_conversion.hpp:8-14) */
OR_EXPORT uint64_t or_special_product(const uint64_t *p, uint64_t k, uint64_t q) {
  uint64_t r = 1;
  for (uint64_t j = 0; j < k; ++j) r = or_mul_mod(r, p[j] % q, q);
  return r;
}
/* ModUp of one polynomial (src/resident.cpp:431-447): Q towers -> evaluation QP.
 * in_eval != 0: input is evaluation Q; otherwise coefficient Q. */
OR_EXPORT int or_mod_up(const uint64_t *in, int in_eval, uint64_t n, const uint64_t *q,
                        const uint64_t *qpsi, uint64_t L, const uint64_t *p, const uint64_t *ppsi,
                        uint64_t K, uint64_t *out) {
  uint64_t *coeff = (uint64_t *)malloc(L * n * sizeof(uint64_t));
  if (!coeff) return -6;
  memcpy(coeff, in, L * n * sizeof(uint64_t));
  memcpy(out, in, L * n * sizeof(uint64_t));
  int err = in_eval ? or_ntt_towers(coeff, n, L, q, qpsi, 1) : or_ntt_towers(out, n, L, q, qpsi, 0);
  if (!err) err = or_basis_switch(coeff, n, q, L, p, K, out + L * n);
  if (!err) err = or_ntt_towers(out + L * n, n, K, p, ppsi, 0);
  free(coeff);
  return err;
}
/* ModDown (src/resident.cpp:448-457, correction :404-430): evaluation QP -> evaluation Q */
OR_EXPORT int or_mod_down(const uint64_t *in, uint64_t n, const uint64_t *q, const uint64_t *qpsi,
                          uint64_t L, const uint64_t *p, const uint64_t *ppsi, uint64_t K,
                          uint64_t *out) {
  uint64_t *pc = (uint64_t *)malloc(K * n * sizeof(uint64_t));
  uint64_t *corr = (uint64_t *)malloc(L * n * sizeof(uint64_t));
  int err = (!pc || !corr) ? -6 : 0;
  if (!err) {
    memcpy(pc, in + L * n, K * n * sizeof(uint64_t));
    err = or_ntt_towers(pc, n, K, p, ppsi, 1);
  }
  if (!err) err = or_basis_switch(pc, n, p, K, q, L, corr);
  if (!err) err = or_ntt_towers(corr, n, L, q, qpsi, 0);
  if (!err)
    for (uint64_t l = 0; l < L; ++l) {
      uint64_t pinv = or_inv_mod(or_special_product(p, K, q[l]), q[l], 0);
      for (uint64_t i = 0; i < n; ++i)
        out[l * n + i] = or_mul_mod(or_sub_mod(in[l * n + i], corr[l * n + i], q[l]), pinv, q[l]);
    }
  free(pc);
  free(corr);
  return err;
}
/* Hybrid digits (src/resident.cpp:458-491): evaluation Q (L towers) ->
 * ceil(L/alpha) evaluation QP polynomials, digit-major, each (L+K)*n. */
OR_EXPORT int or_hybrid_digits(const uint64_t *in, uint64_t n, const uint64_t *q,
                               const uint64_t *qpsi, uint64_t L, const uint64_t *p,
                               const uint64_t *ppsi, uint64_t K, uint64_t alpha, uint64_t *out) {
  uint64_t T = L + K;
  uint64_t *mod = (uint64_t *)malloc(T * sizeof(uint64_t)), *root = (uint64_t *)malloc(T * sizeof(uint64_t));
  uint64_t *coeff = (uint64_t *)malloc(alpha * n * sizeof(uint64_t));
  uint64_t *tq = (uint64_t *)malloc(T * sizeof(uint64_t)), *tr = (uint64_t *)malloc(T * sizeof(uint64_t));
  uint64_t *conv = (uint64_t *)malloc(T * n * sizeof(uint64_t));
  int err = (!mod || !root || !coeff || !tq || !tr || !conv) ? -6 : 0;
  if (!err) {
    for (uint64_t l = 0; l < L; ++l) { mod[l] = q[l]; root[l] = qpsi[l]; }
    for (uint64_t k = 0; k < K; ++k) { mod[L + k] = p[k]; root[L + k] = ppsi[k]; }
  }
  for (uint64_t first = 0, d = 0; !err && first < L; first += alpha, ++d) {
    uint64_t count = alpha < L - first ? alpha : L - first;
    uint64_t *digit = out + d * T * n;
    memcpy(coeff, in + first * n, count * n * sizeof(uint64_t));
    err = or_ntt_towers(coeff, n, count, q + first, qpsi + first, 1);
    uint64_t c = 0;
    for (uint64_t l = 0; l < T; ++l)
      if (l < first || l >= first + count) { tq[c] = mod[l]; tr[c] = root[l]; ++c; }
    /* chunks of at most 40 targets are independent: one call covers them all */
    for (uint64_t b = 0; !err && b < c; b += 40) {
      uint64_t len = c - b < 40 ? c - b : 40;
      err = or_basis_switch(coeff, n, q + first, count, tq + b, len, conv + b * n);
    }
    if (!err) err = or_ntt_towers(conv, n, c, tq, tr, 0);
    c = 0;
    for (uint64_t l = 0; !err && l < T; ++l) {
      if (l >= first && l < first + count)
        memcpy(digit + l * n, in + l * n, n * sizeof(uint64_t));
      else
        memcpy(digit + l * n, conv + (c++) * n, n * sizeof(uint64_t));
    }
  }
  free(mod); free(root); free(coeff); free(tq); free(tr); free(conv);
  return err;
}
/* Active key tower for active index i (include/ntt/backend_contract.hpp:112-119) */
static uint64_t key_index(uint64_t i, uint64_t L, uint64_t M) { return i < L ? i : i + (M - L); }
/* MAC over all digits (src/resident.cpp:530-570, include/ntt/resident_conversion.hpp:50-54).
 * key_b/key_a hold the full-context key: digit-major, each (M+K)*n. */
OR_EXPORT void or_hybrid_mac(const uint64_t *digits, uint64_t beta, const uint64_t *key_b,
                             const uint64_t *key_a, uint64_t n, const uint64_t *q, uint64_t L,
                             const uint64_t *p, uint64_t K, uint64_t M, uint64_t *out_b,
                             uint64_t *out_a) {
  uint64_t T = L + K, full = M + K;
#pragma omp parallel for schedule(static)
  for (long l = 0; l < (long)T; ++l) {
    uint64_t mod = (uint64_t)l < L ? q[l] : p[l - L], kl = key_index((uint64_t)l, L, M);
    for (uint64_t i = 0; i < n; ++i) {
      uint64_t zb = 0, za = 0;
      for (uint64_t d = 0; d < beta; ++d) {
        uint64_t x = digits[(d * T + l) * n + i];
        zb = or_add_mod(zb, or_mul_mod(x, key_b[(d * full + kl) * n + i], mod), mod);
        za = or_add_mod(za, or_mul_mod(x, key_a[(d * full + kl) * n + i], mod), mod);
      }
      out_b[l * n + i] = zb;
      out_a[l * n + i] = za;
    }
  }
}
/* Automorphism address map (include/ntt/resident_conversion.hpp:15-20) */
OR_EXPORT uint64_t or_automorphism_source(uint64_t dest, uint64_t n, uint32_t k) {
  uint64_t lg = or_log2(n);
  uint64_t odd = 2 * or_bit_reverse(dest, lg) + 1;
  uint64_t src = (odd * k) & (2 * n - 1);
  return or_bit_reverse(src / 2, lg);
}
/* sigma_k(B + P*c0), sigma_k(A) on evaluation QP (src/resident.cpp:571-617,
 * include/ntt/resident_conversion.hpp:22-29, src/cuda/resident_adapter.inc:44-59) */
OR_EXPORT void or_hybrid_rotate(const uint64_t *ext_b, const uint64_t *ext_a, const uint64_t *c0,
                                uint64_t n, const uint64_t *q, uint64_t L, const uint64_t *p,
                                uint64_t K, uint32_t k, uint64_t *out_b, uint64_t *out_a) {
  for (uint64_t t = 0; t < L + K; ++t) {
    uint64_t mod = t < L ? q[t] : p[t - L];
    uint64_t scale = t < L ? or_special_product(p, K, mod) : 0;
    for (uint64_t slot = 0; slot < n; ++slot) {
      uint64_t src = t * n + or_automorphism_source(slot, n, k);
      uint64_t b = ext_b[src];
      if (t < L) b = or_add_mod(b, or_mul_mod(c0[src], scale, mod), mod);
      out_b[t * n + slot] = b;
      out_a[t * n + slot] = ext_a[src];
    }
  }
}
/* KeySwitchHybrid core (include/ntt/openfhe_hybrid_adapter.hpp:163-181,212-219):
 * digits -> MAC -> ModDown(B), ModDown(A). in: evaluation Q (L towers). */
OR_EXPORT int or_key_switch(const uint64_t *in, uint64_t n, const uint64_t *q, const uint64_t *qpsi,
                            uint64_t L, uint64_t M, const uint64_t *p, const uint64_t *ppsi,
                            uint64_t K, uint64_t alpha, const uint64_t *key_b,
                            const uint64_t *key_a, uint64_t *out_b, uint64_t *out_a) {
  uint64_t beta = (L + alpha - 1) / alpha, T = L + K;
  uint64_t *digits = (uint64_t *)malloc(beta * T * n * sizeof(uint64_t));
  uint64_t *zb = (uint64_t *)malloc(T * n * sizeof(uint64_t)), *za = (uint64_t *)malloc(T * n * sizeof(uint64_t));
  int err = (!digits || !zb || !za) ? -6 : 0;
  if (!err) err = or_hybrid_digits(in, n, q, qpsi, L, p, ppsi, K, alpha, digits);
  if (!err) {
    or_hybrid_mac(digits, beta, key_b, key_a, n, q, L, p, K, M, zb, za);
    err = or_mod_down(zb, n, q, qpsi, L, p, ppsi, K, out_b);
  }
  if (!err) err = or_mod_down(za, n, q, qpsi, L, p, ppsi, K, out_a);
  free(digits); free(zb); free(za);
  return err;
}

/* elementwise a*b mod q per tower (reference src/cuda/ntt_kernels.cu:220-234 semantics) */
OR_EXPORT void or_pointwise_towers(const uint64_t *a, const uint64_t *b, uint64_t *out, uint64_t n,
                                   uint64_t towers, const uint64_t *q) {
  for (uint64_t k = 0; k < towers; ++k)
    for (uint64_t i = 0; i < n; ++i) out[k * n + i] = or_mul_mod(a[k * n + i], b[k * n + i], q[k]);
}

/* thread count of the OpenMP tower loops (bench CPU baseline reports it) */
OR_EXPORT void or_set_threads(int threads) { omp_set_num_threads(threads); }
OR_EXPORT int or_max_threads(void) { return omp_get_max_threads(); }

/* ---- seeded inputs: include/ntt/benchmark_inputs.hpp:9-22 (MMIX LCG) ---- */
OR_EXPORT uint64_t or_lcg_fill(uint64_t state, uint64_t *out, uint64_t n, uint64_t towers,
                               const uint64_t *q) {
  for (uint64_t k = 0; k < towers; ++k)
    for (uint64_t i = 0; i < n; ++i) {
      state = state * 6364136223846793005ULL + 1442695040888963407ULL;
      out[k * n + i] = state % q[k];
    }
  return state;
}
/* splitmix finaliser used by the reference fixtures (tests/openfhe_hybrid_tests.cpp:153-158) */
OR_EXPORT uint64_t or_mix(uint64_t x) {
  x += 0x9e3779b97f4a7c15ULL;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
  return x ^ (x >> 31);
}
