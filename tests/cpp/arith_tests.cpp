// Host-side checks of the wide-lazy NTT arithmetic (ntt-gpu_amd/csrc/modarith.hpp)
// against exact 128-bit arithmetic: the same __host__ __device__ functions the
// gfx950 kernels inline. Every claimed range is checked on the extremes
// (y = 2^64 - 1, x = 8q - 1, w = q - 1, q just below 2^61) and on seeded
// random operands for OpenFHE-sized and adversarial primes.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <initializer_list>
#include <random>
#include <string>
#include <vector>

#include "../../ntt-gpu_amd/csrc/modarith.hpp"

using u128 = unsigned __int128;
using namespace nttamd;

namespace {
size_t checks = 0;
void check(bool ok, const std::string &why) {
  ++checks;
  if (!ok) {
    std::fprintf(stderr, "FAIL: %s\n", why.c_str());
    std::exit(1);
  }
}
uint64_t mulmod(uint64_t a, uint64_t b, uint64_t q) { return (uint64_t)((u128)a * b % q); }
uint64_t shoup_companion(uint64_t w, uint64_t q) { return (uint64_t)(((u128)w << 64) / q); }
uint64_t shoup_lazy_host(uint64_t x, uint64_t w, uint64_t wp, uint64_t q) {
  return x * w - (uint64_t)(((u128)x * wp) >> 64) * q;
}
uint64_t reduce_once(uint64_t x, uint64_t q) { return x >= q ? x - q : x; }
// reduce2_est's per-prime constants, as the library computes them
void est_constants(uint64_t q, uint32_t &shift, uint32_t &mul) {
  const int bits = 64 - __builtin_clzll(q);
  shift = (uint32_t)(bits - 24);
  mul = (uint32_t)(((u128)1 << (bits + 31)) / q);
}

void one(uint64_t q, uint64_t w, uint64_t x, uint64_t y) {
  const Lazy8 m = lazy8_of(q);
  const uint64_t wp = shoup_companion(w, q);
  // quotient estimate: short by 0, 1 or 2
  const uint64_t hi = (uint64_t)(((u128)y * wp) >> 64), est = mulhi_est(y, wp);
  check(est <= hi && hi - est <= 2, "mulhi_est within 2 of the high product");
  // multiply: congruent, below 4q
  const uint64_t t = shoup4(y, w, wp, m);
  check(t < 4 * q && t % q == mulmod(y % q, w, q), "shoup4 in [0, 4q) and congruent");
  // forward butterfly on [0, 8q)
  uint64_t cx = x % (8 * q), cy = y % (8 * q);
  const uint64_t fx = cx, fy = cy;
  ct8(cx, cy, w, wp, m);
  const uint64_t prod = mulmod(fy % q, w, q);
  check(cx < 8 * q && cy < 8 * q, "ct8 outputs below 8q");
  check(cx % q == (fx % q + prod) % q && cy % q == (fx % q + q - prod) % q, "ct8 congruent");
  // inverse butterfly on [0, 4q)
  uint64_t gx = x % (4 * q), gy = y % (4 * q);
  const uint64_t ix = gx, iy = gy;
  gs8(gx, gy, w, wp, m);
  check(gx < 4 * q && gy < 4 * q, "gs8 outputs below 4q");
  check(gx % q == (ix % q + iy % q) % q && gy % q == mulmod((ix % q + q - iy % q) % q, w, q),
        "gs8 congruent");
  // canonical form of [0, 8q)
  check(canon8(x % (8 * q), m) == x % (8 * q) % q, "canon8");
}
}  // namespace

uint64_t powmod(uint64_t b, uint64_t e, uint64_t q) {
  uint64_t r = 1 % q;
  for (b %= q; e; e >>= 1, b = mulmod(b, b, q))
    if (e & 1) r = mulmod(r, b, q);
  return r;
}
bool is_prime(uint64_t n) {
  if (n < 2) return false;
  for (uint64_t p : {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37})
    if (n % p == 0) return n == p;
  uint64_t d = n - 1;
  int s = 0;
  while (!(d & 1)) d >>= 1, ++s;
  for (uint64_t a : {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37}) {
    uint64_t x = powmod(a, d, n);
    if (x == 1 || x == n - 1) continue;
    bool composite = true;
    for (int r = 1; r < s && composite; ++r) composite = (x = mulmod(x, x, n)) != n - 1;
    if (composite) return false;
  }
  return true;
}
uint32_t brev(uint32_t x, int bits) {
  uint32_t r = 0;
  for (int i = 0; i < bits; ++i) r = (r << 1) | ((x >> i) & 1);
  return r;
}

// The forward reduction schedule of every headroom class H, run over a whole
// 2^16-point transform with the largest NTT prime of that class: every
// butterfly's exact value stays below 2^64, every residue below the plan's
// bound, and the canonicalised output equals an exact 128-bit transform.
void planned_forward_transforms(std::mt19937_64 &rng) {
  const int n = 1 << 16, lg = 16;
  const int reductions[6] = {16, 7, 3, 1, 0, 0};
  for (int H = 1; H <= 5; ++H) {
    const FwdPlanData plan = make_fwd_plan(H);
    int count = 0;
    for (int g = 0; g < 16; ++g) count += plan.red[g] >= 0;
    check(count == reductions[H], "reductions of plan " + std::to_string(H));
    uint64_t q = (uint64_t{1} << (61 - H)) - (uint64_t{1} << 17) + 1;
    while (!is_prime(q)) q -= uint64_t{1} << 17;
    uint64_t psi = 0;
    for (uint64_t g = 2; !psi; ++g) {
      const uint64_t z = powmod(g, (q - 1) / (2 * n), q);
      if (powmod(z, n, q) == q - 1) psi = z;
    }
    std::vector<uint64_t> tw(n), twp(n);
    for (uint32_t e = 0; e < (uint32_t)n; ++e) {
      tw[brev(e, lg)] = powmod(psi, e, q);
    }
    for (int k = 0; k < n; ++k) twp[k] = shoup_companion(tw[k], q);
    const Lazy8 m = lazy8_of(q);
    for (int pattern = 0; pattern < 2; ++pattern) {
      std::vector<uint64_t> a(n), want(n);
      for (int i = 0; i < n; ++i) a[i] = pattern == 0 ? q - 1 : rng() % q;
      want = a;
      uint64_t bound = 1;  // in units of q, tracked from the plan
      for (int g = 0; g < lg; ++g) {
        const int mm = 1 << g, d = n >> (g + 1);
        const int k = plan.red[g];
        bound = k >= 0 ? (uint64_t{1} << k) + 4 : bound + 4;
        for (int grp = 0; grp < mm; ++grp) {
          const uint64_t w = tw[mm + grp], wp = twp[mm + grp];
          for (int j = 0; j < d; ++j) {
            const int i = 2 * grp * d + j;
            // exact reference on canonical values
            const uint64_t t = mulmod(want[i + d], w, q);
            const uint64_t rx = (want[i] + t) % q, ry = (want[i] + q - t) % q;
            want[i] = rx;
            want[i + d] = ry;
            // the planned butterfly
            const uint64_t x = a[i], y = a[i + d];
            const uint64_t xr = k >= 0 ? sub_if_ge_neg(x, 0 - (q << k)) : x;
            const uint64_t xn = mad_lo(xr, y, w, mulhi_est(y, wp), m.nq);
            const u128 exact = (u128)xr + ((u128)xn - xr);  // xr + T as computed
            check(xn >= xr && xn - xr < 4 * q && (xn - xr) % q == mulmod(y % q, w, q) &&
                      exact < ((u128)1 << 64),
                  "planned butterfly: T in [0, 4q), no wrap");
            a[i] = xn;
            a[i + d] = 2 * xr + m.q4 - xn;
            check((u128)a[i] < (u128)bound * q && (u128)a[i + d] < (u128)bound * q, "residues within the plan bound");
          }
        }
      }
      check(bound <= (uint64_t{1} << plan.bound_log), "plan bound_log covers the output");
      uint32_t shift, mul;
      est_constants(q, shift, mul);
      for (int i = 0; i < n; ++i) {
        uint64_t x = a[i], y = a[i];
        for (int k = plan.bound_log - 1; k >= 0; --k) x = sub_if_ge_neg(x, 0 - (q << k));
        y = sub_if_ge_neg(reduce2_est(y, shift, mul, m.nq), 0 - q);
        if (x != want[i] || y != want[i])
          check(false, "planned transform equals the exact transform, H = " + std::to_string(H));
      }
      ++checks;
    }
  }
}

// The inverse reduction schedule of every headroom class H (make_inv_plan),
// run over a whole 2^16-point Gentleman-Sande transform with the largest NTT
// prime of the class: d = x + 2^c q - y never wraps and stays below 2^64, the
// sums and products stay below the plan's bound for the next stage, every
// estimate's input is below 2^7 q, and the last stage (n^-1 folded in, its
// outputs reduced by the exact Shoup product and one subtract) returns the
// exact inverse transform.
void planned_inverse_transforms(std::mt19937_64 &rng) {
  const int n = 1 << 16, lg = 16;
  const int estimates[6] = {0, 7, 4, 3, 2, 2};
  for (int H = 0; H <= 5; ++H) {
    const InvPlanData plan = make_inv_plan(H);
    int est = 0, sub = 0;
    for (int g = 0; g < 16; ++g) est += plan.red[g] == 2, sub += plan.red[g] == 1;
    check(est == estimates[H] && (H == 0 ? sub == 14 : sub == 0), "reductions of inverse plan " + std::to_string(H));
    uint64_t q = (uint64_t{1} << (61 - H)) - (uint64_t{1} << 17) + 1;
    while (!is_prime(q)) q -= uint64_t{1} << 17;
    uint64_t psi = 0;
    for (uint64_t g = 2; !psi; ++g) {
      const uint64_t z = powmod(g, (q - 1) / (2 * n), q);
      if (powmod(z, n, q) == q - 1) psi = z;
    }
    const uint64_t psi_inv = powmod(psi, 2 * (uint64_t)n - 1, q);
    std::vector<uint64_t> tw(n), twp(n);
    for (uint32_t e = 0; e < (uint32_t)n; ++e) tw[brev(e, lg)] = powmod(psi_inv, e, q);
    for (int k = 0; k < n; ++k) twp[k] = shoup_companion(tw[k], q);
    const uint64_t ninv = powmod(n, q - 2, q), last_w = mulmod(tw[1], ninv, q);
    const uint64_t ninv_p = shoup_companion(ninv, q), last_wp = shoup_companion(last_w, q);
    uint32_t shift, mul;
    est_constants(q, shift, mul);
    const Lazy8 m = lazy8_of(q);
    for (int pattern = 0; pattern < 2; ++pattern) {
      std::vector<uint64_t> a(n), want(n);
      for (int i = 0; i < n; ++i) a[i] = pattern == 0 ? q - 1 : rng() % q;
      want = a;
      for (int g = 0; g < lg; ++g) {
        const int d = 1 << g, groups = n >> (g + 1);
        const uint64_t c = q << plan.c_log[g];
        const u128 in_bound = (u128)c;  // inputs below 2^c_log q
        for (int grp = 0; grp < groups; ++grp) {
          const uint64_t w = tw[groups + grp], wp = twp[groups + grp];
          for (int j = 0; j < d; ++j) {
            const int i = 2 * grp * d + j;
            const uint64_t x = a[i], y = a[i + d];
            check((u128)x < in_bound && (u128)y < in_bound, "inverse inputs within 2^c_log q");
            const u128 s = (u128)x + y, dd = (u128)x + c - y;
            check(s < ((u128)1 << 64) && dd < ((u128)1 << 64), "inverse sum and difference fit 64 bits");
            if (g == lg - 1) {
              // exact reference: (x + y) n^-1, (x - y) W n^-1
              const uint64_t ex = mulmod((want[i] + want[i + d]) % q, ninv, q);
              const uint64_t ey = mulmod((want[i] + q - want[i + d]) % q, last_w, q);
              want[i] = ex;
              want[i + d] = ey;
              a[i] = reduce_once(shoup_lazy_host((uint64_t)s, ninv, ninv_p, q), q);
              a[i + d] = reduce_once(shoup_lazy_host((uint64_t)dd, last_w, last_wp, q), q);
              continue;
            }
            const uint64_t ex = (want[i] + want[i + d]) % q, ey = mulmod((want[i] + q - want[i + d]) % q, w, q);
            want[i] = ex;
            want[i + d] = ey;
            uint64_t xs = (uint64_t)s;
            if (plan.red[g] == 2) {
              check(s < ((u128)q << 7), "estimate input below 2^7 q");
              xs = reduce2_est(xs, shift, mul, m.nq);
              check(xs < 2 * q, "estimate output below 2q");
            } else if (plan.red[g] == 1) {
              xs = sub_if_ge_neg(xs, m.nq4);
            }
            const uint64_t ys = shoup4((uint64_t)dd, w, wp, m);
            check(ys < 4 * q, "inverse product below 4q");
            check(xs % q == ex && ys % q == ey, "planned inverse butterfly congruent");
            a[i] = xs;
            a[i + d] = ys;
          }
        }
      }
      for (int i = 0; i < n; ++i)
        if (a[i] != want[i]) check(false, "planned inverse equals the exact transform, H = " + std::to_string(H));
      ++checks;
    }
  }
}

int main() {
  std::mt19937_64 rng(0x7251d9);
  planned_forward_transforms(rng);
  planned_inverse_transforms(rng);
  // OpenFHE-shaped 60- and 50-bit primes, a prime just below 2^61 and tiny ones
  const uint64_t primes[] = {1152921504606584833ULL, 1125899906826241ULL, 2305843009213693951ULL,
                             1152921504598720513ULL, 562949953216513ULL,  65537ULL, 17ULL};
  for (uint64_t q : primes) {
    const uint64_t ws[] = {0, 1, 2, q - 1, q - 2, q / 2, q / 3 + 1};
    const uint64_t xs[] = {0, 1, q - 1, q, 2 * q - 1, 4 * q - 1, 4 * q, 8 * q - 1, ~0ULL};
    for (uint64_t w : ws)
      for (uint64_t x : xs)
        for (uint64_t y : xs) one(q, w, x, y);
    for (int i = 0; i < 200000; ++i) {
      const uint64_t w = rng() % q;
      const uint64_t x = rng(), y = (i & 1) ? rng() : ~0ULL - (rng() & 0xffff);
      one(q, w, x, y);
    }
  }
  // quotient-estimate reduction of x < 2^7 q (and x < 2^64) to [0, 2q), bit
  // lengths 24..63, at the edges (k q - 1, k q, the bound - 1) and at random
  for (int i = 0; i < 3000000; ++i) {
    const int bits = 24 + i % 40;
    uint64_t q = (uint64_t{1} << (bits - 1)) | (rng() >> (65 - bits)) | 1;
    if (i % 9 == 0) q = (bits == 64 ? ~0ULL : (uint64_t{1} << bits) - 1) - 2 * (rng() & 0xff);
    if (i % 11 == 0) q = (uint64_t{1} << (bits - 1)) + 1 + 2 * (rng() & 0xff);
    uint32_t shift, mul;
    est_constants(q, shift, mul);
    const u128 bound = std::min((u128)q << 7, (u128)1 << 64);
    const uint64_t k = 1 + rng() % 127;
    uint64_t x;
    switch (i % 6) {
      case 0: x = (uint64_t)(bound - 1); break;
      case 1: x = (u128)k * q < bound ? k * q : 0; break;
      case 2: x = (u128)k * q - 1 < bound ? k * q - 1 : q - 1; break;
      default: x = (uint64_t)((((u128)rng() << 64) | rng()) % bound);
    }
    const uint64_t r = reduce2_est(x, shift, mul, 0 - q);
    check(mul >= (1u << 31) && r < 2 * q && r % q == x % q, "reduce2_est in [0, 2q) and congruent");
  }
  // mulhi_est on arbitrary 64-bit operands (not only Shoup companions)
  for (int i = 0; i < 1000000; ++i) {
    const uint64_t y = rng() | ((i & 3) == 0 ? 0xffffffff00000000ULL : 0), p = rng() | ((i & 7) == 1 ? ~0ULL : 0);
    const uint64_t hi = (uint64_t)(((u128)y * p) >> 64), est = mulhi_est(y, p);
    check(est <= hi && hi - est <= 2, "mulhi_est on arbitrary operands");
  }
  // sign-selected conditional subtract: c <= 2^63, x < 2c
  for (int i = 0; i < 2000000; ++i) {
    const uint64_t c = (i % 7 == 0) ? (1ULL << 63) - (rng() & 0xff) : (rng() >> (1 + i % 40)) | 1;
    const uint64_t x = (i % 5 == 0) ? (c - 1 + (i % 3)) : (c >> 63) ? rng() : rng() % (2 * c);  // x < 2c
    check(sub_if_ge_neg(x, 0 - c) == (x >= c ? x - c : x), "sub_if_ge_neg");
  }
  const uint64_t edges[] = {1, 2, uint64_t{1} << 63, (uint64_t{1} << 62) + 5, 2305843009213693951ULL};
  for (const uint64_t c : edges)
    for (const uint64_t x : {uint64_t{0}, c - 1, c, 2 * c - 1})
      check(sub_if_ge_neg(x, 0 - c) == (x >= c ? x - c : x), "sub_if_ge_neg at the edges");
  // register conversion kernel: 21-bit source limbs y_kj against
  // h_kj = hat_k 2^(21 j) mod m in two 32-bit-aligned columns; up to 16
  // sources below 2^63, m below 2^63; the columns never overflow and the
  // Montgomery reduction of C_lo + 2^32 C_hi is (sum_k y_k hat_k) 2^-64 mod m
  for (int i = 0; i < 300000; ++i) {
    const int s = 1 + (int)(rng() % 16);
    const bool extreme = (i % 5) == 0;
    uint64_t m = ((rng() >> (1 + i % 33)) | 1) | (uint64_t{1} << 24);
    if (i % 7 == 0) m = (uint64_t{1} << 63) - 1 - 2 * (rng() & 0xffff);
    u128 lo = 0, hi = 0, want = 0;
    for (int k = 0; k < s; ++k) {
      const uint64_t y = extreme ? (uint64_t{1} << 63) - 1 : rng() >> 1;
      uint64_t h = extreme ? m - 1 : rng() % m;
      want = (want + (u128)y * h) % m;
      for (int j = 0; j < 3; ++j) {
        const uint32_t yj = j < 2 ? (uint32_t)(y >> (21 * j)) & 0x1fffff : (uint32_t)(y >> 42);
        lo += (u128)yj * (uint32_t)h;
        hi += (u128)yj * (uint32_t)(h >> 32);
        h = (uint64_t)(((u128)h << 21) % m);
      }
    }
    check(lo < ((u128)1 << 64) && hi < ((u128)1 << 64), "conversion columns fit 64 bits");
    const uint64_t l64 = (uint64_t)lo, h64 = (uint64_t)hi;
    U128 sum{l64 + (h64 << 32), h64 >> 32};
    sum.hi += sum.lo < l64 ? 1u : 0u;
    check(((((u128)sum.hi << 64) | sum.lo) % m) == want && (((u128)sum.hi << 64) | sum.lo) < ((u128)m << 64),
          "conversion columns congruent and below m 2^64");
    const uint64_t qinv_neg = [&] {
      uint64_t inv = 1;
      for (int b = 0; b < 6; ++b) inv *= 2 - m * inv;
      return 0 - inv;
    }();
    const uint64_t r = mont_reduce(sum, m, qinv_neg);
    check(((u128)r << 64) % m == want && r < m, "Montgomery reduction of the columns");
  }
  // the transform kernels' launch-index split y / polys as y * ceil(2^32 / polys) >> 32,
  // used whenever (largest y + 1) * polys <= 2^32
  for (int i = 0; i < 3000000; ++i) {
    const uint64_t p = 1 + (i % 3 == 0 ? rng() % 64 : rng() % 70000);
    const uint64_t y_end = (uint64_t{1} << 32) / p;  // the largest range the launcher allows
    const uint64_t y = (i % 4 == 0) ? y_end - 1 - (rng() % 3 < y_end ? rng() % 3 : 0) : rng() % y_end;
    const uint64_t mul = ((uint64_t{1} << 32) + p - 1) / p;
    check(((y * mul) >> 32) == y / p, "launch index split by reciprocal");
  }
  // 64 x 63-bit products of the Montgomery reduction and the key inner product
  for (int i = 0; i < 2000000; ++i) {
    const uint64_t a = (i % 5 == 0) ? ~0ULL - (rng() & 0xff) : rng();
    const uint64_t b = (i % 7 == 0) ? (uint64_t{1} << 63) - 1 - (rng() & 0xff) : rng() >> 1;
    const U128 p = mul_wide63(a, b);
    const u128 w = (u128)a * b;
    check(p.lo == (uint64_t)w && p.hi == (uint64_t)(w >> 64), "mul_wide63");
  }
  // Barrett variable x variable products: lazy form below 2^62, exact form below 2^63
  for (int i = 0; i < 3000000; ++i) {
    uint64_t q = (rng() >> (1 + i % 40)) | 1;
    if (i % 11 == 0) q = (uint64_t{1} << 63) - 1 - 2 * (rng() & 0xffff);
    if (i % 13 == 0) q = (uint64_t{1} << 62) - 1 - 2 * (rng() & 0xffff);
    if (q < 3) q = 3;
    const u128 mu = ~(u128)0 / q;
    const uint64_t ml = (uint64_t)mu, mh = (uint64_t)(mu >> 64);
    const uint64_t a = (i % 17 == 0) ? q - 1 : rng() % q, b = (i % 19 == 0) ? q - 1 : rng() % q;
    const uint64_t want = (uint64_t)(((u128)a * b) % q);
    check(barrett_mul_exact(a, b, q, ml, mh) == want, "barrett_mul_exact");
    if (q < (uint64_t{1} << 62)) check(barrett_mul_lazy(a, b, q, ml, mh) == want, "barrett_mul_lazy");
    check(barrett_mul(a, b, q, ml, mh) == want, "barrett_mul");
  }
  // narrow towers (q < 2^30): 32-bit Shoup with the high word of the 64-bit
  // companion, Harvey ranges of ct32 / gs32, canonicalisation
  const uint32_t narrow_primes[] = {17, 7681, 12289, 786433, 998244353, 1073479681, 1073741789, 1073707009};
  for (int i = 0; i < 3000000; ++i) {
    const uint32_t q = (i % 4 == 0) ? narrow_primes[i % 8] : (uint32_t)((rng() >> 34) | 1) % ((1u << 30) - 3) + 3;
    const uint32_t w = (uint32_t)(rng() % q);
    const uint64_t wp64 = (uint64_t)(((u128)w << 64) / q);
    const uint32_t wp = (uint32_t)(wp64 >> 32);
    check(wp == (uint32_t)(((uint64_t)w << 32) / q), "narrow companion = high word of the 64-bit companion");
    const Narrow m = narrow_of(q);
    const uint32_t y = (i % 3 == 0) ? (uint32_t)rng() : (uint32_t)(rng() % (4ull * q));  // any 32-bit input
    const uint32_t t = shoup32(y, w, wp, q);
    check(t < 2 * q && t % q == (uint32_t)(((uint64_t)y * w) % q), "shoup32 in [0, 2q) and congruent");
    const uint32_t fx = (uint32_t)(rng() % (4ull * q)), fy = (uint32_t)(rng() % (4ull * q));
    uint32_t cx = fx, cy = fy;
    ct32(cx, cy, w, wp, m);
    const uint64_t prod = ((uint64_t)(fy % q) * w) % q;
    check(cx < 4 * q && cy < 4 * q, "ct32 outputs below 4q");
    check(cx % q == (fx % q + prod) % q && cy % q == (fx % q + q - prod) % q, "ct32 congruent");
    const uint32_t ix = (uint32_t)(rng() % (2ull * q)), iy = (uint32_t)(rng() % (2ull * q));
    uint32_t gx = ix, gy = iy;
    gs32(gx, gy, w, wp, m);
    check(gx < 2 * q && gy < 2 * q, "gs32 outputs below 2q");
    check(gx % q == (ix % q + iy % q) % q && gy % q == ((uint64_t)((ix % q + q - iy % q) % q) * w) % q,
          "gs32 congruent");
    check(canon4_32(fx, m) == fx % q, "canon4_32");
    const uint32_t c = (uint32_t)(rng() >> 33) | 1, x = (uint32_t)(rng() % (2ull * c));
    check(min_sub(x, c) == (x >= c ? x - c : x), "min_sub");
  }
  for (const uint32_t q : narrow_primes)
    for (const uint32_t x : {0u, q - 1, q, 2 * q - 1, 4 * q - 1}) {
      check(canon4_32(x, narrow_of(q)) == x % q, "canon4_32 at the edges");
      for (const uint32_t w : {0u, 1u, q - 1}) {
        const uint32_t wp = (uint32_t)(((uint64_t)w << 32) / q);
        check(shoup32(x, w, wp, q) % q == (uint32_t)(((uint64_t)x * w) % q) && shoup32(x, w, wp, q) < 2 * q,
              "shoup32 at the edges");
        check(shoup32(0xffffffffu, w, wp, q) < 2 * q, "shoup32 on the largest input");
      }
    }
  std::printf("%zu ARITH CHECKS PASSED\n", checks);
  return 0;
}
