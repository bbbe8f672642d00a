#pragma once
// Host parameter search (reference include/ntt/param_gen.hpp:25-96 interface):
// Miller-Rabin with 12 fixed bases, ascending prime walk q = 1 mod 2n,
// psi of order exactly 2n, plus OpenFHE's minimum-root convention.
#include <cstdint>
#include <stdexcept>
#include <vector>

#include "ntt/mod_arith.hpp"

namespace ntt {
namespace detail {
inline bool is_prime_mr(uint64_t n) {
  static constexpr uint64_t kBases[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
  if (n < 2) return false;
  for (uint64_t p : kBases)
    if (n % p == 0) return n == p;
  uint64_t d = n - 1, s = 0;
  while ((d & 1) == 0) d >>= 1, ++s;
  for (uint64_t a : kBases) {
    uint64_t x = pow_mod(a, d, n);
    if (x == 1 || x == n - 1) continue;
    bool composite = true;
    for (uint64_t r = 1; r < s && composite; ++r) {
      x = mul_mod(x, x, n);
      composite = x != n - 1;
    }
    if (composite) return false;
  }
  return true;
}
}  // namespace detail

inline std::vector<uint64_t> find_primes(uint64_t n, uint64_t count, uint64_t bits = 46) {
  const uint64_t step = 2 * n;
  std::vector<uint64_t> out;
  for (uint64_t q = ((uint64_t{1} << (bits - 1)) / step + 1) * step + 1; out.size() < count; q += step) {
    if (q >= (uint64_t{1} << bits)) throw std::runtime_error("prime search exhausted the bit range");
    if (detail::is_prime_mr(q)) out.push_back(q);
  }
  return out;
}

inline uint64_t find_psi(uint64_t q, uint64_t n) {
  uint64_t s = (q ^ n) ? (q ^ n) : 1;
  for (;;) {
    s ^= s << 13;
    s ^= s >> 7;
    s ^= s << 17;
    const uint64_t zeta = pow_mod(2 + s % (q - 3), (q - 1) / (2 * n), q);
    if (pow_mod(zeta, n, q) == q - 1) return zeta;
  }
}

// OpenFHE's root: the smallest primitive 2n-th root, min over psi^(2j+1).
inline uint64_t min_root(uint64_t q, uint64_t n) {
  const uint64_t psi = find_psi(q, n), square = mul_mod(psi, psi, q);
  uint64_t best = psi, power = psi;
  for (uint64_t j = 0; j < n; ++j, power = mul_mod(power, square, q))
    if (power < best) best = power;
  return best;
}

}  // namespace ntt
