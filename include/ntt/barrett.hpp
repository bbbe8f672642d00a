#pragma once
// Barrett multiplication with a 128-bit reciprocal (reference include/ntt/barrett.hpp:6-49).
#include "ntt/mod_arith.hpp"

namespace ntt {

struct BarrettModulus {
  uint64_t modulus, reciprocal_low, reciprocal_high;
};

// mu = floor(2^128 / q), 1 < q < 2^63.
inline BarrettModulus barrett_precompute(uint64_t q) {
  if (q <= 1 || q >= (uint64_t{1} << 63)) throw std::invalid_argument("Barrett requires 1 < q < 2^63");
  const unsigned __int128 all = ~static_cast<unsigned __int128>(0);
  const unsigned __int128 mu = all / q + (all % q == q - 1 ? 1 : 0);
  return {q, static_cast<uint64_t>(mu), static_cast<uint64_t>(mu >> 64)};
}

namespace barrett_detail {
NTT_HD uint64_t high_product(uint64_t a, uint64_t b) {
  return static_cast<uint64_t>((static_cast<unsigned __int128>(a) * b) >> 64);
}
// Low limb of floor((high:low) * mu / 2^128); for a*b < q^2 the quotient fits one limb.
NTT_HD uint64_t quotient(uint64_t low, uint64_t high, BarrettModulus m) {
  const uint64_t cross = low * m.reciprocal_high + high * m.reciprocal_low;
  const uint64_t first = low * m.reciprocal_high;
  const uint64_t mid = cross + high_product(low, m.reciprocal_low);
  const uint64_t carry = static_cast<uint64_t>(cross < first) + static_cast<uint64_t>(mid < cross);
  return high * m.reciprocal_high + high_product(low, m.reciprocal_high) + high_product(high, m.reciprocal_low) +
         carry;
}
}  // namespace barrett_detail

NTT_HD uint64_t barrett_multiply(uint64_t a, uint64_t b, BarrettModulus m) {
  const uint64_t low = a * b;
  const uint64_t est = barrett_detail::quotient(low, barrett_detail::high_product(a, b), m);
  const uint64_t r = low - est * m.modulus;
  return r >= m.modulus ? r - m.modulus : r;
}

}  // namespace ntt
