#pragma once
// Shoup multiplication by a fixed constant (reference include/ntt/shoup.hpp:7-26).
#include "ntt/mod_arith.hpp"

namespace ntt {

// floor(w * 2^64 / q) for canonical w and 1 < q < 2^63.
inline uint64_t shoup_precompute(uint64_t w, uint64_t q) {
  if (q <= 1 || q >= (uint64_t{1} << 63) || w >= q)
    throw std::invalid_argument("Shoup requires canonical constant and 1 < q < 2^63");
  return static_cast<uint64_t>((static_cast<unsigned __int128>(w) << 64) / q);
}

// x * w mod q for x < q: the quotient estimate is at most one short, so one
// conditional subtraction makes the residue canonical.
NTT_HD uint64_t shoup_multiply(uint64_t x, uint64_t w, uint64_t wp, uint64_t q) {
  const uint64_t est = static_cast<uint64_t>((static_cast<unsigned __int128>(x) * wp) >> 64);
  const uint64_t r = x * w - est * q;
  return r >= q ? r - q : r;
}

}  // namespace ntt
