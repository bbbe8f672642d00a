#pragma once
// Bit reversal (reference include/ntt/bitrev.hpp:23-64 interface).
#include <cstdint>
#include <utility>
#include <vector>

#ifndef NTT_HD
#define NTT_HD inline
#endif

namespace ntt {

NTT_HD uint64_t log2_of_pow2(uint64_t n) {
  uint64_t lg = 0;
  while ((uint64_t{1} << lg) < n) ++lg;
  return lg;
}

NTT_HD uint64_t bit_reverse(uint64_t x, uint64_t lg) {
  uint64_t r = 0;
  for (uint64_t i = 0; i < lg; ++i, x >>= 1) r = (r << 1) | (x & 1);
  return r;
}

// In-place permutation v[i] <-> v[bitrev(i)]; each pair swapped once.
inline void bit_reverse_permute(std::vector<uint64_t> &v) {
  const uint64_t lg = log2_of_pow2(v.size());
  for (uint64_t i = 0; i < v.size(); ++i) {
    const uint64_t j = bit_reverse(i, lg);
    if (i < j) std::swap(v[i], v[j]);
  }
}

}  // namespace ntt
