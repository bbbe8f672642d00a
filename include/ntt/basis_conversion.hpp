#pragma once
// Coefficient-domain approximate CRT switch with the unreduced lift of pinned
// OpenFHE (reference include/ntt/basis_conversion.hpp:13-114 interface):
// out_m = [ sum_i [x_i * Bhat_i^-1]_{b_i} * Bhat_i ]_m, no reduction modulo B.
#include <algorithm>
#include <span>
#include <vector>

#include "ntt/param_gen.hpp"
#include "ntt/shoup.hpp"

namespace ntt {

// One output coefficient from source-major input and a source-major hat table.
inline uint64_t approximate_basis_value(const uint64_t *input, uint64_t dimension, uint64_t coefficient,
                                        uint64_t target_index, uint64_t source_count, const uint64_t *source_moduli,
                                        const uint64_t *target_moduli, uint64_t target_count,
                                        const uint64_t *inverses, const uint64_t *inverse_quotients,
                                        const uint64_t *source_hats_mod_target) {
  const uint64_t m = target_moduli[target_index];
  uint64_t sum = 0;
  for (uint64_t k = 0; k < source_count; ++k) {
    const uint64_t y = shoup_multiply(input[k * dimension + coefficient], inverses[k], inverse_quotients[k],
                                      source_moduli[k]);
    sum = add_mod(sum, mul_mod(y % m, source_hats_mod_target[k * target_count + target_index], m), m);
  }
  return sum;
}

class ApproximateBasisPlan {
  std::vector<uint64_t> source_, target_, inverses_, quotients_, hats_;

 public:
  ApproximateBasisPlan(std::vector<uint64_t> source, std::vector<uint64_t> target)
      : source_(std::move(source)), target_(std::move(target)) {
    if (source_.empty() || source_.size() > 40 || target_.empty() || target_.size() > 40)
      throw std::invalid_argument("basis switch requires 1..40 source and target primes");
    std::vector<uint64_t> all = source_;
    all.insert(all.end(), target_.begin(), target_.end());
    std::sort(all.begin(), all.end());
    if (std::adjacent_find(all.begin(), all.end()) != all.end())
      throw std::invalid_argument("basis switch requires distinct disjoint primes");
    for (uint64_t m : all)
      if (m < 3 || m >= (uint64_t{1} << 63) || !detail::is_prime_mr(m))
        throw std::invalid_argument("basis switch requires odd primes below 2^63");
    for (size_t k = 0; k < source_.size(); ++k) {
      const uint64_t b = source_[k];
      uint64_t hat = 1;
      for (size_t o = 0; o < source_.size(); ++o)
        if (o != k) hat = mul_mod(hat, source_[o] % b, b);
      inverses_.push_back(inv_mod(hat, b));
      quotients_.push_back(shoup_precompute(inverses_.back(), b));
      for (uint64_t m : target_) {
        uint64_t h = 1;
        for (size_t o = 0; o < source_.size(); ++o)
          if (o != k) h = mul_mod(h, source_[o] % m, m);
        hats_.push_back(h);
      }
    }
  }
  const std::vector<uint64_t> &source_moduli() const { return source_; }
  const std::vector<uint64_t> &target_moduli() const { return target_; }
  const std::vector<uint64_t> &inverses() const { return inverses_; }
  const std::vector<uint64_t> &inverse_quotients() const { return quotients_; }
  const std::vector<uint64_t> &source_hats_mod_target() const { return hats_; }

  void validate_input(uint64_t dimension, std::span<const uint64_t> input) const {
    if (dimension < 2 || dimension > 65536 || (dimension & (dimension - 1)) ||
        input.size() != dimension * source_.size())
      throw std::invalid_argument("basis switch requires power-of-two N in [2,65536] and source*N values");
    for (size_t k = 0; k < source_.size(); ++k)
      for (size_t i = 0; i < dimension; ++i)
        if (input[k * dimension + i] >= source_[k])
          throw std::invalid_argument("basis switch requires canonical input residues");
  }

  std::vector<uint64_t> convert(uint64_t dimension, std::span<const uint64_t> input) const {
    validate_input(dimension, input);
    std::vector<uint64_t> out(dimension * target_.size());
    for (size_t m = 0; m < target_.size(); ++m)
      for (size_t i = 0; i < dimension; ++i)
        out[m * dimension + i] = approximate_basis_value(input.data(), dimension, i, m, source_.size(),
                                                         source_.data(), target_.data(), target_.size(),
                                                         inverses_.data(), quotients_.data(), hats_.data());
    return out;
  }
};

}  // namespace ntt
