#pragma once
// Synchronous host-vector NTT on the GPU (reference include/ntt/cuda_ntt.hpp:34-79
// interface): limb-major host vectors of limbs()*n canonical residues; every
// call uploads, runs the whole transform on device and downloads. The result
// equals NegacyclicNtt on every limb, bit for bit. Free of HIP headers: the
// implementation calls the C ABI of include/ntt_hip.h.
#include <cstddef>
#include <cstdint>
#include <memory>
#include <span>
#include <stdexcept>
#include <vector>

#include "ntt/ntt.hpp"
#include "ntt/param_gen.hpp"

namespace ntt::hip {

namespace detail {
// Host preflight before any device call (reference include/ntt/cuda_validation.hpp:17-48).
struct Shape {
  size_t values, twiddles;
};
inline Shape validate_parameters(uint64_t n, std::span<const uint64_t> q, std::span<const uint64_t> roots) {
  if (n < 2 || n > 65536 || (n & (n - 1))) throw std::invalid_argument("HIP NTT requires power-of-two N in [2,65536]");
  if (q.empty() || q.size() != roots.size() || q.size() > 65535)
    throw std::invalid_argument("one (q, psi) pair per limb; at most 65535 limbs");
  for (size_t k = 0; k < q.size(); ++k) {
    if (q[k] <= n || q[k] >= (uint64_t{1} << 63) || q[k] % (2 * n) != 1 || !::ntt::detail::is_prime_mr(q[k]))
      throw std::invalid_argument("HIP NTT requires prime q < 2^63 with 2N dividing q-1");
    if (roots[k] == 0 || roots[k] >= q[k] || pow_mod(roots[k], n, q[k]) != q[k] - 1)
      throw std::invalid_argument("HIP NTT requires a canonical primitive 2N-th root");
  }
  return {q.size() * static_cast<size_t>(n), q.size() * static_cast<size_t>(n + 1)};
}
inline void validate_values(uint64_t n, std::span<const uint64_t> q, std::span<const uint64_t> values) {
  if (values.size() != q.size() * n) throw std::length_error("HIP buffer must have limbs*N elements");
  for (size_t k = 0; k < q.size(); ++k)
    for (size_t i = 0; i < n; ++i)
      if (values[k * n + i] >= q[k]) throw std::invalid_argument("HIP requires canonical residues");
}
}  // namespace detail

class HipNttTransform {
 public:
  HipNttTransform(uint64_t transform_length, const std::vector<uint64_t> &moduli,
                  const std::vector<uint64_t> &primitive_roots, int device = 0);
  ~HipNttTransform();
  HipNttTransform(const HipNttTransform &) = delete;
  HipNttTransform &operator=(const HipNttTransform &) = delete;

  uint64_t n() const;
  uint64_t limbs() const;
  uint64_t q(uint64_t limb) const;
  const ntt::NegacyclicNtt &ref(uint64_t limb) const;

  void forward(std::vector<uint64_t> &values);
  void inverse(std::vector<uint64_t> &values);
  void pointwise_multiply(std::vector<uint64_t> &left, const std::vector<uint64_t> &right);
  std::vector<uint64_t> multiply(std::vector<uint64_t> left, const std::vector<uint64_t> &right);

 private:
  struct Impl;
  std::unique_ptr<Impl> impl_;
};

}  // namespace ntt::hip
