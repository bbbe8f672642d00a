#pragma once
// Host negacyclic NTT over Z_q[x]/(x^N + 1) (reference include/ntt/ntt.hpp:53-214
// interface). Same convention: psi of order 2N (psi^N = -1); forward maps
// natural-order coefficients to A_k = a(psi^(2 rev(k) + 1)), stored in
// bit-reversed order; inverse is its exact inverse including N^-1.
//
// This is the host-side API class (used by callers that want a CPU
// transform and by CudaNttTransform::ref). Our implementation uses the
// bit-reversed power tables and one multiply per butterfly; device work
// never routes through it.
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "ntt/bitrev.hpp"
#include "ntt/mod_arith.hpp"

namespace ntt {

class NegacyclicNtt {
 public:
  NegacyclicNtt(uint64_t n, uint64_t q, uint64_t psi) : n_(n), q_(q), psi_(psi) {
    if (n < 2 || (n & (n - 1)) != 0) throw std::invalid_argument("n must be a power of two >= 2");
    if (q < 2 || (q - 1) % (2 * n) != 0) throw std::invalid_argument("q must be prime with (2n) | (q-1)");
    if (pow_mod(psi, n, q) != q - 1) throw std::invalid_argument("psi^n != -1 mod q (psi must have order 2n)");
    if (q <= n) throw std::invalid_argument("require q > n for N^-1 to exist");
    n_inv_ = inv_mod(n % q, q);
    fwd_ = table(psi);
    inv_ = table(inv_mod(psi, q));
  }
  uint64_t n() const { return n_; }
  uint64_t q() const { return q_; }

  void forward(std::vector<uint64_t> &a) const {
    check_size(a);
    for (uint64_t m = 1; m < n_; m <<= 1) {
      const uint64_t d = n_ / (2 * m);
      for (uint64_t g = 0; g < m; ++g) {
        const uint64_t w = fwd_[m + g];
        for (uint64_t i = 2 * g * d; i < 2 * g * d + d; ++i) {
          const uint64_t x = a[i], y = mul_mod(a[i + d], w, q_);
          a[i] = add_mod(x, y, q_);
          a[i + d] = sub_mod(x, y, q_);
        }
      }
    }
  }

  void inverse(std::vector<uint64_t> &a) const {
    check_size(a);
    for (uint64_t m = n_ / 2; m >= 1; m >>= 1) {
      const uint64_t d = n_ / (2 * m);
      for (uint64_t g = 0; g < m; ++g) {
        const uint64_t w = inv_[m + g];
        for (uint64_t i = 2 * g * d; i < 2 * g * d + d; ++i) {
          const uint64_t x = a[i], y = a[i + d];
          a[i] = add_mod(x, y, q_);
          a[i + d] = mul_mod(sub_mod(x, y, q_), w, q_);
        }
      }
    }
    for (auto &v : a) v = mul_mod(v, n_inv_, q_);
  }

 private:
  std::vector<uint64_t> table(uint64_t root) const {
    std::vector<uint64_t> t(n_);
    const uint64_t lg = log2_of_pow2(n_);
    uint64_t power = 1;
    for (uint64_t e = 0; e < n_; ++e, power = mul_mod(power, root, q_)) t[bit_reverse(e, lg)] = power;
    return t;
  }
  void check_size(const std::vector<uint64_t> &a) const {
    if (a.size() != n_)
      throw std::length_error("vector size " + std::to_string(a.size()) + " != n " + std::to_string(n_));
  }
  uint64_t n_, q_, psi_, n_inv_ = 0;
  std::vector<uint64_t> fwd_, inv_;
};

// Ring product in Z_q[x]/(x^N+1) (reference ntt.hpp:205-214).
inline std::vector<uint64_t> ntt_multiply(const NegacyclicNtt &t, std::vector<uint64_t> a,
                                          std::vector<uint64_t> b) {
  t.forward(a);
  t.forward(b);
  for (size_t i = 0; i < a.size(); ++i) a[i] = mul_mod(a[i], b[i], t.q());
  t.inverse(a);
  return a;
}

}  // namespace ntt
