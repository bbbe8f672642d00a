// KeySwitchHybrid through the include/ntt Session API on the HIP transport,
// one request at a time: the reference's own call graph
// (openfhe_hybrid_adapter.hpp:163-181 -> Session::hybrid_digits,
// hybrid_accumulate, mod_down x2; src/resident.cpp:448-570), with the input
// and the evaluation key resident and each request waited for before the
// next. It is the "plain" point of comparison for the fused batched path.
//
// stdin: "n L K alpha" then L Q moduli, L Q roots, K P moduli, K P roots.
// argv: [requests] [warmup]. stdout: one JSON line.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <iostream>
#include <vector>

#include "ntt/resident.hpp"

using namespace ntt::resident;
using namespace ntt::proposal;

namespace {
uint64_t lcg(uint64_t &s) { return s = s * 6364136223846793005ULL + 1442695040888963407ULL; }

double pct(std::vector<double> v, double q) {
  std::sort(v.begin(), v.end());
  const double pos = (v.size() - 1) * q;
  const size_t lo = (size_t)pos, hi = std::min(lo + 1, v.size() - 1);
  return v[lo] + (v[hi] - v[lo]) * (pos - lo);
}
}  // namespace

int main(int argc, char **argv) {
  try {
    const int requests = argc > 1 ? std::atoi(argv[1]) : 16;
    const int warmup = argc > 2 ? std::atoi(argv[2]) : 2;
    size_t n, L, K, alpha;
    if (!(std::cin >> n >> L >> K >> alpha)) throw std::runtime_error("expected n L K alpha on stdin");
    std::vector<uint64_t> q(L), qr(L), p(K), pr(K);
    for (auto *v : {&q, &qr, &p, &pr})
      for (auto &x : *v)
        if (!(std::cin >> x)) throw std::runtime_error("expected moduli and roots on stdin");
    std::vector<Tower> qt, pt;
    for (size_t i = 0; i < L; ++i) qt.push_back({q[i], qr[i]});
    for (size_t i = 0; i < K; ++i) pt.push_back({p[i], pr[i]});
    const auto params = Parameters::create(n, qt, pt, alpha);
    const PolynomialLayout base{params, L, qp_basis(*params, L), Domain::Coefficient, Order::Natural};
    const PolynomialLayout full{params, L, qp_basis(*params, L), Domain::Evaluation, Order::BitReversed};
    const PolynomialLayout source{params, L, q_prefix(*params, L), Domain::Evaluation, Order::BitReversed};
    const size_t T = L + K, digits = (L + alpha - 1) / alpha;
    std::vector<uint64_t> kb(digits * T * n), ka(kb.size()), c(L * n);
    uint64_t s1 = 0xbbb, s2 = 0xaaa, s3 = 0x7251d9;
    for (size_t i = 0; i < kb.size(); ++i) {
      const size_t t = (i / n) % T;
      const uint64_t m = t < L ? q[t] : p[t - L];
      kb[i] = lcg(s1) % m;
      ka[i] = lcg(s2) % m;
    }
    for (size_t i = 0; i < c.size(); ++i) c[i] = lcg(s3) % q[i / n];
    Session s = make_hip_session(base, 0, {size_t{64} << 30, size_t{16} << 30, size_t{1} << 20});
    std::vector<HybridKeyDigit> rows;
    for (size_t g = 0; g < digits; ++g)
      rows.push_back({std::span<const uint64_t>(kb).subspan(g * T * n, T * n),
                      std::span<const uint64_t>(ka).subspan(g * T * n, T * n)});
    const auto key = s.upload_hybrid_key(full, rows);
    const auto input = s.upload(source, c);
    auto one = [&] {
      const auto z = s.hybrid_accumulate(s.hybrid_digits(input), key);
      const auto b = s.mod_down(z.b), a = s.mod_down(z.a);
      b.readiness().wait();
      a.readiness().wait();
    };
    for (int i = 0; i < warmup; ++i) one();
    std::vector<double> lat;
    const auto t0 = std::chrono::steady_clock::now();
    for (int i = 0; i < requests; ++i) {
      const auto r0 = std::chrono::steady_clock::now();
      one();
      lat.push_back(std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - r0).count());
    }
    const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::printf("{\"keyswitch_per_s\": %.2f, \"p50_ms\": %.3f, \"p99_ms\": %.3f, \"requests\": %d}\n",
                requests / wall, pct(lat, 0.5), pct(lat, 0.99), requests);
    s.close();
  } catch (const std::exception &e) {
    std::fprintf(stderr, "session bench failed: %s\n", e.what());
    return 1;
  }
  return 0;
}
