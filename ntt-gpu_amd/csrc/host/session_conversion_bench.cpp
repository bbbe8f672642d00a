// The reference's resident conversion benchmark (src/resident_conversion_bench.cpp:97-300)
// on the HIP transport, through the include/ntt Session: ModUp, ModDown and
// ModUp followed by ModDown with the input already on the device (host wall
// around the operation and its readiness wait), and the same pair at the host
// boundary (upload, both operations, download). Q = the vector file's primes
// (at most 40, each below 2^60), P = the first three 60-bit primes of the
// reference generator that are not in Q, three towers per digit; inputs from
// the reference's fixed-seed operand stream. Every sample's result is checked
// outside the timer against a CPU computation with the include/ntt reference
// NTT and approximate basis conversion.
//
// Usage: session_conversion_bench <parameter-vector> <warmups> <samples>
// stdout: JSONL (workload, sample and summary lines per operation, complete).
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <functional>
#include <iostream>
#include <string>
#include <vector>

#include "ntt/basis_conversion.hpp"
#include "ntt/measurement.hpp"
#include "ntt/ntt.hpp"
#include "ntt/param_gen.hpp"
#include "ntt/resident.hpp"
#include "ntt/resident_conversion.hpp"
#include "ntt/vector_io.hpp"

using namespace ntt::resident;
using namespace ntt::proposal;
namespace ms = ntt::measurement;

namespace {

using Values = std::vector<uint64_t>;

Values tower_slice(const Values &v, size_t n, size_t first, size_t count) {
  return Values(v.begin() + first * n, v.begin() + (first + count) * n);
}

// CPU ModUp of coefficient Q residues: evaluation Q, then the converted and
// transformed P towers (Session::mod_up semantics).
Values cpu_mod_up(const Values &x, size_t n, const std::vector<Tower> &q, const std::vector<Tower> &p) {
  std::vector<uint64_t> qm, pm;
  for (const Tower &t : q) qm.push_back(t.q);
  for (const Tower &t : p) pm.push_back(t.q);
  Values out;
  for (size_t k = 0; k < q.size(); ++k) {
    Values t = tower_slice(x, n, k, 1);
    ntt::NegacyclicNtt(n, q[k].q, q[k].psi).forward(t);
    out.insert(out.end(), t.begin(), t.end());
  }
  const Values conv = ntt::ApproximateBasisPlan(qm, pm).convert(n, x);
  for (size_t k = 0; k < p.size(); ++k) {
    Values t = tower_slice(conv, n, k, 1);
    ntt::NegacyclicNtt(n, p[k].q, p[k].psi).forward(t);
    out.insert(out.end(), t.begin(), t.end());
  }
  return out;
}

// CPU ModDown of evaluation QP residues: (y_Q - NTT(conv(INTT(y_P)))) P^-1 mod q.
Values cpu_mod_down(const Values &y, size_t n, const std::vector<Tower> &q, const std::vector<Tower> &p,
                    const Parameters &params) {
  std::vector<uint64_t> qm, pm;
  for (const Tower &t : q) qm.push_back(t.q);
  for (const Tower &t : p) pm.push_back(t.q);
  Values pc;
  for (size_t k = 0; k < p.size(); ++k) {
    Values t = tower_slice(y, n, q.size() + k, 1);
    ntt::NegacyclicNtt(n, p[k].q, p[k].psi).inverse(t);
    pc.insert(pc.end(), t.begin(), t.end());
  }
  const Values conv = ntt::ApproximateBasisPlan(pm, qm).convert(n, pc);
  Values out(q.size() * n);
  for (size_t k = 0; k < q.size(); ++k) {
    Values t = tower_slice(conv, n, k, 1);
    ntt::NegacyclicNtt(n, q[k].q, q[k].psi).forward(t);
    const uint64_t m = q[k].q;
    const uint64_t pinv = ntt::inv_mod(detail::special_prime_product(params, m), m);
    for (size_t i = 0; i < n; ++i)
      out[k * n + i] = ntt::mul_mod(ntt::sub_mod(y[k * n + i], t[i], m), pinv, m);
  }
  return out;
}

double wall_ms(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

int main(int argc, char **argv) {
  try {
    if (argc != 4) throw std::invalid_argument("usage: session_conversion_bench <parameter-vector> <warmups> <samples>");
    const auto vector = ntt::read_vector_file(argv[1]);
    const size_t warmups = std::stoul(argv[2]), samples = std::stoul(argv[3]);
    ms::validate_counts(warmups, samples);
    require(vector.q.size() <= 40, "conversion benchmark supports at most 40 Q towers");
    for (uint64_t m : vector.q) require(m < (uint64_t{1} << 60), "conversion benchmark requires Q primes below 2^60");
    const size_t n = vector.n;
    std::vector<Tower> q, p;
    for (size_t k = 0; k < vector.q.size(); ++k) q.push_back({vector.q[k], vector.psi[k]});
    for (uint64_t m : ntt::find_primes(n, vector.q.size() + 3, 60))
      if (std::find(vector.q.begin(), vector.q.end(), m) == vector.q.end() && p.size() < 3)
        p.push_back({m, ntt::find_psi(m, n)});
    require(p.size() == 3, "three disjoint special primes required");
    const auto params = Parameters::create(n, q, p, 3);
    const size_t L = q.size();
    const PolynomialLayout full{params, L, qp_basis(*params, L), Domain::Coefficient, Order::Natural};
    const PolynomialLayout source{params, L, q_prefix(*params, L), Domain::Coefficient, Order::Natural};
    const PolynomialLayout source_eval{params, L, q_prefix(*params, L), Domain::Evaluation, Order::BitReversed};
    const PolynomialLayout full_eval{params, L, qp_basis(*params, L), Domain::Evaluation, Order::BitReversed};
    std::vector<uint64_t> qm = vector.q, all = vector.q;
    for (const Tower &t : p) all.push_back(t.q);
    const Values coefficients = ms::benchmark_inputs(n, qm).first;
    const Values arbitrary = ms::benchmark_inputs(n, all).first;
    Values evaluation;
    for (size_t k = 0; k < L; ++k) {
      Values t = tower_slice(coefficients, n, k, 1);
      ntt::NegacyclicNtt(n, q[k].q, q[k].psi).forward(t);
      evaluation.insert(evaluation.end(), t.begin(), t.end());
    }
    const Values expected_up = cpu_mod_up(coefficients, n, q, p);
    const Values expected_chain = cpu_mod_down(expected_up, n, q, p, *params);
    const Values expected_down = cpu_mod_down(arbitrary, n, q, p, *params);

    std::cout << "{\"type\":\"workload\",\"schema\":1,\"timing\":\"host_wall_resident_conversion\",\"n\":" << n
              << ",\"towers\":" << L << ",\"p_towers\":3,\"alpha\":3,\"warmups\":" << warmups
              << ",\"samples\":" << samples << ",\"seed\":" << ms::kInputSeed
              << ",\"input_policy\":\"fixed_seed_reset_outside_timer\",\"p_policy\":\"first_disjoint_60_bit_primes\""
              << ",\"transport\":\"hip\"}\n";
    Session s = make_hip_session(full, 0, {size_t{64} << 30, size_t{16} << 30, size_t{1} << 20});
    auto check = [](const std::vector<uint64_t> &got, const Values &want, const char *what) {
      if (got != want) throw std::runtime_error(std::string("result differs from the CPU reference: ") + what);
    };
    auto device_case = [&](const char *name, const PolynomialLayout &layout, const Values &input, const Values &want,
                           const std::function<Polynomial(const Polynomial &)> &op) {
      const Polynomial resident = s.upload(layout, input);
      resident.readiness().wait();
      std::vector<double> times;
      for (size_t i = 0; i < warmups + samples; ++i) {
        const auto t0 = std::chrono::steady_clock::now();
        const Polynomial result = op(resident);
        result.readiness().wait();
        const double ms_ = wall_ms(t0);
        const Download d = s.download(result);  // owns the staging the span points into
        const auto values = d.wait_values();
        check(std::vector<uint64_t>(values.begin(), values.end()), want, name);
        if (i >= warmups) times.push_back(ms_);
      }
      ms::write_samples(std::cout, name, times);
    };
    device_case("resident_modup_coefficient", source, coefficients, expected_up,
                [&](const Polynomial &x) { return s.mod_up(x); });
    device_case("resident_modup_evaluation", source_eval, evaluation, expected_up,
                [&](const Polynomial &x) { return s.mod_up(x); });
    device_case("resident_moddown_evaluation", full_eval, arbitrary, expected_down,
                [&](const Polynomial &x) { return s.mod_down(x); });
    device_case("resident_modup_moddown", source, coefficients, expected_chain,
                [&](const Polynomial &x) { return s.mod_down(s.mod_up(x)); });
    std::vector<double> times;
    for (size_t i = 0; i < warmups + samples; ++i) {
      const auto t0 = std::chrono::steady_clock::now();
      const Download d = s.download(s.mod_down(s.mod_up(s.upload(source, coefficients))));
      const auto values = d.wait_values();
      const double ms_ = wall_ms(t0);
      check(std::vector<uint64_t>(values.begin(), values.end()), expected_chain, "host boundary");
      if (i >= warmups) times.push_back(ms_);
    }
    ms::write_samples(std::cout, "resident_conversion_host_boundary", times);
    s.close();
    std::cout << "{\"type\":\"complete\",\"operations\":5}\n";
  } catch (const std::exception &e) {
    std::fprintf(stderr, "session conversion bench failed: %s\n", e.what());
    return 1;
  }
  return 0;
}
