// The include/ntt C++ API over the gfx950 transport (needs a GPU):
//  * make_hip_session / make_cuda_session preflight (reference
//    tests/resident_cuda_preflight.cpp:1-21) before any device call;
//  * the resident chain, ModUp/ModDown and a full Hybrid key switch through
//    the Session on the device, bit-exact against the oracle;
//  * the synchronous transform class on the pinned OpenFHE towers with the
//    reference's seven patterns (tests/cuda_openfhe_tests.cpp:61-106),
//    compared with the host NegacyclicNtt and the oracle.
#include <array>
#include <iostream>

#include "ntt/cuda_ntt.hpp"
#include "ntt/resident.hpp"
#include "support/oracle_api.h"

using namespace ntt::resident;
using namespace ntt::proposal;

namespace {
size_t checks = 0;
void check(bool ok, const std::string &why) {
  ++checks;
  if (!ok) throw std::runtime_error(why);
}
// Equality of residue arrays; the failure names the count and first differing
// (tower, index) so a device mismatch says which stage and tower went wrong.
// `got` is a download's span: it must be compared while that download lives,
// so callers pass `s.download(x).wait_values()` inside this call's expression.
void same(std::span<const uint64_t> got, const std::vector<uint64_t> &want, size_t n, const std::string &why) {
  size_t bad = 0, first = 0;
  for (size_t i = 0; i < std::min(got.size(), want.size()); ++i)
    if (got[i] != want[i] && bad++ == 0) first = i;
  if (got.size() == want.size() && bad == 0) {
    ++checks;
    return;
  }
  std::string detail = " (" + std::to_string(bad) + " of " + std::to_string(want.size()) + " differ";
  if (bad)
    detail += "; first at tower " + std::to_string(first / n) + " index " + std::to_string(first % n) + ": got " +
              std::to_string(got[first]) + " want " + std::to_string(want[first]);
  if (got.size() != want.size()) detail += "; size " + std::to_string(got.size()) + " vs " + std::to_string(want.size());
  check(false, why + detail + ")");
}

template <class F>
void rejects(F f, const char *why) {
  ++checks;
  try {
    f();
  } catch (const ContractError &) {
    return;
  } catch (const std::invalid_argument &) {
    return;
  }
  throw std::runtime_error(why);
}

const std::array<uint64_t, 8> kQ = {1152921504606584833ULL, 1125899902124033ULL,    1125899911168001ULL,
                                    1125899903827969ULL,    1125899908022273ULL,    1152921504598720513ULL,
                                    1152921504597016577ULL, 1152921504595968001ULL};
const std::array<uint64_t, 8> kPsi = {18043022392882ULL, 15995550378ULL,  21670117260ULL,    938640682ULL,
                                      6511788946ULL,     800790938143ULL, 17749908910371ULL, 11469071954203ULL};
constexpr size_t N = 65536;

ParameterHandle pinned(size_t alpha = 3) {
  std::vector<Tower> q, p;
  for (size_t i = 0; i < 5; ++i) q.push_back({kQ[i], kPsi[i]});
  for (size_t i = 5; i < 8; ++i) p.push_back({kQ[i], kPsi[i]});
  return Parameters::create(N, q, p, alpha);
}

uint64_t lcg(uint64_t &s) { return s = s * 6364136223846793005ULL + 1442695040888963407ULL; }

std::vector<uint64_t> seeded(const PolynomialLayout &l, uint64_t seed) {
  std::vector<uint64_t> v(bytes(l) / 8);
  const size_t n = l.parameters->n();
  for (size_t t = 0; t < l.towers.size(); ++t)
    for (size_t i = 0; i < n; ++i) v[t * n + i] = lcg(seed) % l.parameters->tower(l.towers[t]).q;
  return v;
}

void preflight() {
  auto params = pinned();
  const PolynomialLayout base{params, 5, qp_basis(*params, 5), Domain::Coefficient, Order::Natural};
  auto eval = base;
  eval.domain = Domain::Evaluation;
  eval.order = Order::BitReversed;
  rejects([&] { make_hip_session(eval, 0); }, "evaluation base rejected");
  rejects([&] { make_cuda_session(base, -1); }, "negative device rejected");
  rejects([&] { make_hip_session(base, 0, {1 << 20, 1 << 20, 16}); }, "tables over budget rejected");
  rejects([&] { make_hip_session(base, 0, {~size_t{0}, ~size_t{0}, 0}); }, "empty operation budget rejected");
  rejects([] { ntt::cuda::CudaNttTransform(3, {kQ[0]}, {kPsi[0]}); }, "non power of two");
  rejects([] { ntt::cuda::CudaNttTransform(N, {kQ[0]}, {kPsi[0] + 1}); }, "bad root");
}

void session_chain_and_conversion() {
  auto params = pinned();
  const PolynomialLayout base{params, 5, qp_basis(*params, 5), Domain::Coefficient, Order::Natural};
  PolynomialLayout q = base;
  q.towers = q_prefix(*params, 5);
  Session s = make_cuda_session(base, 0, {size_t{4} << 30, size_t{1} << 30, 1024});
  const auto a = seeded(q, 1), b = seeded(q, 2);
  // ring product through the device chain == oracle transforms + pointwise
  auto da = s.upload(q, a), db = s.upload(q, b);
  auto prod = s.download(s.inverse(s.pointwise(s.forward(da), s.forward(db))));
  std::vector<uint64_t> qm, qr, pm, pr;
  for (size_t i = 0; i < 5; ++i) qm.push_back(kQ[i]), qr.push_back(kPsi[i]);
  for (size_t i = 5; i < 8; ++i) pm.push_back(kQ[i]), pr.push_back(kPsi[i]);
  auto fa = a, fb = b;
  or_ntt_towers(fa.data(), N, 5, qm.data(), qr.data(), 0);
  or_ntt_towers(fb.data(), N, 5, qm.data(), qr.data(), 0);
  for (size_t i = 0; i < fa.size(); ++i) fa[i] = ntt::mul_mod(fa[i], fb[i], kQ[i / N]);
  or_ntt_towers(fa.data(), N, 5, qm.data(), qr.data(), 1);
  same(prod.wait_values(), fa, N, "device ring product equals the oracle");
  // ModUp stages one by one: Q -> P basis switch of the coefficients, then the
  // P-tower forward transform (towers 5..7 of the session basis)
  PolynomialLayout p_coeff = base;
  p_coeff.towers.erase(p_coeff.towers.begin(), p_coeff.towers.begin() + 5);
  std::vector<uint64_t> want_p(3 * N);
  check(or_basis_switch(a.data(), N, qm.data(), 5, pm.data(), 3, want_p.data()) == 0, "oracle basis switch ran");
  const auto dp = s.switch_basis(da, p_coeff);
  same(s.download(dp).wait_values(), want_p, N, "device Q->P basis switch equals the oracle");
  or_ntt_towers(want_p.data(), N, 3, pm.data(), pr.data(), 0);
  same(s.download(s.forward(dp)).wait_values(), want_p, N, "device P-tower forward equals the oracle");
  // ModUp (coefficient input) then ModDown on the device
  auto up = s.mod_up(da);
  std::vector<uint64_t> want_up(8 * N), want_down(5 * N);
  or_mod_up(a.data(), 0, N, qm.data(), qr.data(), 5, pm.data(), pr.data(), 3, want_up.data());
  same(s.download(up).wait_values(), want_up, N, "device ModUp equals the oracle");
  or_mod_down(want_up.data(), N, qm.data(), qr.data(), 5, pm.data(), pr.data(), 3, want_down.data());
  same(s.download(s.mod_down(up)).wait_values(), want_down, N, "device ModDown equals the oracle");
  check(s.stats().uploads == 2 && s.stats().downloads == 5, "no hidden transfers");
  s.close();
}

void session_key_switch(size_t L) {
  auto params = pinned();
  const size_t M = 5, K = 3, alpha = 3, T = M + K, digits_full = 2;
  const PolynomialLayout base{params, L, qp_basis(*params, L), Domain::Coefficient, Order::Natural};
  const PolynomialLayout full{params, M, qp_basis(*params, M), Domain::Evaluation, Order::BitReversed};
  const PolynomialLayout source{params, L, q_prefix(*params, L), Domain::Evaluation, Order::BitReversed};
  std::vector<uint64_t> kb(digits_full * T * N), ka(kb.size());
  uint64_t s1 = 0xbbb, s2 = 0xaaa;
  for (size_t i = 0; i < kb.size(); ++i) {
    const uint64_t m = kQ[(i / N) % T];
    kb[i] = lcg(s1) % m;
    ka[i] = lcg(s2) % m;
  }
  const auto c = seeded(source, 77 + L);
  Session s = make_hip_session(base, 0, {size_t{8} << 30, size_t{2} << 30, 1024});
  std::vector<HybridKeyDigit> rows;
  for (size_t g = 0; g < digits_full; ++g)
    rows.push_back({std::span<const uint64_t>(kb).subspan(g * T * N, T * N),
                    std::span<const uint64_t>(ka).subspan(g * T * N, T * N)});
  const auto key = s.upload_hybrid_key(full, rows);
  const auto z = s.hybrid_accumulate(s.hybrid_digits(s.upload(source, c)), key);
  const auto ob = s.download(s.mod_down(z.b)), oa = s.download(s.mod_down(z.a));
  std::vector<uint64_t> qm(kQ.begin(), kQ.begin() + L), qr(kPsi.begin(), kPsi.begin() + L);
  std::vector<uint64_t> pm(kQ.begin() + 5, kQ.end()), pr(kPsi.begin() + 5, kPsi.end());
  std::vector<uint64_t> wb(L * N), wa(L * N);
  check(or_key_switch(c.data(), N, qm.data(), qr.data(), L, M, pm.data(), pr.data(), K, alpha, kb.data(), ka.data(),
                      wb.data(), wa.data()) == 0,
        "oracle ran");
  same(ob.wait_values(), wb, N, "device Session key switch B equals the oracle, L=" + std::to_string(L));
  same(oa.wait_values(), wa, N, "device Session key switch A equals the oracle, L=" + std::to_string(L));
  // Two more accumulations queued back to back (the driver merges each one's
  // per-digit MACs into one pass): a second input, then the first input again.
  const auto c2 = seeded(source, 177 + L);
  const auto z2 = s.hybrid_accumulate(s.hybrid_digits(s.upload(source, c2)), key);
  const auto z3 = s.hybrid_accumulate(s.hybrid_digits(s.upload(source, c)), key);
  const auto ob2 = s.download(s.mod_down(z2.b)), oa2 = s.download(s.mod_down(z2.a));
  const auto ob3 = s.download(s.mod_down(z3.b)), oa3 = s.download(s.mod_down(z3.a));
  std::vector<uint64_t> wb2(L * N), wa2(L * N);
  check(or_key_switch(c2.data(), N, qm.data(), qr.data(), L, M, pm.data(), pr.data(), K, alpha, kb.data(), ka.data(),
                      wb2.data(), wa2.data()) == 0,
        "oracle ran");
  same(ob2.wait_values(), wb2, N, "second Session key switch B equals the oracle");
  same(oa2.wait_values(), wa2, N, "second Session key switch A equals the oracle");
  same(ob3.wait_values(), wb, N, "repeated Session key switch B unchanged");
  same(oa3.wait_values(), wa, N, "repeated Session key switch A unchanged");
  s.close();
}

void transform_class_patterns() {
  for (const std::vector<size_t> &sel :
       {std::vector<size_t>{0, 1, 2, 3, 4}, std::vector<size_t>{0, 1, 2, 5, 6, 7}, std::vector<size_t>{0}}) {
    std::vector<uint64_t> q, r;
    for (size_t i : sel) q.push_back(kQ[i]), r.push_back(kPsi[i]);
    ntt::cuda::CudaNttTransform gpu(N, q, r);
    for (unsigned pattern = 0; pattern < 7; ++pattern) {
      std::vector<uint64_t> in(N * q.size());
      uint64_t state = 0x7251d9;
      for (size_t k = 0; k < q.size(); ++k)
        for (size_t i = 0; i < N; ++i) {
          const uint64_t m = q[k];
          const uint64_t v = lcg(state) % m;
          in[k * N + i] = pattern == 0 ? 0 : pattern == 1 ? 1 : pattern == 2 ? m - 1 : pattern == 3 ? (i % 2 ? m - 1 : 0)
                        : pattern == 4 ? (i == 17) : pattern == 5 ? (i == N - 1) : v;
        }
      auto want = in;
      for (size_t k = 0; k < q.size(); ++k) {
        std::vector<uint64_t> row(in.begin() + k * N, in.begin() + (k + 1) * N);
        gpu.ref(k).forward(row);
        std::copy(row.begin(), row.end(), want.begin() + k * N);
      }
      auto oracle = in;
      or_ntt_towers(oracle.data(), N, q.size(), q.data(), r.data(), 0);
      check(want == oracle, "host NegacyclicNtt equals the oracle");
      auto v = in;
      gpu.forward(v);
      check(v == want, "GPU forward, pattern " + std::to_string(pattern));
      gpu.inverse(v);
      check(v == in, "GPU round trip, pattern " + std::to_string(pattern));
      auto sq = want;
      gpu.pointwise_multiply(sq, want);
      gpu.inverse(sq);
      check(gpu.multiply(in, in) == sq, "fused ring product equals staged product");
    }
  }
}
// The synchronous class on 16 towers (over 2^20 residues: the parallel
// check-and-stage path): bit-exact forward / inverse against the host
// reference, and the contract errors (noncanonical residue, wrong size,
// a bad right operand) thrown with the caller's vectors untouched.
void transform_class_staging_and_rejections() {
  const auto q = ntt::find_primes(N, 16, 60);
  std::vector<uint64_t> r;
  for (uint64_t m : q) r.push_back(ntt::find_psi(m, N));
  ntt::cuda::CudaNttTransform gpu(N, q, r);
  std::vector<uint64_t> in(N * q.size());
  uint64_t state = 0x5eed;
  for (size_t i = 0; i < in.size(); ++i) in[i] = lcg(state) % q[i / N];
  auto v = in;
  gpu.forward(v);
  for (size_t k : {size_t{0}, size_t{7}, size_t{15}}) {
    std::vector<uint64_t> row(in.begin() + k * N, in.begin() + (k + 1) * N);
    gpu.ref(k).forward(row);
    check(std::equal(row.begin(), row.end(), v.begin() + k * N), "staged forward, tower " + std::to_string(k));
  }
  gpu.inverse(v);
  check(v == in, "staged round trip");
  auto bad = in;
  bad.back() = q.back();  // q itself is not a canonical residue
  auto left = in;
  rejects([&] { gpu.forward(bad); }, "noncanonical residue in the last tower");
  check(bad.back() == q.back(), "rejected vector untouched");
  rejects([&] { gpu.pointwise_multiply(left, bad); }, "noncanonical right operand");
  check(left == in, "left operand untouched when the right one is rejected");
  std::vector<uint64_t> shorter(in.begin(), in.end() - 1);
  bool length = false;
  try {
    gpu.forward(shorter);
  } catch (const std::length_error &) {
    length = true;
  }
  check(length, "wrong length rejected");
  auto again = in;
  gpu.forward(again);  // the transform still works after the rejections
  gpu.inverse(again);
  check(again == in, "round trip after rejections");
}
}  // namespace

int main() {
  try {
    preflight();
    session_chain_and_conversion();
    session_key_switch(5);
    session_key_switch(4);
    transform_class_patterns();
    transform_class_staging_and_rejections();
    std::cout << checks << " HIP RESIDENT CHECKS PASSED (device)\n";
  } catch (const std::exception &e) {
    std::cerr << "FAIL: " << e.what() << '\n';
    return 1;
  }
  return 0;
}
