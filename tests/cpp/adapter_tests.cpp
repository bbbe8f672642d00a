// The OpenFHE Hybrid adapter (include/ntt/openfhe_hybrid_adapter.hpp) on the
// CPU: OpenFHE is replaced by the test double in support/openfhe_double (plain
// vectors), the Session by the replay transport, and every result is
// compared with the C oracle. Covers the five entry points, their commit
// order (c0 + B and A, or c0 + B and c1 + A; relinearization in key order),
// rotations via 5^k, hoisted rotations with identity shifts, and the
// transaction rule: on any failure the caller's ciphertext is unchanged.
#include <iostream>

#include "ntt/openfhe_hybrid_adapter.hpp"
#include "ntt/param_gen.hpp"
#include "support/replay_transport.hpp"

extern "C" void or_hybrid_mac(const uint64_t *digits, uint64_t beta, const uint64_t *key_b, const uint64_t *key_a,
                              uint64_t n, const uint64_t *q, uint64_t L, const uint64_t *p, uint64_t K, uint64_t M,
                              uint64_t *out_b, uint64_t *out_a);

using namespace lbcrypto;
using ntt::proposal::ContractError;
using ntt::proposal::PolynomialLayout;
namespace core = ntt::openfhe::core;

namespace {
size_t g_checks = 0;
void expect(bool ok, const std::string &what) {
  ++g_checks;
  if (!ok) throw std::runtime_error(what);
}
template <class E = std::exception, class F>
void expect_throw(F &&f, const std::string &what) {
  ++g_checks;
  try {
    f();
  } catch (const E &) {
    return;
  } catch (const std::exception &e) {
    throw std::runtime_error(what + " (wrong exception: " + e.what() + ")");
  }
  throw std::runtime_error(what + " (nothing thrown)");
}

constexpr uint32_t n = 16, M = 5, K = 2, alpha = 2, L = 4;

struct World {
  std::vector<uint64_t> q, qr, p, pr;  // full Q chain and P
  std::vector<std::shared_ptr<ILNativeParams>> qt, pt;
  std::shared_ptr<ILDCRTParams> full_q, level, qp;
  CryptoContext<DCRTPoly> ctx;
  size_t sessions = 0;
  bool fail_transport = false;

  World() {
    const auto primes = ntt::find_primes(n, M + K, 30);
    for (size_t i = 0; i < M + K; ++i) {
      const uint64_t psi = ntt::find_psi(primes[i], n);
      (i < M ? q : p).push_back(primes[i]);
      (i < M ? qr : pr).push_back(psi);
      (i < M ? qt : pt).push_back(std::make_shared<ILNativeParams>(2 * n, primes[i], psi));
    }
    full_q = std::make_shared<ILDCRTParams>(2 * n, qt);
    level = std::make_shared<ILDCRTParams>(2 * n, std::vector<std::shared_ptr<ILNativeParams>>(qt.begin(), qt.begin() + L));
    auto all = qt;
    all.insert(all.end(), pt.begin(), pt.end());
    qp = std::make_shared<ILDCRTParams>(2 * n, all);
    auto crypto = std::make_shared<CryptoParametersCKKSRNS>();
    crypto->q = full_q;
    crypto->p = std::make_shared<ILDCRTParams>(2 * n, pt);
    crypto->per_part = alpha;
    ctx = std::make_shared<CryptoContextImpl<DCRTPoly>>();
    ctx->params = crypto;
    ctx->cyclotomic_order = 2 * n;
  }
  // seeded residues as a DCRTPoly over `params`
  DCRTPoly poly(const std::shared_ptr<ILDCRTParams> &params, uint64_t seed, Format f = EVALUATION) const {
    DCRTPoly out(params, f, false);
    uint64_t x = seed * 0x9e3779b97f4a7c15ull + 7;
    for (size_t t = 0; t < params->GetParams().size(); ++t) {
      NativePoly tower(params->GetParams()[t], f);
      std::vector<uint64_t> w(n);
      for (auto &v : w) {
        x ^= x << 13;
        x ^= x >> 7;
        x ^= x << 17;
        v = x % params->GetParams()[t]->GetModulus().ConvertToInt();
      }
      tower.set_words(w, f);
      out.SetElementAtIndex(t, std::move(tower));
    }
    return out;
  }
  EvalKey<DCRTPoly> key(uint64_t seed) const {
    auto k = std::make_shared<EvalKeyImpl<DCRTPoly>>();
    for (uint32_t d = 0; d < (M + alpha - 1) / alpha; ++d) {
      k->b.push_back(poly(qp, seed + d));
      k->a.push_back(poly(qp, seed + 100 + d));
    }
    k->context = ctx;
    return k;
  }
  auto maker() {
    return [this](const PolynomialLayout &l) {
      ++sessions;
      auto t = std::make_shared<replay::Transport>(l);
      if (fail_transport)
        t->log->fault = [](size_t, replay::Call c) {
          return c == replay::Call::Switch ? replay::Fault::Throw : replay::Fault::None;
        };
      return ntt::resident::Session(t);
    };
  }
};

std::vector<uint64_t> words(const DCRTPoly &p) {
  std::vector<uint64_t> out;
  for (const auto &t : p.GetAllElements()) {
    const auto w = t.words();
    out.insert(out.end(), w.begin(), w.end());
  }
  return out;
}
std::vector<uint64_t> flat(const std::vector<DCRTPoly> &ps) {
  std::vector<uint64_t> out;
  for (const auto &p : ps) {
    const auto w = words(p);
    out.insert(out.end(), w.begin(), w.end());
  }
  return out;
}
std::vector<uint64_t> add(const std::vector<uint64_t> &a, const std::vector<uint64_t> &b, const std::vector<uint64_t> &q) {
  std::vector<uint64_t> out(a.size());
  for (size_t i = 0; i < a.size(); ++i) out[i] = (a[i] + b[i]) % q[i / n];
  return out;
}

// (B, A) of the oracle's key switch of evaluation words c at level L
std::pair<std::vector<uint64_t>, std::vector<uint64_t>> oracle_ks(const World &w, const std::vector<uint64_t> &c,
                                                                  const EvalKey<DCRTPoly> &k) {
  std::vector<uint64_t> b(L * n), a(L * n);
  const auto kb = flat(k->b), ka = flat(k->a);
  or_key_switch(c.data(), n, w.q.data(), w.qr.data(), L, M, w.p.data(), w.pr.data(), K, alpha, kb.data(), ka.data(),
                b.data(), a.data());
  return {b, a};
}
// rotation by automorphism k of the ciphertext (c0, c1): sigma_k(B + P c0), sigma_k(A) before ModDown
std::pair<std::vector<uint64_t>, std::vector<uint64_t>> oracle_rot(const World &w, const std::vector<uint64_t> &c0,
                                                                   const std::vector<uint64_t> &c1,
                                                                   const EvalKey<DCRTPoly> &key, uint32_t k) {
  const size_t T = L + K, beta = (L + alpha - 1) / alpha;
  std::vector<uint64_t> dig(beta * T * n), zb(T * n), za(T * n), xb(T * n), xa(T * n), b(L * n), a(L * n);
  const auto kb = flat(key->b), ka = flat(key->a);
  or_hybrid_digits(c1.data(), n, w.q.data(), w.qr.data(), L, w.p.data(), w.pr.data(), K, alpha, dig.data());
  or_hybrid_mac(dig.data(), beta, kb.data(), ka.data(), n, w.q.data(), L, w.p.data(), K, M, zb.data(), za.data());
  or_hybrid_rotate(zb.data(), za.data(), c0.data(), n, w.q.data(), L, w.p.data(), K, k, xb.data(), xa.data());
  or_mod_down(xb.data(), n, w.q.data(), w.qr.data(), L, w.p.data(), w.pr.data(), K, b.data());
  or_mod_down(xa.data(), n, w.q.data(), w.qr.data(), L, w.p.data(), w.pr.data(), K, a.data());
  return {b, a};
}

Ciphertext<DCRTPoly> ciphertext(const World &w, std::vector<DCRTPoly> parts) {
  auto c = std::make_shared<CiphertextImpl<DCRTPoly>>();
  c->context = w.ctx;
  c->elements = std::move(parts);
  return c;
}

// ---------------------------------------------------------------- scenarios
void core_parts() {
  expect(core::normalize_slot_shift(3, 16) == 3 && core::normalize_slot_shift(9, 16) == 1, "shift mod N/2");
  expect(core::normalize_slot_shift(-9, 16) == -1, "negative shift keeps its sign");
  expect(core::normalize_slot_shift(31, 16) == 31, "conjugation index 2N - 1 passes");
  const int32_t shifts[] = {0, 2, 0};
  const auto plan = core::plan_rotations(shifts);
  expect(plan.size() == 3 && plan[0].identity && !plan[1].identity && !core::all_identity(plan), "rotation plan");
  const int32_t zeros[] = {0, 0};
  expect(core::all_identity(core::plan_rotations(zeros)), "all-zero shifts");
  World w;
  core::BasisDescription d{n, {}, {}, alpha, L};
  for (size_t i = 0; i < M; ++i) d.q.push_back({w.q[i], w.qr[i]});
  for (size_t i = 0; i < K; ++i) d.p.push_back({w.p[i], w.pr[i]});
  const auto src = core::source_layout(d);
  expect(src.active_q == L && src.towers.size() == L && src.domain == ntt::proposal::Domain::Evaluation, "source layout");
  expect(core::session_basis(src).towers.size() == L + K && core::full_key_layout(src).towers.size() == M + K,
         "session basis and key layout");
  expect(core::full_key_digits(src) == 3, "key digits of the full chain");
  auto bad = d;
  bad.active_q = M + 1;
  expect_throw<ContractError>([&] { core::source_layout(bad); }, "level above the chain");
  bad = d;
  bad.p.clear();
  expect_throw<ContractError>([&] { core::source_layout(bad); }, "no special primes");
  // commit order with a value type that records its history
  struct Tag {
    std::string s;
    Tag &operator+=(const Tag &o) {
      s += "+" + o.s;
      return *this;
    }
  };
  auto ev = [](Tag &t) { t.s = "E(" + t.s + ")"; };
  const Tag two[] = {{"c0"}, {"c1"}};
  auto c2 = core::key_switch_candidate<Tag>(two, Tag{"B"}, Tag{"A"}, ev);
  expect(c2.size() == 2 && c2[0].s == "E(c0)+B" && c2[1].s == "A", "two components: c0 + B, A");
  const Tag three[] = {{"c0"}, {"c1"}, {"c2"}};
  auto c3 = core::key_switch_candidate<Tag>(three, Tag{"B"}, Tag{"A"}, ev);
  expect(c3.size() == 2 && c3[0].s == "E(c0)+B" && c3[1].s == "E(c1)+A", "three components: c0 + B, c1 + A");
  const Tag four[] = {{"c0"}, {"c1"}, {"c2"}, {"c3"}};
  auto r4 = core::relinearize_candidate<Tag>(four, ev, [](size_t j) {
    return std::make_pair(Tag{"B" + std::to_string(j)}, Tag{"A" + std::to_string(j)});
  });
  expect(r4[0].s == "E(c0)+B0+B1" && r4[1].s == "E(c1)+A0+A1", "relinearization adds in key order");
}

void key_switch_core() {
  World w;
  const auto key = w.key(10);
  const DCRTPoly c = w.poly(w.level, 1);
  const auto got = ntt::openfhe::hybrid_key_switch_core(c, key, w.maker());
  const auto [wb, wa] = oracle_ks(w, words(c), key);
  expect(words(got.b) == wb && words(got.a) == wa, "core (B, A) equal the oracle's key switch");
  expect(got.b.GetFormat() == EVALUATION && got.b.GetParams() == w.level, "results over the input's parameters");
  // a coefficient-form input is rejected before any session
  const size_t before = w.sessions;
  expect_throw<ContractError>([&] { ntt::openfhe::hybrid_key_switch_core(w.poly(w.level, 2, COEFFICIENT), key, w.maker()); },
                              "coefficient input");
  auto short_key = w.key(20);
  short_key->b.pop_back();
  expect_throw<ContractError>([&] { ntt::openfhe::hybrid_key_switch_core(c, short_key, w.maker()); }, "key digit count");
  expect(w.sessions == before, "rejections open no session");
}

void key_switch_in_place() {
  World w;
  const auto key = w.key(30);
  for (int parts : {2, 3}) {
    const DCRTPoly c0 = w.poly(w.level, 41, COEFFICIENT), c1 = w.poly(w.level, 42), c2 = w.poly(w.level, 43);
    auto ct = parts == 2 ? ciphertext(w, {c0, c1}) : ciphertext(w, {c0, c1, c2});
    ntt::openfhe::hybrid_key_switch_in_place(ct, key, w.maker());
    DCRTPoly e0 = c0;
    e0.SetFormat(EVALUATION);
    const auto [b, a] = oracle_ks(w, words(parts == 2 ? c1 : c2), key);
    const auto qs = std::vector<uint64_t>(w.q.begin(), w.q.begin() + L);
    expect(ct->GetElements().size() == 2, "two components afterwards");
    expect(words(ct->GetElements()[0]) == add(words(e0), b, qs), "c0 + B");
    expect(words(ct->GetElements()[1]) == (parts == 2 ? a : add(words(c1), a, qs)), parts == 2 ? "A" : "c1 + A");
  }
  // a failure inside the transaction leaves the ciphertext as it was
  auto ct = ciphertext(w, {w.poly(w.level, 51), w.poly(w.level, 52)});
  const auto before = flat(ct->GetElements());
  w.fail_transport = true;
  expect_throw<std::runtime_error>([&] { ntt::openfhe::hybrid_key_switch_in_place(ct, key, w.maker()); },
                                   "transport failure surfaces");
  expect(flat(ct->GetElements()) == before && ct->GetElements().size() == 2, "ciphertext unchanged on failure");
  w.fail_transport = false;
  auto other = w.key(60);
  other->context = std::make_shared<CryptoContextImpl<DCRTPoly>>(*w.ctx);
  expect_throw<ContractError>([&] { ntt::openfhe::hybrid_key_switch_in_place(ct, other, w.maker()); },
                              "key of another context");
  expect(flat(ct->GetElements()) == before, "unchanged after a rejection");
}

void relinearize() {
  World w;
  const auto k0 = w.key(70), k1 = w.key(80);
  w.ctx->mult_keys["key"] = {k0, k1};
  const auto qs = std::vector<uint64_t>(w.q.begin(), w.q.begin() + L);
  for (int extra : {0, 1, 2}) {
    std::vector<DCRTPoly> parts = {w.poly(w.level, 91), w.poly(w.level, 92, COEFFICIENT)};
    for (int j = 0; j < extra; ++j) parts.push_back(w.poly(w.level, 93 + j, j == 1 ? COEFFICIENT : EVALUATION));
    auto ct = ciphertext(w, parts);
    ntt::openfhe::hybrid_relinearize_in_place(ct, w.maker());
    DCRTPoly e1 = parts[1];
    e1.SetFormat(EVALUATION);
    auto want0 = words(parts[0]), want1 = words(e1);
    for (int j = 0; j < extra; ++j) {
      DCRTPoly hj = parts[2 + j];
      hj.SetFormat(EVALUATION);
      const auto [b, a] = oracle_ks(w, words(hj), j == 0 ? k0 : k1);
      want0 = add(want0, b, qs);
      want1 = add(want1, a, qs);
    }
    expect(ct->GetElements().size() == 2 && words(ct->GetElements()[0]) == want0 &&
               words(ct->GetElements()[1]) == want1,
           "relinearization with " + std::to_string(extra) + " higher components");
  }
  w.ctx->mult_keys["key"] = {k0};
  auto ct = ciphertext(w, {w.poly(w.level, 1), w.poly(w.level, 2), w.poly(w.level, 3), w.poly(w.level, 4)});
  expect_throw<ContractError>([&] { ntt::openfhe::hybrid_relinearize_in_place(ct, w.maker()); }, "too few keys");
}

void rotations() {
  World w;
  auto keys = std::make_shared<std::map<uint32_t, EvalKey<DCRTPoly>>>();
  for (int32_t s : {1, 3, -1}) (*keys)[w.ctx->FindAutomorphismIndex((uint32_t)s)] = w.key(200 + 10 * s);
  w.ctx->rotation_keys["key"] = keys;
  const DCRTPoly c0 = w.poly(w.level, 301), c1 = w.poly(w.level, 302);
  ConstCiphertext<DCRTPoly> ct = ciphertext(w, {c0, c1});
  expect(w.ctx->FindAutomorphismIndex(3) == 125 % 32 && w.ctx->FindAutomorphismIndex((uint32_t)-1) * 5 % 32 == 1,
         "CKKS automorphism indices 5^k mod 2N");
  for (int32_t s : {3, 3 + (int32_t)n / 2, -1}) {
    const auto out = ntt::openfhe::hybrid_rotate(ct, s, w.maker());
    const uint32_t k = w.ctx->FindAutomorphismIndex((uint32_t)core::normalize_slot_shift(s, n));
    const auto [b, a] = oracle_rot(w, words(c0), words(c1), keys->at(k), k);
    expect(words(out->GetElements()[0]) == b && words(out->GetElements()[1]) == a,
           "rotation by " + std::to_string(s) + " equals the oracle");
  }
  expect_throw<ContractError>([&] { ntt::openfhe::hybrid_rotate(ct, 2, w.maker()); }, "missing rotation key");
  // hoisted: one session, identity shifts cloned
  const int32_t shifts[] = {1, 0, 3, -1};
  const size_t before = w.sessions;
  const auto many = ntt::openfhe::hybrid_rotate_many(ct, shifts, w.maker());
  expect(w.sessions == before + 1 && many.size() == 4, "hoisted rotations share one session");
  expect(flat(many[1]->GetElements()) == flat(ct->GetElements()), "zero shift returns the input");
  for (size_t r : {0u, 2u, 3u}) {
    const auto one = ntt::openfhe::hybrid_rotate(ct, shifts[r], w.maker());
    expect(flat(many[r]->GetElements()) == flat(one->GetElements()), "hoisted equals single rotation");
  }
  const int32_t zeros[] = {0, 0};
  const size_t s2 = w.sessions;
  const auto ids = ntt::openfhe::hybrid_rotate_many(ct, zeros, w.maker());
  expect(ids.size() == 2 && w.sessions == s2, "all-zero shifts open no session");
  const int32_t with_missing[] = {1, 2};
  const size_t s3 = w.sessions;
  expect_throw<ContractError>([&] { ntt::openfhe::hybrid_rotate_many(ct, with_missing, w.maker()); },
                              "a missing key is found before the session opens");
  expect(w.sessions == s3, "no session for a rejected batch");
}
}  // namespace

int main() {
  const std::pair<const char *, void (*)()> scenarios[] = {
      {"core_parts", core_parts},           {"key_switch_core", key_switch_core},
      {"key_switch_in_place", key_switch_in_place}, {"relinearize", relinearize},
      {"rotations", rotations},
  };
  for (const auto &[name, run] : scenarios) {
    try {
      run();
    } catch (const std::exception &e) {
      std::cerr << "FAILED " << name << ": " << e.what() << "\n";
      return 1;
    }
    std::cout << "ok " << name << "\n";
  }
  std::cout << g_checks << " ADAPTER CHECKS PASSED\n";
  return 0;
}
