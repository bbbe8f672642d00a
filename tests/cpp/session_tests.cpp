// The include/ntt Session runtime on the deferred CPU transport (no GPU).
// Each scenario pins one part of the Session contract (include/ntt/resident.hpp;
// reference behaviour src/resident.cpp:1-662): the driver-call order of every
// operation, fresh immutable outputs, completion and release of pinned
// resources, byte and operation limits, preflight rejections that reach no
// driver, failure at every driver call, quarantine when quiescence cannot be
// proven, validation of downloaded values, and the conversion and Hybrid
// key-switching chains, whose results are compared with the C oracle.
#include <cstring>
#include <iostream>
#include <optional>

#include "ntt/param_gen.hpp"
#include "ntt/resident.hpp"
#include "support/replay_transport.hpp"

extern "C" void or_hybrid_mac(const uint64_t *digits, uint64_t beta, const uint64_t *key_b, const uint64_t *key_a,
                              uint64_t n, const uint64_t *q, uint64_t L, const uint64_t *p, uint64_t K, uint64_t M,
                              uint64_t *out_b, uint64_t *out_a);

using namespace ntt::resident;
using namespace ntt::proposal;
using replay::Call;
using replay::Fault;
using replay::Transport;

namespace {

size_t g_checks = 0;
std::string g_scenario;

void expect(bool ok, const std::string &what) {
  ++g_checks;
  if (!ok) throw std::runtime_error(g_scenario + ": " + what);
}

template <class E = std::exception, class F>
std::string expect_throw(F &&f, const std::string &what) {
  ++g_checks;
  try {
    f();
  } catch (const E &e) {
    return e.what();
  } catch (const std::exception &e) {
    throw std::runtime_error(g_scenario + ": " + what + " (wrong exception: " + e.what() + ")");
  }
  throw std::runtime_error(g_scenario + ": " + what + " (nothing thrown)");
}

// A basis of `q` Q towers and `p` P towers of ring dimension n (primes of
// `bits` bits), as the session's coefficient-form layout.
PolynomialLayout basis(size_t q, size_t p = 0, uint64_t n = 16, size_t alpha = 1, uint64_t bits = 30,
                       size_t active = 0) {
  const auto primes = ntt::find_primes(n, q + p, bits);
  std::vector<Tower> tq, tp;
  for (size_t i = 0; i < q + p; ++i) (i < q ? tq : tp).push_back({primes[i], ntt::find_psi(primes[i], n)});
  auto par = Parameters::create(n, tq, tp, alpha);
  const size_t L = active ? active : q;
  return {par, L, p ? qp_basis(*par, L) : q_prefix(*par, L), Domain::Coefficient, Order::Natural};
}

std::vector<uint64_t> moduli(const PolynomialLayout &l) {
  std::vector<uint64_t> m;
  for (auto t : l.towers) m.push_back(l.parameters->tower(t).q);
  return m;
}
std::vector<uint64_t> roots(const PolynomialLayout &l) {
  std::vector<uint64_t> r;
  for (auto t : l.towers) r.push_back(l.parameters->tower(t).psi);
  return r;
}

// Seeded canonical residues (splitmix-style stream).
std::vector<uint64_t> values(const std::vector<uint64_t> &mods, uint64_t n, uint64_t seed) {
  std::vector<uint64_t> v(mods.size() * n);
  uint64_t x = seed * 0x9e3779b97f4a7c15ull + 1;
  for (size_t i = 0; i < v.size(); ++i) {
    x ^= x >> 31;
    x *= 0xbf58476d1ce4e5b9ull;
    x ^= x >> 29;
    v[i] = x % mods[i / n];
  }
  return v;
}
std::vector<uint64_t> values(const PolynomialLayout &l, uint64_t seed) {
  return values(moduli(l), l.parameters->n(), seed);
}

// Negacyclic schoolbook product per tower (independent of every transform).
std::vector<uint64_t> ring_product(const PolynomialLayout &l, const std::vector<uint64_t> &a,
                                   const std::vector<uint64_t> &b) {
  const size_t n = l.parameters->n();
  const auto m = moduli(l);
  std::vector<uint64_t> out(a.size(), 0);
  for (size_t t = 0; t < m.size(); ++t)
    for (size_t i = 0; i < n; ++i)
      for (size_t j = 0; j < n; ++j) {
        const uint64_t v = (uint64_t)((unsigned __int128)a[t * n + i] * b[t * n + j] % m[t]);
        uint64_t &r = out[t * n + (i + j) % n];
        r = i + j < n ? (r + v) % m[t] : (r + m[t] - v) % m[t];
      }
  return out;
}

bool same(std::span<const uint64_t> a, const std::vector<uint64_t> &b) {
  return a.size() == b.size() && std::equal(a.begin(), a.end(), b.begin());
}

struct Rig {
  std::shared_ptr<Transport> transport;
  std::shared_ptr<replay::Log> log;
  std::optional<Session> session;
  explicit Rig(const PolynomialLayout &l, Limits limits = {}) : transport(std::make_shared<Transport>(l)) {
    log = transport->log;
    session.emplace(transport, limits);
  }
  Session &s() { return *session; }
};

// ---------------------------------------------------------------- scenarios
void ring_product_chain() {
  const auto l = basis(2);
  Rig r(l);
  const auto a = values(l, 1), b = values(l, 2);
  size_t mark = r.log->calls.size();
  Polynomial pa = r.s().upload(a);
  expect(r.log->since(mark) == "event allocate upload record", "upload call order: " + r.log->since(mark));
  Polynomial pb = r.s().upload(b);
  mark = r.log->calls.size();
  Polynomial fa = r.s().forward(pa);
  expect(r.log->since(mark) == "event allocate copy transform record", "forward call order: " + r.log->since(mark));
  Polynomial fb = r.s().forward(pb);
  mark = r.log->calls.size();
  Polynomial prod = r.s().pointwise(fa, fb);
  expect(r.log->since(mark) == "event allocate copy pointwise record", "pointwise order: " + r.log->since(mark));
  Polynomial back = r.s().inverse(prod);
  mark = r.log->calls.size();
  Download d = r.s().download(back);
  expect(r.log->since(mark) == "event download record", "download call order: " + r.log->since(mark));

  const Stats st = r.s().stats();
  expect(st.submitted == 7 && st.pending == 7 && !st.quiescent, "seven pending journal entries");
  expect(r.log->jobs_run == 0, "submission runs nothing on the host");
  expect(st.uploads == 2 && st.downloads == 1 && st.device_copies == 4 && st.parameter_uploads == 0,
         "upload/download/copy counters");
  expect(pa.view().allocation != fa.view().allocation && fa.view().allocation != prod.view().allocation &&
             prod.view().allocation != back.view().allocation,
         "every output is a fresh allocation");
  expect(fa.layout().domain == Domain::Evaluation && back.layout().domain == Domain::Coefficient,
         "domains follow the transforms");
  expect(pa.readiness().query().outcome == Outcome::Pending, "query before execution is pending");

  expect(same(d.wait_values(), ring_product(l, a, b)), "product equals the schoolbook product");
  const Stats after = r.s().stats();
  expect(after.pending == 0 && after.quiescent, "waiting on the last entry settles everything");
  const size_t waits = r.log->count(Call::Wait);
  expect(pa.readiness().query().outcome == Outcome::Succeeded, "earlier entries complete with the later one");
  expect(r.log->count(Call::Wait) == waits, "a settled entry needs no further driver wait");
  const auto h = r.s().history();
  expect(h.size() == 7 && h.front().operation.id == 1 && h.back().operation.id == 7, "journal ids 1..7");
  for (const auto &rec : h) expect(rec.outcome == Outcome::Succeeded, "every entry succeeded");

  pa = pb = fa = fb = prod = back = Polynomial();
  d = Download();
  expect(r.log->live_buffers == 0 && r.log->live_events == 0, "released handles free buffers and events");
  expect(r.s().stats().device_bytes == 0 && r.s().stats().host_bytes == 0, "budgets return to zero");
  r.s().close();
  expect(r.s().stats().closed, "closed");
}

void inputs_are_immutable() {
  const auto l = basis(3);
  Rig r(l);
  const auto a = values(l, 5);
  Polynomial x = r.s().upload(a);
  Polynomial f1 = r.s().forward(x), f2 = r.s().forward(x);
  Polynomial i1 = r.s().inverse(f1);
  const auto w1 = r.s().download(f1), w2 = r.s().download(f2), wx = r.s().download(x), wi = r.s().download(i1);
  expect(same(w1.wait_values(), std::vector<uint64_t>(w2.wait_values().begin(), w2.wait_values().end())),
         "the same input transforms to the same output twice");
  expect(same(wx.wait_values(), a), "the input is unchanged by operations that read it");
  expect(same(wi.wait_values(), a), "inverse(forward(x)) == x");
  std::vector<uint64_t> want = a;
  auto m = moduli(l), psi = roots(l);
  or_ntt_towers(want.data(), l.parameters->n(), m.size(), m.data(), psi.data(), 0);
  expect(same(w1.wait_values(), want), "forward equals the oracle's transform");
}

void preflight_rejections() {
  const auto l = basis(2);
  Rig r(l);
  auto good = values(l, 9);
  Polynomial x = r.s().upload(good);
  Polynomial fx = r.s().forward(x);
  const size_t calls = r.log->calls.size();
  auto bad = good;
  bad[3] = moduli(l)[0];
  expect_throw<ContractError>([&] { r.s().upload(bad); }, "non-canonical residue");
  expect_throw<ContractError>([&] { r.s().upload(std::span<const uint64_t>(good).first(good.size() - 1)); },
                              "short host array");
  expect_throw<ContractError>([&] { r.s().forward(fx); }, "forward of evaluation input");
  expect_throw<ContractError>([&] { r.s().inverse(x); }, "inverse of coefficient input");
  expect_throw<ContractError>([&] { r.s().pointwise(x, x); }, "pointwise of coefficient input");
  expect_throw<ContractError>([&] { r.s().subview(x, 1, 2, 2); }, "subview past its parent");
  expect_throw<ContractError>([&] { r.s().subview(r.s().subview(x, 0, 2, 1), 0, 1, 2); }, "subview raising level");
  expect_throw<ContractError>([&] { r.s().mod_up(x); }, "ModUp on a session without P towers");
  expect_throw<ContractError>([&] { r.s().forward(Polynomial()); }, "empty polynomial");
  Rig other(l);
  Polynomial y = other.s().upload(good);
  expect_throw<ContractError>([&] { r.s().forward(y); }, "polynomial of another session");
  const Completion foreign[] = {y.readiness()};
  expect_throw<ContractError>([&] { r.s().forward(x, foreign); }, "wait on another session");
  expect(r.log->calls.size() == calls, "rejections reach no driver call");
  expect(!r.s().stats().poisoned, "rejections do not poison");
  expect_throw<ContractError>([&] { (void)Polynomial().layout(); }, "empty handle has no layout");
  expect_throw<ContractError>([&] { Session s(nullptr); }, "session without driver");
  // one transport per session
  expect_throw<ContractError>([&] { Session again(r.transport); }, "second session on one transport");
  expect_throw<ContractError>([&] { Rig z(l, Limits{1 << 20, 1 << 20, 0}); }, "zero operation limit");
  // a session is opened on coefficient layouts only
  auto eval = transform_layout(l, false);
  expect_throw<ContractError>([&] { Session s(std::make_shared<Transport>(eval)); }, "evaluation session basis");
  // explicit waits on earlier completions of the same session are accepted
  const Completion mine[] = {x.readiness(), fx.readiness()};
  Polynomial z = r.s().inverse(fx, mine);
  expect(same(r.s().download(z).wait_values(), good), "explicit waits are honoured");
}

void views() {
  const auto l = basis(4);
  Rig r(l);
  const auto a = values(l, 11);
  const size_t n = l.parameters->n();
  Polynomial x = r.s().upload(a);
  Polynomial mid = r.s().subview(x, 1, 3, 4);
  Polynomial inner = r.s().subview(mid, 1, 2, 4);
  expect(inner.view().first_tower == 2 && inner.view().tower_count == 2, "nested view offsets add");
  expect(inner.view().allocation == x.view().allocation, "a view shares its parent's allocation");
  const std::vector<uint64_t> slice(a.begin() + 2 * n, a.begin() + 4 * n);
  expect(same(r.s().download(inner).wait_values(), slice), "a view downloads its towers");
  // a transform of a view reads only the view's towers
  Polynomial f = r.s().forward(inner);
  std::vector<uint64_t> want = slice;
  auto m = moduli(l), psi = roots(l);
  or_ntt_towers(want.data(), n, 2, m.data() + 2, psi.data() + 2, 0);
  expect(same(r.s().download(f).wait_values(), want), "transform of a view");
}

void limits() {
  const auto l = basis(2);
  const size_t poly = bytes(l);
  {  // device bytes: two polynomials fit, a third does not until one is released
    Rig r(l, Limits{2 * poly, 64 * poly, 100});
    Polynomial a = r.s().upload(values(l, 1));
    Polynomial b = r.s().upload(values(l, 2));
    const size_t calls = r.log->calls.size();
    expect_throw<ContractError>([&] { r.s().upload(values(l, 3)); }, "device limit");
    expect(r.log->calls.size() == calls && !r.s().stats().poisoned, "a refusal submits nothing");
    r.s().download(b).wait_values();
    a = Polynomial();
    Polynomial c = r.s().upload(values(l, 3));  // fits again
    expect(r.s().stats().device_bytes == 2 * poly, "device bytes of live polynomials");
  }
  {  // host staging is held until its entry completes
    Rig r(l, Limits{64 * poly, poly, 100});
    Polynomial a = r.s().upload(values(l, 1));
    expect(r.s().stats().host_bytes == poly, "staging charged while the upload is pending");
    expect_throw<ContractError>([&] { r.s().upload(values(l, 2)); }, "host limit");
    a.readiness().wait();
    expect(r.s().stats().host_bytes == 0, "staging released on completion");
    Polynomial b = r.s().upload(values(l, 2));
  }
  {  // the journal is append-only: its limit counts every submission
    Rig r(l, Limits{64 * poly, 64 * poly, 3});
    Polynomial a = r.s().upload(values(l, 1));
    Polynomial f = r.s().forward(a);
    f.readiness().wait();
    Polynomial g = r.s().inverse(f);
    g.readiness().wait();
    const size_t calls = r.log->calls.size();
    expect_throw<ContractError>([&] { r.s().forward(a); }, "operation limit");
    expect(r.log->calls.size() == calls && !r.s().stats().poisoned, "journal refusal submits nothing");
  }
}

// A short chain used by the fault sweeps.
void fault_chain(Session &s, const PolynomialLayout &l) {
  Polynomial a = s.upload(values(l, 1));
  Polynomial f = s.forward(a);
  Polynomial p = s.pointwise(f, f);
  s.download(s.inverse(p)).wait_values();
}

void failure_at_every_call() {
  const auto l = basis(2);
  size_t total = 0;
  {
    Rig clean(l);
    fault_chain(clean.s(), l);
    total = clean.log->calls.size();
  }
  expect(total > 20, "the chain makes driver calls");
  for (size_t k = 1; k <= total; ++k)
    for (Fault mode : {Fault::Throw, Fault::ThrowAfterQueue}) {
      Rig r(l);
      r.log->fault = [k, mode](size_t i, Call) { return i == k ? mode : Fault::None; };
      bool thrown = false;
      try {
        fault_chain(r.s(), l);
      } catch (const std::runtime_error &) {
        thrown = true;
      }
      // ThrowAfterQueue only fires on calls that queue work
      const Call c = r.log->calls.size() >= k ? r.log->calls[k - 1] : Call::Record;
      const bool queues = c != Call::Allocate && c != Call::MakeEvent && c != Call::Record && c != Call::Ready &&
                          c != Call::Wait;
      if (mode == Fault::ThrowAfterQueue && !queues) continue;
      expect(thrown, "failure at call " + std::to_string(k) + " surfaces");
      const Stats st = r.s().stats();
      expect(st.poisoned && st.quiescent && st.pending == 0, "failure poisons and drains (call " + std::to_string(k) + ")");
      const auto h = r.s().history();
      expect(!h.empty() && h.back().outcome == Outcome::Failed, "the failing entry is recorded as failed");
      expect_throw<ContractError>([&] { r.s().upload(values(l, 2)); }, "a poisoned session admits nothing");
      expect_throw<std::runtime_error>([&] { r.s().close(); }, "close reports the first failure");
      r.session.reset();
      expect(r.log->live_buffers == 0 && r.log->live_events == 0, "nothing stays pinned after a proven drain");
    }
}

void failure_while_running() {
  const auto l = basis(2);
  for (size_t job = 1; job <= 6; ++job) {
    Rig r(l);
    size_t queued = 0;
    r.log->fault = [&queued, job](size_t, Call c) {
      const bool q = c != Call::Allocate && c != Call::MakeEvent && c != Call::Record && c != Call::Ready &&
                     c != Call::Wait;
      return q && ++queued == job ? Fault::FailWhenRun : Fault::None;
    };
    Polynomial a = r.s().upload(values(l, 1));
    Polynomial f = r.s().forward(a);
    Polynomial p = r.s().pointwise(f, f);
    Download d = r.s().download(r.s().inverse(p));
    expect(!r.s().stats().poisoned, "a queued failure is not visible at submission");
    r.transport->run_all();
    // query reports the failed record instead of throwing; the session is poisoned
    expect(d.completion().query().outcome == Outcome::Failed, "query reports the transport error as a failure");
    expect(r.s().stats().poisoned && r.s().stats().quiescent, "the transport error poisons and drains");
    expect_throw<std::runtime_error>([&] { d.wait_values(); }, "wait rethrows the first failure");
  }
}

void quarantine() {
  const auto l = basis(2);
  Rig r(l);
  r.log->quiesce_ok = false;
  Polynomial a = r.s().upload(values(l, 1));
  Polynomial f = r.s().forward(a);
  r.log->fault = [](size_t, Call c) { return c == Call::Record ? Fault::Throw : Fault::None; };
  expect_throw<std::runtime_error>([&] { r.s().download(f); }, "failure with an unprovable drain");
  const Stats st = r.s().stats();
  expect(st.poisoned && !st.quiescent, "poisoned and not quiescent");
  const size_t held = r.log->live_buffers;
  a = f = Polynomial();
  expect(held >= 2 && r.log->live_buffers == held, "queued work keeps its buffers while unproven");
  expect(r.s().stats().host_bytes > 0, "and its staging");
  expect_throw<std::runtime_error>([&] { r.s().close(); }, "close cannot prove the drain either");
  expect(r.log->live_buffers == held, "still quarantined after close");
  r.log->quiesce_ok = true;
  r.log->fault = nullptr;
  expect_throw<std::runtime_error>([&] { r.s().close(); }, "a later close still reports the first failure");
  expect(r.s().stats().quiescent && r.log->live_buffers == 0 && r.log->live_events == 0,
         "a proven drain reclaims the quarantine");
  expect(r.s().stats().host_bytes == 0 && r.s().stats().device_bytes == 0, "and its budget");
}

void quarantine_outlives_the_session() {
  const auto l = basis(2);
  auto transport = std::make_shared<Transport>(l);
  auto log = transport->log;
  {
    Session s(transport);
    log->quiesce_ok = false;
    Polynomial a = s.upload(values(l, 1));
    log->fault = [](size_t, Call c) { return c == Call::Transform ? Fault::Throw : Fault::None; };
    expect_throw<std::runtime_error>([&] { s.forward(a); }, "failure");
  }  // the Session and every handle are gone; the queued upload may still run
  expect(log->live_buffers >= 1, "device memory of unproven work is never freed");
}

void bad_download_and_health() {
  const auto l = basis(2);
  {
    Rig r(l);
    r.log->corrupt_downloads = true;
    Polynomial a = r.s().upload(values(l, 1));
    Download d = r.s().download(a);
    expect_throw<ContractError>([&] { d.wait_values(); }, "a non-canonical download is rejected");
    expect(r.s().stats().poisoned && r.s().history().back().outcome == Outcome::Failed,
           "and fails its entry and the session");
  }
  {
    Rig r(l);
    r.log->unhealthy = true;
    Polynomial a = r.s().upload(values(l, 1));
    a.readiness().wait();
    expect_throw<std::runtime_error>([&] { r.s().close(); }, "health failure reported at close");
    expect(r.s().stats().closed && r.s().stats().poisoned, "closed and poisoned");
    expect(a.readiness().query().outcome == Outcome::Succeeded, "completed records survive close");
  }
  {
    Rig r(l);
    Polynomial a = r.s().upload(values(l, 1));
    r.s().close();
    expect_throw<ContractError>([&] { r.s().forward(a); }, "a closed session admits nothing");
    Session moved = std::move(*r.session);
    expect_throw<ContractError>([&] { (void)r.s().stats(); }, "a moved-from session");
    expect(moved.stats().closed, "the move carries the state");
  }
}

// ModUp (coefficient and evaluation input) and ModDown against the oracle;
// neither chain waits on the host.
void conversion_chain() {
  const auto l = basis(3, 2, 16, 2, 30);
  const size_t n = 16, L = 3, K = 2;
  Rig r(l);
  const auto mods = moduli(l), rts = roots(l);
  const std::vector<uint64_t> qmods(mods.begin(), mods.begin() + L), qroots(rts.begin(), rts.begin() + L);
  const std::vector<uint64_t> pmods(mods.begin() + L, mods.end()), proots(rts.begin() + L, rts.end());
  const auto x = values(qmods, n, 21);
  PolynomialLayout qlayout = l;
  qlayout.towers.resize(L);
  Polynomial c = r.s().upload(qlayout, x);
  Polynomial e = r.s().forward(c);
  const size_t tables = r.s().stats().parameter_uploads;
  Polynomial up_c = r.s().mod_up(c), up_e = r.s().mod_up(e);
  expect(r.s().stats().parameter_uploads == tables + 2, "one conversion table per ModUp");
  Polynomial down = r.s().mod_down(up_e);
  expect(r.s().stats().parameter_uploads == tables + 4, "ModDown uploads its conversion and correction tables");
  Polynomial down2 = r.s().mod_down(up_e);
  expect(r.s().stats().parameter_uploads == tables + 6, "a repeated ModDown uploads them again");
  expect(r.log->jobs_run == 0 && r.log->count(Call::Wait) == 0, "the chains stay queued");
  std::vector<uint64_t> want_up((L + K) * n), xe = x, want_down(L * n);
  or_mod_up(x.data(), 0, n, qmods.data(), qroots.data(), L, pmods.data(), proots.data(), K, want_up.data());
  expect(same(r.s().download(up_c).wait_values(), want_up), "ModUp of coefficient input");
  or_ntt_towers(xe.data(), n, L, qmods.data(), qroots.data(), 0);
  or_mod_up(xe.data(), 1, n, qmods.data(), qroots.data(), L, pmods.data(), proots.data(), K, want_up.data());
  expect(same(r.s().download(up_e).wait_values(), want_up), "ModUp of evaluation input");
  or_mod_down(want_up.data(), n, qmods.data(), qroots.data(), L, pmods.data(), proots.data(), K, want_down.data());
  expect(same(r.s().download(down).wait_values(), want_down), "ModDown");
  expect(same(r.s().download(down2).wait_values(), want_down), "ModDown again");
  // conversion metadata errors are preflight
  const size_t calls = r.log->calls.size();
  expect_throw<ContractError>([&] { r.s().mod_down(c); }, "ModDown of a Q-only polynomial");
  expect_throw<ContractError>([&] { r.s().mod_up(up_e); }, "ModUp of a QP polynomial");
  expect_throw<ContractError>([&] { r.s().hybrid_digits(c); }, "digits of coefficient input");
  expect(r.log->calls.size() == calls, "no driver call for rejected conversions");
}

// Digits -> key inner product -> ModDown at a reduced level (L = 4 of M = 5),
// and a rotation before ModDown, against the oracle.
void hybrid_chain() {
  const size_t n = 16, M = 5, L = 4, K = 2, alpha = 2;
  const auto l = basis(M, K, n, alpha, 30, L);
  Rig r(l);
  const auto par = l.parameters;
  std::vector<uint64_t> fullq, fullr, pm, pr;
  for (const auto &t : par->q()) fullq.push_back(t.q), fullr.push_back(t.psi);
  for (const auto &t : par->p()) pm.push_back(t.q), pr.push_back(t.psi);
  std::vector<uint64_t> fullmods = fullq;
  fullmods.insert(fullmods.end(), pm.begin(), pm.end());
  const size_t beta_full = (M + alpha - 1) / alpha, beta = (L + alpha - 1) / alpha, T = L + K;
  std::vector<std::vector<uint64_t>> kb, ka;
  std::vector<uint64_t> kb_flat, ka_flat;
  for (size_t d = 0; d < beta_full; ++d) {
    kb.push_back(values(fullmods, n, 100 + d));
    ka.push_back(values(fullmods, n, 200 + d));
    kb_flat.insert(kb_flat.end(), kb.back().begin(), kb.back().end());
    ka_flat.insert(ka_flat.end(), ka.back().begin(), ka.back().end());
  }
  std::vector<HybridKeyDigit> digits;
  for (size_t d = 0; d < beta_full; ++d) digits.push_back({kb[d], ka[d]});
  const PolynomialLayout full{par, M, qp_basis(*par, M), Domain::Evaluation, Order::BitReversed};
  // key admission
  expect_throw<ContractError>([&] { r.s().upload_hybrid_key(full, std::span(digits).first(beta_full - 1)); },
                              "wrong digit count");
  PolynomialLayout active_only = full;
  active_only.active_q = L;
  active_only.towers = qp_basis(*par, L);
  expect_throw<ContractError>([&] { r.s().upload_hybrid_key(active_only, digits); }, "key over the active basis");
  const size_t uploads = r.s().stats().uploads;
  HybridKey key = r.s().upload_hybrid_key(full, digits);
  expect(r.s().stats().uploads == uploads + 2 * beta, "one upload per active digit and channel");

  const std::vector<uint64_t> qm(fullq.begin(), fullq.begin() + L), qr(fullr.begin(), fullr.begin() + L);
  const auto c = values(qm, n, 31), c0 = values(qm, n, 32);
  PolynomialLayout qe = transform_layout(l, false);
  qe.towers.resize(L);
  Polynomial pc = r.s().upload(qe, c), pc0 = r.s().upload(qe, c0);
  const auto ds = r.s().hybrid_digits(pc);
  expect(ds.size() == beta, "one digit per alpha active towers");
  HybridContributions acc = r.s().hybrid_accumulate(ds, key);
  Polynomial ob = r.s().mod_down(acc.b), oa = r.s().mod_down(acc.a);
  std::vector<uint64_t> wb(L * n), wa(L * n);
  or_key_switch(c.data(), n, qm.data(), qr.data(), L, M, pm.data(), pr.data(), K, alpha, kb_flat.data(),
                ka_flat.data(), wb.data(), wa.data());
  expect(same(r.s().download(ob).wait_values(), wb), "key switch B at a reduced level");
  expect(same(r.s().download(oa).wait_values(), wa), "key switch A at a reduced level");

  // rotation of the extended contributions, then ModDown
  const uint32_t k = 5;
  HybridContributions rot = r.s().hybrid_rotate_extended(acc, pc0, k);
  Polynomial rb = r.s().mod_down(rot.b), ra = r.s().mod_down(rot.a);
  std::vector<uint64_t> dig(beta * T * n), zb(T * n), za(T * n), xb(T * n), xa(T * n);
  or_hybrid_digits(c.data(), n, qm.data(), qr.data(), L, pm.data(), pr.data(), K, alpha, dig.data());
  or_hybrid_mac(dig.data(), beta, kb_flat.data(), ka_flat.data(), n, qm.data(), L, pm.data(), K, M, zb.data(),
                za.data());
  or_hybrid_rotate(zb.data(), za.data(), c0.data(), n, qm.data(), L, pm.data(), K, k, xb.data(), xa.data());
  or_mod_down(xb.data(), n, qm.data(), qr.data(), L, pm.data(), pr.data(), K, wb.data());
  or_mod_down(xa.data(), n, qm.data(), qr.data(), L, pm.data(), pr.data(), K, wa.data());
  expect(same(r.s().download(rb).wait_values(), wb), "rotated B");
  expect(same(r.s().download(ra).wait_values(), wa), "rotated A");
  expect_throw<ContractError>([&] { r.s().hybrid_rotate_extended(acc, pc0, 4); }, "even automorphism");
  expect_throw<ContractError>([&] { r.s().hybrid_accumulate(std::span(ds).first(1), key); }, "digit count");
}

// A complement of more than 40 towers is converted in runs of at most 40.
void wide_digit_complements() {
  const size_t n = 8, L = 40, K = 3, alpha = 2;  // digit 0's complement: 38 Q + 3 P = 41 towers
  const auto l = basis(L, K, n, alpha, 30);
  Rig r(l);
  const auto mods = moduli(l), rts = roots(l);
  std::vector<uint64_t> qm(mods.begin(), mods.begin() + L), qr(rts.begin(), rts.begin() + L);
  std::vector<uint64_t> pm(mods.begin() + L, mods.end()), pr(rts.begin() + L, rts.end());
  const auto x = values(qm, n, 41);
  PolynomialLayout qe = transform_layout(l, false);
  qe.towers.resize(L);
  Polynomial px = r.s().upload(qe, x);
  const size_t switches = r.log->count(Call::Switch);
  const auto ds = r.s().hybrid_digits(px);
  // each side of a digit is converted in ceil(towers / 40) runs
  const size_t beta = (L + alpha - 1) / alpha;
  size_t want_switches = 0;
  for (size_t d = 0; d < beta; ++d) {
    const size_t before = d * alpha, after = L + K - std::min(L, (d + 1) * alpha);
    want_switches += (before + 39) / 40 + (after + 39) / 40;
  }
  expect(want_switches == 2 * beta && ds.size() == beta && r.log->count(Call::Switch) == switches + want_switches,
         "complements split into runs of at most 40 targets");
  std::vector<uint64_t> want(beta * (L + K) * n);
  or_hybrid_digits(x.data(), n, qm.data(), qr.data(), L, pm.data(), pr.data(), K, alpha, want.data());
  for (size_t d = 0; d < ds.size(); ++d) {
    const std::vector<uint64_t> part(want.begin() + d * (L + K) * n, want.begin() + (d + 1) * (L + K) * n);
    expect(same(r.s().download(ds[d]).wait_values(), part), "digit " + std::to_string(d) + " equals the oracle's");
  }
}

}  // namespace

int main() {
  const std::pair<const char *, void (*)()> scenarios[] = {
      {"ring_product_chain", ring_product_chain},
      {"inputs_are_immutable", inputs_are_immutable},
      {"preflight_rejections", preflight_rejections},
      {"views", views},
      {"limits", limits},
      {"failure_at_every_call", failure_at_every_call},
      {"failure_while_running", failure_while_running},
      {"quarantine", quarantine},
      {"quarantine_outlives_the_session", quarantine_outlives_the_session},
      {"bad_download_and_health", bad_download_and_health},
      {"conversion_chain", conversion_chain},
      {"hybrid_chain", hybrid_chain},
      {"wide_digit_complements", wide_digit_complements},
  };
  for (const auto &[name, run] : scenarios) {
    g_scenario = name;
    try {
      run();
    } catch (const std::exception &e) {
      std::cerr << "FAILED " << e.what() << "\n";
      return 1;
    }
    std::cout << "ok " << name << "\n";
  }
  std::cout << g_checks << " SESSION RUNTIME CHECKS PASSED\n";
  return 0;
}
