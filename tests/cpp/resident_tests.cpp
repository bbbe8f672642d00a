// Resident runtime behaviour on the deferred CPU transport (no GPU): the
// operation chain, immutability and fresh outputs, leases, budgets, failure
// injection at every driver step, quarantine, bad output, conversion chain,
// Hybrid digit complements over 40 towers, key admission and closed/moved
// sessions (behaviours of reference tests/resident_tests.cpp:49-427), plus a
// full Hybrid key switch through the Session checked against the oracle.
#include <array>
#include <cstring>
#include <functional>
#include <iostream>
#include <type_traits>

#include "ntt/resident.hpp"
#include "support/host_driver.hpp"
#include "support/oracle_api.h"

using namespace ntt::resident;
using namespace ntt::proposal;
using host_test::HostDriver;

namespace {
size_t checks = 0;
void check(bool ok, const char *why) {
  ++checks;
  if (!ok) throw std::runtime_error(why);
}
template <class F>
void rejects(F f, const char *why) {
  ++checks;
  try {
    f();
  } catch (const std::exception &) {
    return;
  }
  throw std::runtime_error(why);
}

PolynomialLayout layout(size_t q_count = 2, size_t p_count = 0, size_t n = 8, size_t alpha = 1) {
  auto primes = ntt::find_primes(n, q_count + p_count, 20);
  std::vector<Tower> q, p;
  for (size_t i = 0; i < primes.size(); ++i) (i < q_count ? q : p).push_back({primes[i], ntt::find_psi(primes[i], n)});
  auto params = Parameters::create(n, q, p, alpha);
  return {params, q_count, p_count ? qp_basis(*params, q_count) : q_prefix(*params, q_count), Domain::Coefficient,
          Order::Natural};
}

std::vector<uint64_t> input(const PolynomialLayout &p, size_t seed) {
  std::vector<uint64_t> out(bytes(p) / sizeof(uint64_t));
  const size_t n = p.parameters->n();
  for (size_t k = 0; k < p.towers.size(); ++k)
    for (size_t i = 0; i < n; ++i) out[k * n + i] = (seed + i * 17 + k * 31) % p.parameters->tower(p.towers[k]).q;
  return out;
}

// Schoolbook negacyclic product per tower: the independent side of the chain check.
std::vector<uint64_t> naive(const PolynomialLayout &p, std::span<const uint64_t> a, std::span<const uint64_t> b) {
  const size_t n = p.parameters->n();
  std::vector<uint64_t> out(a.size());
  for (size_t k = 0; k < p.towers.size(); ++k) {
    const uint64_t q = p.parameters->tower(p.towers[k]).q;
    for (size_t i = 0; i < n; ++i)
      for (size_t j = 0; j < n; ++j) {
        const uint64_t v = ntt::mul_mod(a[k * n + i], b[k * n + j], q);
        uint64_t &t = out[k * n + (i + j) % n];
        t = i + j < n ? ntt::add_mod(t, v, q) : ntt::sub_mod(t, v, q);
      }
  }
  return out;
}

void equal(std::span<const uint64_t> got, const std::vector<uint64_t> &want, const char *why) {
  check(got.size() == want.size() && std::equal(got.begin(), got.end(), want.begin()), why);
}

void chain_and_leases() {
  static_assert(!std::is_aggregate_v<Polynomial> && !std::is_aggregate_v<Completion>);
  auto p = layout();
  auto d = std::make_shared<HostDriver>(p);
  Session s(d);
  rejects([&] { Session again(d); }, "one transport per session");
  auto a = input(p, 5), b = input(p, 9), original = a;
  auto da = s.upload(a), db = s.upload(b);
  std::fill(a.begin(), a.end(), 0);  // caller memory is not borrowed
  check(da.readiness().query().outcome == Outcome::Pending, "deferred upload stays pending");
  check(d->trace->executed == 0, "submission does not execute the queue");
  auto fa = s.forward(da), fb = s.forward(db);
  auto prod = s.pointwise(fa, fb);
  auto coeff = s.inverse(prod);
  auto result = s.download(coeff);
  check(s.stats().submitted == 7 && s.stats().pending == 7, "seven journalled operations");
  check(s.stats().uploads == 2 && s.stats().downloads == 1 && s.stats().device_copies == 4,
        "two uploads, four device copies, one download");
  check(da.view().allocation != fa.view().allocation && fa.view().allocation != prod.view().allocation,
        "fresh output allocations");
  equal(result.wait_values(), naive(p, original, b), "chain equals schoolbook negacyclic product");
  check(s.stats().pending == 0 && s.stats().quiescent, "completion retires prior work");
  equal(s.download(da).wait_values(), original, "forward leaves its input untouched");
  equal(s.download(db).wait_values(), b, "right input immutable");
  auto done = da.readiness();
  auto eval = s.download(fa);
  auto eval_values = eval.wait_values();
  auto dep = s.download(s.forward(da, std::span<const Completion>(&done, 1)));
  equal(dep.wait_values(), {eval_values.begin(), eval_values.end()}, "explicit same-stream dependency");
  eval = {};
  dep = {};
  equal(s.multiply(original, b).wait_values(), naive(p, original, b), "convenience multiply");
  equal(s.download(s.inverse(s.pointwise(fa, fa))).wait_values(), naive(p, original, original),
        "read-only aliasing with fresh output");
  auto copied = da.view();
  ++copied.allocation.generation;
  check(da.view().allocation != copied.allocation, "descriptor copies cannot alter the registry");
  da = db = fa = fb = prod = coeff = {};
  result = {};
  check(d->trace->live_buffers == 0 && d->trace->live_events == 0, "retired pins free buffers and events");
  check(s.stats().device_bytes == 0 && s.stats().host_bytes == 0, "budget credits return");
}

void rejection_and_views() {
  auto p = layout(3, 2);
  auto d = std::make_shared<HostDriver>(p);
  Session s(d);
  auto a = input(p, 3), bad = a;
  bad[0] = p.parameters->tower(p.towers[0]).q;
  rejects([&] { s.multiply(a, bad); }, "bad second input rejects the whole chain");
  check(d->trace->calls == 0, "invalid inputs submit nothing");
  rejects([&] { s.upload(std::span<const uint64_t>(a).first(1)); }, "short input");
  auto da = s.upload(a);
  const size_t calls = d->trace->calls;
  rejects([&] { s.inverse(da); }, "inverse needs evaluation input");
  rejects([&] { s.pointwise(da, da); }, "pointwise needs evaluation input");
  rejects([&] { s.forward(Polynomial{}); }, "empty polynomial");
  auto d2 = std::make_shared<HostDriver>(p);
  Session other(d2);
  rejects([&] { other.forward(da); }, "foreign session handle");
  auto c = da.readiness();
  rejects([&] { other.upload(a, std::span<const Completion>(&c, 1)); }, "foreign dependency");
  rejects([&] { s.subview(da, 0, 4, 2); }, "reduced QP is not a contiguous prefix");
  rejects([&] { s.subview(da, 4, 2, 3); }, "subview beyond extent");
  check(d->trace->calls == calls && d2->trace->calls == 0, "preflight failures touch no transport");
  auto view = s.subview(da, 1, 2, 3);
  auto nested = s.subview(view, 1, 1, 3);
  equal(s.download(s.inverse(s.forward(nested))).wait_values(), {a.begin() + 16, a.begin() + 24},
        "nested views use allocation-relative offsets");
  check(nested.view().first_tower == 2, "nested view offset");
  auto e1 = s.forward(view), e2 = s.forward(nested);
  rejects([&] { s.pointwise(e1, e2); }, "incompatible tower layouts");
  s.close();
  rejects([&] { s.forward(da); }, "stale handle after close");
  rejects([&] { s.upload(a); }, "closed session refuses admission");
  check(da.readiness().query().outcome == Outcome::Succeeded, "completion metadata survives close");
  rejects([] { Completion{}.wait(); }, "empty completion");
  rejects([] { Download{}.wait_values(); }, "empty download");
}

void budgets() {
  auto p = layout();
  auto a = input(p, 4);
  auto d = std::make_shared<HostDriver>(p);
  Session s(d, {bytes(p), bytes(p), 1});
  auto da = s.upload(a);
  const size_t calls = d->trace->calls;
  rejects([&] { s.forward(da); }, "device budget");
  rejects([&] { s.download(da); }, "host staging budget");
  check(d->trace->calls == calls && !s.stats().poisoned, "budget refusals are preflight");
  da.readiness().wait();
  rejects([&] { s.download(da); }, "bounded operation journal");
  check(d->trace->calls == calls + 1, "journal refusal submits nothing");
}

void failure_matrix() {
  auto p = layout();
  auto a = input(p, 1), b = input(p, 2);
  size_t steps;
  {
    auto d = std::make_shared<HostDriver>(p);
    Session s(d);
    auto r = s.multiply(a, b);
    steps = d->trace->calls;
    r.wait_values();
  }
  for (size_t step = 1; step <= steps + 1; ++step) {
    auto d = std::make_shared<HostDriver>(p);
    auto trace = d->trace;
    {
      Session s(d);
      trace->fail_step = step;
      rejects([&] { s.multiply(a, b).wait_values(); }, "injected synchronous failure propagates");
      check(s.stats().poisoned && s.stats().quiescent, "failure drains and poisons");
      check(!s.history().empty() && s.history().back().outcome == Outcome::Failed, "failed record kept");
      rejects([&] { s.upload(a); }, "no resubmission after failure");
      rejects([&] { s.close(); }, "close rethrows the failure");
    }
    check(trace->live_buffers == 0 && trace->live_events == 0, "failure releases drained resources");
  }
  for (size_t seq : {size_t{1}, size_t{4}, size_t{11}}) {
    auto d = std::make_shared<HostDriver>(p);
    Session s(d);
    d->trace->async_fail_at = seq;
    auto r = s.multiply(a, b);
    rejects([&] { r.wait_values(); }, "asynchronous failure hides host output");
    check(s.stats().poisoned, "asynchronous failure poisons");
    check(r.completion().query().outcome != Outcome::Succeeded, "failed output not successful");
  }
}

void partial_and_quarantine() {
  auto p = layout();
  auto a = input(p, 1), b = input(p, 2);
  std::vector<size_t> enq;
  {
    auto d = std::make_shared<HostDriver>(p);
    Session s(d);
    auto r = s.multiply(a, b);
    enq = d->trace->enqueue_calls;
    r.wait_values();
  }
  for (size_t step : enq) {
    auto d = std::make_shared<HostDriver>(p);
    Session s(d);
    d->trace->partial_step = step;
    rejects([&] { s.multiply(a, b); }, "failure after a partial submission");
    check(d->trace->executed > 0 && d->trace->live_buffers == 0 && d->trace->live_events == 0,
          "partial submission pins resources until drained");
    check(s.history().back().outcome == Outcome::Failed, "partial failure keeps its record");
  }
  auto d = std::make_shared<HostDriver>(p);
  auto trace = d->trace;
  {
    Session s(d);
    auto r = s.multiply(a, b);
    trace->async_fail_at = 3;
    trace->allow_drain = false;
    rejects([&] { r.wait_values(); }, "failed wait with unproven quiescence");
    r = {};
    check(s.stats().poisoned && !s.stats().quiescent && trace->live_buffers == 6,
          "abandoned handles cannot free quarantined buffers");
    check(s.stats().host_bytes == 3 * bytes(p), "quarantine keeps upload and download staging");
    rejects([&] { s.close(); }, "failed close keeps pins");
    check(trace->live_buffers == 6, "unproven drain cannot release");
    trace->allow_drain = true;
    rejects([&] { s.close(); }, "recovered cleanup still reports the original failure");
    check(s.stats().quiescent && trace->live_buffers == 0 && trace->live_events == 0, "proven drain reclaims");
  }
  check(trace->live_buffers == 0 && trace->live_events == 0, "nothing leaked");
}

void abandonment_and_bad_output() {
  auto p = layout();
  auto a = input(p, 1), b = input(p, 2);
  auto d = std::make_shared<HostDriver>(p);
  Download out;
  {
    Session s(d);
    out = s.multiply(a, b);
  }
  equal(out.wait_values(), naive(p, a, b), "closing with a live output drains first");
  out = {};
  check(d->trace->live_buffers == 0 && d->trace->live_events == 0, "abandonment cleanup");
  auto bad = std::make_shared<HostDriver>(p);
  Session s(bad);
  bad->trace->corrupt_download = true;
  auto r = s.multiply(a, b);
  rejects([&] { r.wait_values(); }, "noncanonical output stays hidden");
  check(s.stats().poisoned && s.history().back().outcome == Outcome::Failed, "bad output fails its completion");
}

void query_failure_and_evaluation_input() {
  auto p = layout();
  auto a = input(p, 8);
  auto d = std::make_shared<HostDriver>(p);
  Session s(d);
  auto eval = s.upload_evaluation(a);
  check(eval.layout().domain == Domain::Evaluation && eval.layout().order == Order::BitReversed,
        "evaluation upload layout");
  auto want = a;
  for (size_t k = 0; k < p.towers.size(); ++k) {
    const auto t = p.parameters->tower(p.towers[k]);
    std::vector<uint64_t> row(a.begin() + k * 8, a.begin() + (k + 1) * 8);
    ntt::NegacyclicNtt(8, t.q, t.psi).inverse(row);
    std::copy(row.begin(), row.end(), want.begin() + k * 8);
  }
  equal(s.download(s.inverse(eval)).wait_values(), want, "inverse of an evaluation upload");
  auto other = std::make_shared<HostDriver>(p);
  Session failed(other);
  auto v = failed.upload(a);
  other->trace->fail_step = other->trace->calls + 1;
  check(v.readiness().query().outcome == Outcome::Failed, "query reports driver failure");
  rejects([&] { v.readiness().wait(); }, "wait rethrows");
  check(failed.stats().poisoned, "query errors poison");
  auto cleanup = std::make_shared<HostDriver>(p);
  Session ce(cleanup);
  cleanup->trace->cleanup_error = true;
  rejects([&] { ce.close(); }, "close reports cleanup errors");
  check(ce.stats().poisoned && ce.stats().closed, "cleanup error stops admission");
}

void conversion_chain_and_failures() {
  auto p = layout(3, 2), q = p;
  q.towers = q_prefix(*p.parameters, 3);
  auto d = std::make_shared<HostDriver>(p);
  Session s(d);
  auto values = input(p, 17);
  auto up = s.upload(p, values);
  auto src = s.subview(up, 0, 3, 3);
  auto ext = s.mod_up(src);
  auto red = s.mod_down(ext);
  check(s.stats().downloads == 0 && s.stats().uploads == 1 && s.stats().parameter_uploads == 3,
        "ModUp/ModDown without intermediate transfers");
  check(d->trace->executed == 0, "conversion composition does not wait on the host");
  validate_host_values(q, s.download(red).wait_values());
  equal(s.download(up).wait_values(), values, "conversion preserves its input");
  auto up_values = s.download(ext);
  auto second = s.mod_up(s.forward(src));
  equal(s.download(second).wait_values(), {up_values.wait_values().begin(), up_values.wait_values().end()},
        "coefficient and evaluation ModUp agree");
  auto ppart = s.subview(up, 3, 2, 3);
  auto switched = s.switch_basis(ppart, q);
  const auto &qs = p.parameters->q();
  const auto &ps = p.parameters->p();
  ntt::ApproximateBasisPlan ref({ps[0].q, ps[1].q}, {qs[0].q, qs[1].q, qs[2].q});
  equal(s.download(switched).wait_values(), ref.convert(8, std::span<const uint64_t>(values).subspan(24)),
        "basis switch honours a nonzero view offset");
  // Tables are built once per session but still staged and uploaded per call.
  const size_t tables_before = s.stats().parameter_uploads;
  auto red_again = s.mod_down(ext);
  check(s.stats().parameter_uploads == tables_before + 2, "a repeated ModDown uploads both of its tables again");
  const auto red_first = s.download(red);  // owns the staging its values point into
  const auto first_values = red_first.wait_values();
  equal(s.download(red_again).wait_values(), {first_values.begin(), first_values.end()},
        "a repeated ModDown reuses its tables with the same result");
  const size_t before = d->trace->calls;
  rejects([&] { s.mod_down(up); }, "ModDown rejects coefficient input");
  rejects([&] { s.mod_up(up); }, "ModUp rejects QP input");
  rejects([&] { s.switch_basis(src, q); }, "overlapping bases");
  rejects([&] { s.switch_basis(ext, q); }, "basis switch rejects evaluation input");
  auto wrong = q;
  wrong.active_q = 2;
  wrong.towers.resize(2);
  rejects([&] { s.upload(wrong, std::span<const uint64_t>(values).first(8)); }, "subset upload length");
  check(d->trace->calls == before, "bad conversion metadata submits nothing");

  auto chain = [&](Session &x) {
    auto a = x.upload(q, std::span<const uint64_t>(values).first(24));
    return x.download(x.mod_down(x.mod_up(a)));
  };
  auto base = std::make_shared<HostDriver>(p);
  {
    Session b(base);
    chain(b).wait_values();
  }
  for (size_t step = 1; step <= base->trace->calls; ++step) {
    auto f = std::make_shared<HostDriver>(p);
    f->trace->fail_step = step;
    bool failed = false;
    {
      Session t(f);
      try {
        chain(t).wait_values();
      } catch (...) {
        failed = true;
      }
    }
    check(failed && f->trace->live_buffers == 0 && f->trace->live_events == 0,
          "failed conversion releases tables, outputs and events");
  }
  for (size_t step : base->trace->enqueue_calls) {
    auto f = std::make_shared<HostDriver>(p);
    f->trace->partial_step = step;
    bool failed = false;
    {
      Session t(f);
      try {
        chain(t).wait_values();
      } catch (...) {
        failed = true;
      }
    }
    check(failed && f->trace->live_buffers == 0 && f->trace->live_events == 0,
          "partially submitted conversion drains before release");
  }
  auto partial = std::make_shared<HostDriver>(p);
  Session t(partial);
  auto a = t.upload(q, std::span<const uint64_t>(values).first(24));
  partial->trace->allow_drain = false;
  partial->trace->partial_step = partial->trace->calls + 4;  // event, output, table, table upload
  auto p_target = p;
  p_target.towers.erase(p_target.towers.begin(), p_target.towers.begin() + 3);
  rejects([&] { t.switch_basis(a, p_target); }, "failed constant upload");
  a = {};
  check(t.stats().poisoned && partial->trace->live_buffers >= 3 && t.stats().host_bytes > 0,
        "unproven drain keeps constants, staging and input");
  partial->trace->allow_drain = true;
  rejects([&] { t.close(); }, "later drain still reports the failure");
  check(partial->trace->live_buffers == 0 && t.stats().device_bytes == 0 && t.stats().host_bytes == 0,
        "proven drain reclaims conversion resources");
  auto limited = std::make_shared<HostDriver>(p);
  Session low(limited, {400, 4096, 100});
  auto sl = low.upload(q, std::span<const uint64_t>(values).first(24));
  sl.readiness().wait();
  const size_t prior = limited->trace->calls;
  rejects([&] { low.switch_basis(sl, p_target); }, "table bytes count against the device limit");
  check(limited->trace->calls == prior && low.stats().device_bytes == 192 && !low.stats().poisoned,
        "device admission failure restores credits");
  auto hl = std::make_shared<HostDriver>(p);
  Session lh(hl, {4096, 256, 100});
  auto sh = lh.upload(q, std::span<const uint64_t>(values).first(24));
  const size_t hcalls = hl->trace->calls;
  rejects([&] { lh.switch_basis(sh, p_target); }, "table staging counts against the host limit");
  check(hl->trace->calls == hcalls && lh.stats().host_bytes == 192 && !lh.stats().poisoned,
        "host admission failure keeps earlier staging");
}

void hybrid_digit_complement() {
  const auto full = layout(3, 40);
  auto active = full;
  active.towers = q_prefix(*full.parameters, full.active_q);
  const auto coeffs = input(active, 9);
  auto d = std::make_shared<HostDriver>(full);
  Session s(d);
  const auto src = s.upload(active, coeffs);
  const size_t calls = d->trace->calls;
  rejects([&] { s.hybrid_digits(src); }, "digits need evaluation input");
  check(d->trace->calls == calls, "invalid digits input submits nothing");
  const auto digits = s.hybrid_digits(s.forward(src));
  check(digits.size() == 3 && s.stats().parameter_uploads == 7, "complements over 40 towers split in order");
  check(s.stats().downloads == 0 && d->trace->executed == 0, "digits stay queued and resident");
  const size_t n = full.parameters->n();
  for (size_t digit = 0; digit < digits.size(); ++digit) {
    std::vector<uint64_t> want;
    for (const TowerRef ref : full.towers) {
      const auto t = full.parameters->tower(ref);
      std::vector<uint64_t> row(coeffs.begin() + digit * n, coeffs.begin() + (digit + 1) * n);
      for (auto &v : row) v %= t.q;  // a one-tower digit lifts its residues directly
      ntt::NegacyclicNtt(n, t.q, t.psi).forward(row);
      want.insert(want.end(), row.begin(), row.end());
    }
    equal(s.download(digits[digit]).wait_values(), want, "complement residues and QP order");
  }
  s.close();
}

void hybrid_key_admission_and_lifetime() {
  const auto base = layout(2, 1);
  const auto eval = transform_layout(base, false);
  const auto values = input(eval, 7);
  auto d = std::make_shared<HostDriver>(base);
  {
    Session s(d);
    auto invalid = values;
    invalid.front() = base.parameters->q().front().q;
    const std::array bad{HybridKeyDigit{values, values}, HybridKeyDigit{values, invalid}};
    rejects([&] { s.upload_hybrid_key(eval, bad); }, "invalid residue in a later key digit");
    check(d->trace->calls == 0, "every key row validated before submission");
    const std::array rows{HybridKeyDigit{values, values}, HybridKeyDigit{values, values}};
    const auto key = s.upload_hybrid_key(eval, rows);
    const auto digit = s.upload_evaluation(values);
    const std::array digits{digit, digit};
    const size_t before = d->trace->calls;
    rejects([&] { s.hybrid_accumulate(std::span(digits).first(1), key); }, "missing active digit");
    check(d->trace->calls == before, "invalid MAC input submits nothing");
    d->trace->allow_drain = false;
    d->trace->partial_step = before + 5;  // event, two outputs, digit 0, fail after digit 1
    rejects([&] { s.hybrid_accumulate(digits, key); }, "MAC partial submission reported");
    check(s.stats().poisoned && !s.stats().quiescent && d->trace->live_buffers == 7,
          "undrained MAC keeps outputs, key and digit allocations");
    d->trace->allow_drain = true;
    rejects([&] { s.close(); }, "close keeps the original error after draining");
  }
  check(d->trace->live_buffers == 0 && d->trace->live_events == 0, "drained MAC releases everything");
}

void closed_and_moved_admission() {
  const auto base = layout(2, 1);
  const auto values = input(base, 1);
  auto d = std::make_shared<HostDriver>(base);
  Session moved(d);
  Session owner(std::move(moved));
  const Polynomial empty;
  const HybridKey key;
  const HybridContributions channels;
  const std::vector<std::function<void(Session &)>> admissions{
      [&](Session &x) { x.upload(values); },
      [&](Session &x) { x.upload(base, values); },
      [&](Session &x) { x.upload_evaluation(values); },
      [&](Session &x) { x.forward(empty); },
      [&](Session &x) { x.inverse(empty); },
      [&](Session &x) { x.pointwise(empty, empty); },
      [&](Session &x) { x.switch_basis(empty, base); },
      [&](Session &x) { x.mod_up(empty); },
      [&](Session &x) { x.mod_down(empty); },
      [&](Session &x) { x.hybrid_digits(empty); },
      [&](Session &x) { x.upload_hybrid_key(base, {}); },
      [&](Session &x) { x.hybrid_accumulate({}, key); },
      [&](Session &x) { x.hybrid_rotate_extended(channels, empty, 1); },
      [&](Session &x) { x.subview(empty, 0, 1, 1); },
      [&](Session &x) { x.download(empty); },
      [&](Session &x) { x.multiply(values, values); },
  };
  owner.close();
  for (Session *x : {&moved, &owner})
    for (const auto &admit : admissions) rejects([&] { admit(*x); }, "closed or moved session rejects work");
  check(d->trace->calls == 0, "rejected admission never touches the driver");
  check(owner.stats().closed && owner.history().empty(), "closed diagnostics readable");
}

// Whole KeySwitchHybrid core through the Session (digits -> MAC -> ModDown
// of both channels) at a reduced level, against the C oracle.
void hybrid_key_switch_against_oracle() {
  const size_t n = 64, M = 5, K = 3, alpha = 2, L = 4;
  const auto qs = ntt::find_primes(n, M, 40), ps = ntt::find_primes(n, K, 45);
  std::vector<Tower> q, p;
  for (auto v : qs) q.push_back({v, ntt::find_psi(v, n)});
  for (auto v : ps) p.push_back({v, ntt::find_psi(v, n)});
  auto params = Parameters::create(n, q, p, alpha);
  const PolynomialLayout base{params, L, qp_basis(*params, L), Domain::Coefficient, Order::Natural};
  const PolynomialLayout full{params, M, qp_basis(*params, M), Domain::Evaluation, Order::BitReversed};
  PolynomialLayout source{params, L, q_prefix(*params, L), Domain::Evaluation, Order::BitReversed};
  const size_t digits_full = (M + alpha - 1) / alpha;
  std::vector<uint64_t> kb((M + K) * n * digits_full), ka(kb.size()), c(L * n);
  for (size_t i = 0; i < kb.size(); ++i) {
    const size_t t = (i / n) % (M + K);
    const uint64_t m = t < M ? q[t].q : p[t - M].q;
    kb[i] = (i * 2654435761u + 17) % m;
    ka[i] = (i * 40503u + 99) % m;
  }
  for (size_t i = 0; i < c.size(); ++i) c[i] = (i * 97 + 5) % q[i / n].q;
  auto d = std::make_shared<HostDriver>(base);
  Session s(d);
  std::vector<HybridKeyDigit> rows;
  const size_t stride = (M + K) * n;
  for (size_t g = 0; g < digits_full; ++g)
    rows.push_back({std::span<const uint64_t>(kb).subspan(g * stride, stride),
                    std::span<const uint64_t>(ka).subspan(g * stride, stride)});
  const auto key = s.upload_hybrid_key(full, rows);
  const auto digits = s.hybrid_digits(s.upload(source, c));
  const auto z = s.hybrid_accumulate(digits, key);
  const auto ob = s.download(s.mod_down(z.b)), oa = s.download(s.mod_down(z.a));
  std::vector<uint64_t> qm, qr, pm, pr, wb(L * n), wa(L * n);
  for (size_t i = 0; i < L; ++i) qm.push_back(q[i].q), qr.push_back(q[i].psi);
  for (auto &t : p) pm.push_back(t.q), pr.push_back(t.psi);
  check(or_key_switch(c.data(), n, qm.data(), qr.data(), L, M, pm.data(), pr.data(), K, alpha, kb.data(), ka.data(),
                      wb.data(), wa.data()) == 0,
        "oracle key switch ran");
  equal(ob.wait_values(), wb, "Session key switch B equals the oracle");
  equal(oa.wait_values(), wa, "Session key switch A equals the oracle");
  s.close();
}

}  // namespace

int main() {
  try {
    chain_and_leases();
    rejection_and_views();
    budgets();
    failure_matrix();
    partial_and_quarantine();
    abandonment_and_bad_output();
    query_failure_and_evaluation_input();
    conversion_chain_and_failures();
    hybrid_digit_complement();
    hybrid_key_admission_and_lifetime();
    closed_and_moved_admission();
    hybrid_key_switch_against_oracle();
    std::cout << checks << " RESIDENT HOST RUNTIME CHECKS PASSED\n";
  } catch (const std::exception &e) {
    std::cerr << "FAIL: " << e.what() << '\n';
    return 1;
  }
  return 0;
}
