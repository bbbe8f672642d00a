#pragma once
// Deferred CPU transport for the Session runtime tests (test infrastructure:
// no device work). Every driver call is logged by kind, enqueued work runs
// only when an event is waited on, polled past or the transport is drained,
// and the arithmetic is done by the C oracle (oracle/ntt_oracle.c), so a
// Session chain on this transport is compared with the oracle's composites.
//
// Faults are chosen by a predicate over (call index, call kind): throw before
// the call acts, throw after its work was queued, or queue work that throws
// when it runs. Further switches: an unprovable quiesce, downloads that come
// back non-canonical and a health check that fails at close.
#include <cstdint>
#include <deque>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "ntt/resident_driver.hpp"
#include "support/oracle_api.h"

namespace replay {
using namespace ntt::resident;
using ntt::proposal::PolynomialLayout;

enum class Call {
  Allocate, MakeEvent, Upload, Copy, CopyRange, Transform, Pointwise, Mac, Rotate, Switch, Correct, Download,
  Record, Ready, Wait
};

inline const char *name(Call c) {
  static const char *names[] = {"allocate", "event", "upload", "copy", "copy_range", "transform", "pointwise",
                                "mac", "rotate", "switch", "correct", "download", "record", "ready", "wait"};
  return names[int(c)];
}

enum class Fault { None, Throw, ThrowAfterQueue, FailWhenRun };

struct Log {
  std::vector<Call> calls;
  size_t live_buffers = 0, live_events = 0, jobs_run = 0, quiesce_calls = 0;
  std::function<Fault(size_t, Call)> fault;  // index is 1-based over all calls
  bool quiesce_ok = true, corrupt_downloads = false, unhealthy = false;

  size_t count(Call c) const {
    size_t k = 0;
    for (Call x : calls) k += x == c;
    return k;
  }
  // calls since position `from`, as names (for order assertions)
  std::string since(size_t from) const {
    std::string s;
    for (size_t i = from; i < calls.size(); ++i) s += std::string(s.empty() ? "" : " ") + name(calls[i]);
    return s;
  }
};

struct Words final : detail::Buffer {
  std::shared_ptr<Log> log;
  std::vector<uint64_t> v;
  Words(std::shared_ptr<Log> l, size_t n) : log(std::move(l)), v(n) { ++log->live_buffers; }
  ~Words() override { --log->live_buffers; }
};

struct Mark final : detail::Event {
  std::shared_ptr<Log> log;
  size_t fence = 0;
  bool recorded = false;
  explicit Mark(std::shared_ptr<Log> l) : log(std::move(l)) { ++log->live_events; }
  ~Mark() override { --log->live_events; }
};

class Transport final : public detail::Driver {
  struct Job {
    std::function<void()> work;
    bool fails;
  };
  PolynomialLayout basis_;
  std::deque<Job> jobs_;
  size_t base_ = 0;  // index of jobs_.front() in the stream
  bool broken_ = false;

  static Words &w(detail::Buffer &b) { return dynamic_cast<Words &>(b); }
  static const Words &w(const detail::Buffer &b) { return dynamic_cast<const Words &>(b); }

  Fault enter(Call c) {
    log->calls.push_back(c);
    const Fault f = log->fault ? log->fault(log->calls.size(), c) : Fault::None;
    if (f == Fault::Throw) throw std::runtime_error(std::string("injected failure in ") + name(c));
    return f;
  }
  void queue(Call c, std::function<void()> work) {
    const Fault f = enter(c);
    jobs_.push_back({std::move(work), f == Fault::FailWhenRun});
    if (f == Fault::ThrowAfterQueue) throw std::runtime_error(std::string("injected failure after ") + name(c));
  }
  void run_to(size_t fence) {
    if (broken_) throw std::runtime_error("replay stream failed earlier");
    while (base_ < fence) {
      Job job = std::move(jobs_.front());
      jobs_.pop_front();
      ++base_;
      if (job.fails) {
        broken_ = true;
        throw std::runtime_error("injected failure while running queued work");
      }
      job.work();
      ++log->jobs_run;
    }
  }
  static uint64_t modulus(const PolynomialLayout &p, size_t tower) { return p.parameters->tower(p.towers[tower]).q; }
  static uint64_t mulmod(uint64_t a, uint64_t b, uint64_t q) { return (uint64_t)((unsigned __int128)a * b % q); }

 public:
  std::shared_ptr<Log> log = std::make_shared<Log>();
  explicit Transport(PolynomialLayout basis) : basis_(std::move(basis)) { ntt::proposal::validate(basis_); }
  size_t queued() const { return jobs_.size(); }

  int ordinal() const override { return 0; }
  const PolynomialLayout &layout() const override { return basis_; }
  size_t constant_bytes() const override { return 0; }
  std::unique_ptr<detail::Buffer> allocate(size_t n) override {
    enter(Call::Allocate);
    return std::make_unique<Words>(log, n);
  }
  std::unique_ptr<detail::Event> make_event() override {
    enter(Call::MakeEvent);
    return std::make_unique<Mark>(log);
  }
  void upload(detail::Buffer &out, std::span<const uint64_t> from) override {
    Words *o = &w(out);
    queue(Call::Upload, [o, from] { std::copy(from.begin(), from.end(), o->v.begin()); });
  }
  void copy(detail::Buffer &out, const detail::Buffer &in, size_t offset, size_t n) override {
    Words *o = &w(out);
    const Words *i = &w(in);
    queue(Call::Copy, [o, i, offset, n] { std::copy_n(i->v.begin() + offset, n, o->v.begin()); });
  }
  void copy_range(detail::Buffer &out, size_t at, const detail::Buffer &in, size_t from, size_t n) override {
    Words *o = &w(out);
    const Words *i = &w(in);
    if (at + n > o->v.size() || from + n > i->v.size()) throw std::runtime_error("copy_range out of bounds");
    queue(Call::CopyRange, [o, i, at, from, n] { std::copy_n(i->v.begin() + from, n, o->v.begin() + at); });
  }
  void transform(detail::Buffer &buf, const PolynomialLayout &p, bool inverse) override {
    Words *o = &w(buf);
    queue(Call::Transform, [o, p, inverse] {
      std::vector<uint64_t> q, psi;
      for (const auto t : p.towers) {
        q.push_back(p.parameters->tower(t).q);
        psi.push_back(p.parameters->tower(t).psi);
      }
      if (or_ntt_towers(o->v.data(), p.parameters->n(), q.size(), q.data(), psi.data(), inverse ? 1 : 0))
        throw std::runtime_error("oracle transform rejected its shape");
    });
  }
  void pointwise(detail::Buffer &out, const detail::Buffer &right, size_t offset, const PolynomialLayout &p) override {
    Words *o = &w(out);
    const Words *r = &w(right);
    queue(Call::Pointwise, [o, r, offset, p] {
      const size_t n = p.parameters->n();
      for (size_t i = 0; i < p.towers.size() * n; ++i)
        o->v[i] = mulmod(o->v[i], r->v[offset + i], modulus(p, i / n));
    });
  }
  void hybrid_mac(detail::Buffer &rb, detail::Buffer &ra, const detail::Buffer &digit, const detail::Buffer &kb,
                  const detail::Buffer &ka, const PolynomialLayout &p, bool first) override {
    Words *b = &w(rb), *a = &w(ra);
    const Words *d = &w(digit), *xb = &w(kb), *xa = &w(ka);
    queue(Call::Mac, [=] {
      const size_t n = p.parameters->n();
      for (size_t i = 0; i < d->v.size(); ++i) {
        const uint64_t q = modulus(p, i / n);
        const uint64_t pb = mulmod(d->v[i], xb->v[i], q), pa = mulmod(d->v[i], xa->v[i], q);
        b->v[i] = first ? pb : (b->v[i] + pb) % q;
        a->v[i] = first ? pa : (a->v[i] + pa) % q;
      }
    });
  }
  void hybrid_rotate(detail::Buffer &ob, detail::Buffer &oa, const detail::Buffer &eb, const detail::Buffer &ea,
                     const detail::Buffer &c0, size_t c0_offset, const detail::Buffer &scales,
                     const PolynomialLayout &p, uint32_t k) override {
    Words *b = &w(ob), *a = &w(oa);
    const Words *xb = &w(eb), *xa = &w(ea), *c = &w(c0), *sc = &w(scales);
    queue(Call::Rotate, [=] {
      const size_t n = p.parameters->n();
      for (size_t t = 0; t < p.towers.size(); ++t) {
        const uint64_t q = modulus(p, t);
        for (size_t j = 0; j < n; ++j) {
          const size_t src = t * n + or_automorphism_source(j, n, k);
          uint64_t v = xb->v[src];
          if (t < p.active_q) v = (v + mulmod(c->v[c0_offset + src], sc->v[t], q)) % q;
          b->v[t * n + j] = v;
          a->v[t * n + j] = xa->v[src];
        }
      }
    });
  }
  void switch_basis(detail::Buffer &out, const detail::Buffer &in, size_t offset, const detail::Buffer &constants,
                    const PolynomialLayout &source, const PolynomialLayout &target) override {
    Words *o = &w(out);
    const Words *i = &w(in), *table = &w(constants);
    queue(Call::Switch, [o, i, table, offset, source, target] {
      // the table opens with the source then the target moduli: the oracle
      // rebuilds the plan from exactly those
      const size_t s = source.towers.size(), t = target.towers.size();
      for (size_t k = 0; k < s; ++k)
        if (table->v[k] != modulus(source, k)) throw std::runtime_error("conversion table: wrong source modulus");
      for (size_t k = 0; k < t; ++k)
        if (table->v[s + k] != modulus(target, k)) throw std::runtime_error("conversion table: wrong target modulus");
      if (or_basis_switch(i->v.data() + offset, source.parameters->n(), table->v.data(), s, table->v.data() + s, t,
                          o->v.data()))
        throw std::runtime_error("oracle rejected the conversion");
    });
  }
  void correct(detail::Buffer &out, const detail::Buffer &left, size_t lo, const detail::Buffer &right, size_t ro,
               const detail::Buffer &constants, const PolynomialLayout &p) override {
    Words *o = &w(out);
    const Words *l = &w(left), *r = &w(right), *table = &w(constants);
    queue(Call::Correct, [=] {
      const size_t n = p.parameters->n(), T = p.towers.size();
      for (size_t i = 0; i < T * n; ++i) {
        const uint64_t q = table->v[i / n], pinv = table->v[T + i / n];
        const uint64_t d = l->v[lo + i] >= r->v[ro + i] ? l->v[lo + i] - r->v[ro + i] : l->v[lo + i] + q - r->v[ro + i];
        o->v[i] = mulmod(d, pinv, q);
      }
    });
  }
  void download(std::span<uint64_t> to, const detail::Buffer &in, size_t offset) override {
    const Words *i = &w(in);
    const bool corrupt = log->corrupt_downloads;
    queue(Call::Download, [i, to, offset, corrupt] {
      std::copy_n(i->v.begin() + offset, to.size(), to.begin());
      if (corrupt && !to.empty()) to[to.size() - 1] = ~uint64_t{0};
    });
  }
  void record(detail::Event &e) override {
    enter(Call::Record);
    Mark &m = dynamic_cast<Mark &>(e);
    m.fence = base_ + jobs_.size();
    m.recorded = true;
  }
  bool ready(detail::Event &e) override {
    enter(Call::Ready);
    if (broken_) throw std::runtime_error("replay stream failed earlier");
    const Mark &m = dynamic_cast<Mark &>(e);
    return m.recorded && base_ >= m.fence;
  }
  void wait(detail::Event &e) override {
    enter(Call::Wait);
    const Mark &m = dynamic_cast<Mark &>(e);
    if (!m.recorded) throw std::runtime_error("wait on an unrecorded event");
    run_to(m.fence);
  }
  bool quiesce() noexcept override {
    ++log->quiesce_calls;
    if (!log->quiesce_ok) return false;
    // a host stream can always prove it is empty: run what still can, drop the rest
    try {
      run_to(base_ + jobs_.size());
    } catch (...) {
    }
    base_ += jobs_.size();
    jobs_.clear();
    return true;
  }
  void check_health() const override {
    if (log->unhealthy) throw std::runtime_error("injected health failure");
  }
  // tests: run queued work as a device would, without an event; a failure is
  // kept for the next driver call to report, as a device reports it
  void run_all() {
    try {
      run_to(base_ + jobs_.size());
    } catch (const std::runtime_error &) {
    }
  }
};

}  // namespace replay
