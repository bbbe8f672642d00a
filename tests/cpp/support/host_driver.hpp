#pragma once
// Deferred CPU transport for runtime tests (test support only; never device
// execution). Work is queued on enqueue and executed only when an event is
// waited on or the queue is drained, so tests can see what the Session
// submits. Fault injection mirrors the reference's host fake
// (tests/support/resident_host_driver.hpp:10-195): a synchronous failure at
// driver call k, a failure right after an enqueue, an asynchronous failure of
// queued item k, unproven drains, corrupted downloads and cleanup errors.
#include <functional>
#include <memory>
#include <stdexcept>
#include <vector>

#include "ntt/ntt.hpp"
#include "ntt/resident_conversion.hpp"
#include "ntt/resident_driver.hpp"

namespace host_test {
using namespace ntt::resident;
using ntt::proposal::PolynomialLayout;

struct Trace {
  size_t calls = 0, live_buffers = 0, live_events = 0, executed = 0, drains = 0;
  size_t fail_step = 0, partial_step = 0, async_fail_at = 0;
  bool allow_drain = true, corrupt_download = false, cleanup_error = false;
  std::vector<size_t> enqueue_calls;
};

struct HostBuffer final : detail::Buffer {
  std::shared_ptr<Trace> trace;
  std::vector<uint64_t> values;
  HostBuffer(std::shared_ptr<Trace> t, size_t n) : trace(std::move(t)), values(n) { ++trace->live_buffers; }
  ~HostBuffer() override { --trace->live_buffers; }
};

struct HostEvent final : detail::Event {
  std::shared_ptr<Trace> trace;
  size_t fence = 0;
  bool recorded = false;
  explicit HostEvent(std::shared_ptr<Trace> t) : trace(std::move(t)) { ++trace->live_events; }
  ~HostEvent() override { --trace->live_events; }
};

class HostDriver final : public detail::Driver {
  PolynomialLayout layout_;
  std::vector<std::function<void()>> queue_;
  size_t done_ = 0;
  bool failed_ = false;

  void tick() {
    if (++trace->calls == trace->fail_step) throw std::runtime_error("injected transport failure");
  }
  void enqueue(std::function<void()> work) {
    tick();
    queue_.push_back(std::move(work));
    trace->enqueue_calls.push_back(trace->calls);
    if (trace->calls == trace->partial_step) throw std::runtime_error("failure after enqueue");
  }
  void run_until(size_t fence) {
    if (failed_) throw std::runtime_error("host test queue failed");
    while (done_ < fence) {
      if (done_ + 1 == trace->async_fail_at) {
        failed_ = true;
        throw std::runtime_error("injected asynchronous failure");
      }
      queue_[done_]();
      queue_[done_] = {};
      ++done_;
      ++trace->executed;
    }
  }
  static HostBuffer &hb(detail::Buffer &b) { return dynamic_cast<HostBuffer &>(b); }
  static const HostBuffer &hb(const detail::Buffer &b) { return dynamic_cast<const HostBuffer &>(b); }

 public:
  std::shared_ptr<Trace> trace = std::make_shared<Trace>();
  explicit HostDriver(PolynomialLayout p) : layout_(std::move(p)) { ntt::proposal::validate(layout_); }
  int ordinal() const override { return 0; }
  const PolynomialLayout &layout() const override { return layout_; }
  size_t constant_bytes() const override { return 0; }
  std::unique_ptr<detail::Buffer> allocate(size_t n) override {
    tick();
    return std::make_unique<HostBuffer>(trace, n);
  }
  std::unique_ptr<detail::Event> make_event() override {
    tick();
    return std::make_unique<HostEvent>(trace);
  }
  void upload(detail::Buffer &out, std::span<const uint64_t> src) override {
    auto *b = &hb(out);
    enqueue([b, src] { std::copy(src.begin(), src.end(), b->values.begin()); });
  }
  void copy(detail::Buffer &out, const detail::Buffer &in, size_t offset, size_t n) override {
    auto *a = &hb(out);
    auto *b = &hb(in);
    enqueue([a, b, offset, n] { std::copy_n(b->values.begin() + offset, n, a->values.begin()); });
  }
  void copy_range(detail::Buffer &out, size_t dst, const detail::Buffer &in, size_t src, size_t n) override {
    auto *a = &hb(out);
    auto *b = &hb(in);
    ntt::proposal::require(dst <= a->values.size() && n <= a->values.size() - dst && src <= b->values.size() &&
                               n <= b->values.size() - src,
                           "host copy range extent");
    enqueue([a, b, dst, src, n] { std::copy_n(b->values.begin() + src, n, a->values.begin() + dst); });
  }
  void transform(detail::Buffer &out, const PolynomialLayout &p, bool inverse) override {
    auto *b = &hb(out);
    enqueue([b, p, inverse] {
      const size_t n = p.parameters->n();
      for (size_t k = 0; k < p.towers.size(); ++k) {
        const auto t = p.parameters->tower(p.towers[k]);
        std::vector<uint64_t> row(b->values.begin() + k * n, b->values.begin() + (k + 1) * n);
        ntt::NegacyclicNtt plan(n, t.q, t.psi);
        inverse ? plan.inverse(row) : plan.forward(row);
        std::copy(row.begin(), row.end(), b->values.begin() + k * n);
      }
    });
  }
  void pointwise(detail::Buffer &out, const detail::Buffer &right, size_t offset, const PolynomialLayout &p) override {
    auto *a = &hb(out);
    auto *b = &hb(right);
    enqueue([a, b, offset, p] {
      const size_t n = p.parameters->n();
      for (size_t k = 0; k < p.towers.size(); ++k) {
        const uint64_t q = p.parameters->tower(p.towers[k]).q;
        for (size_t i = 0; i < n; ++i)
          a->values[k * n + i] = ntt::mul_mod(a->values[k * n + i], b->values[offset + k * n + i], q);
      }
    });
  }
  void hybrid_mac(detail::Buffer &rb, detail::Buffer &ra, const detail::Buffer &digit, const detail::Buffer &kb,
                  const detail::Buffer &ka, const PolynomialLayout &layout, bool first) override {
    auto *b = &hb(rb), *a = &hb(ra);
    const auto *d = &hb(digit), *k1 = &hb(kb), *k2 = &hb(ka);
    enqueue([b, a, d, k1, k2, layout, first] {
      const size_t n = layout.parameters->n();
      for (size_t i = 0; i < d->values.size(); ++i) {
        const uint64_t q = layout.parameters->tower(layout.towers[i / n]).q;
        detail::hybrid_mac_value(b->values[i], d->values[i], k1->values[i], q, first);
        detail::hybrid_mac_value(a->values[i], d->values[i], k2->values[i], q, first);
      }
    });
  }
  void hybrid_rotate(detail::Buffer &ob, detail::Buffer &oa, const detail::Buffer &eb, const detail::Buffer &ea,
                     const detail::Buffer &c0, size_t c0_offset, const detail::Buffer &scales,
                     const PolynomialLayout &layout, uint32_t k) override {
    auto *b = &hb(ob), *a = &hb(oa);
    const auto *xb = &hb(eb), *xa = &hb(ea), *first = &hb(c0), *table = &hb(scales);
    enqueue([=] {
      const size_t n = layout.parameters->n(), lg = ntt::log2_of_pow2(n);
      for (size_t t = 0; t < layout.towers.size(); ++t) {
        const uint64_t q = layout.parameters->tower(layout.towers[t]).q;
        for (size_t slot = 0; slot < n; ++slot) {
          const size_t src = t * n + detail::automorphism_source_index(slot, n, lg, k);
          b->values[t * n + slot] = detail::rotated_b_value(xb->values[src], first->values.data() + c0_offset, src, t,
                                                            layout.active_q, table->values.data(), q);
          a->values[t * n + slot] = xa->values[src];
        }
      }
    });
  }
  void switch_basis(detail::Buffer &out, const detail::Buffer &in, size_t offset, const detail::Buffer &constants,
                    const PolynomialLayout &source, const PolynomialLayout &target) override {
    auto *a = &hb(out);
    const auto *b = &hb(in), *table = &hb(constants);
    enqueue([a, b, table, offset, source, target] {
      for (size_t i = 0; i < a->values.size(); ++i)
        a->values[i] = detail::resident_basis_value(b->values.data() + offset, source.parameters->n(), i,
                                                    source.towers.size(), target.towers.size(), table->values.data());
    });
  }
  void correct(detail::Buffer &out, const detail::Buffer &left, size_t lo, const detail::Buffer &right, size_t ro,
               const detail::Buffer &constants, const PolynomialLayout &layout) override {
    auto *a = &hb(out);
    const auto *b = &hb(left), *c = &hb(right), *table = &hb(constants);
    enqueue([a, b, c, table, lo, ro, layout] {
      for (size_t i = 0; i < a->values.size(); ++i)
        a->values[i] = detail::resident_correction_value(b->values[lo + i], c->values[ro + i],
                                                         i / layout.parameters->n(), layout.towers.size(),
                                                         table->values.data());
    });
  }
  void download(std::span<uint64_t> dst, const detail::Buffer &in, size_t offset) override {
    const auto *b = &hb(in);
    const bool corrupt = trace->corrupt_download;
    enqueue([b, dst, offset, corrupt] {
      std::copy_n(b->values.begin() + offset, dst.size(), dst.begin());
      if (corrupt) dst[0] = UINT64_MAX;
    });
  }
  void record(detail::Event &e) override {
    tick();
    auto &he = dynamic_cast<HostEvent &>(e);
    he.fence = queue_.size();
    he.recorded = true;
  }
  bool ready(detail::Event &e) override {
    tick();
    auto &he = dynamic_cast<HostEvent &>(e);
    if (failed_) throw std::runtime_error("host test queue failed");
    return he.recorded && done_ >= he.fence;
  }
  void wait(detail::Event &e) override {
    tick();
    auto &he = dynamic_cast<HostEvent &>(e);
    if (!he.recorded) throw std::runtime_error("unrecorded host test event");
    run_until(he.fence);
  }
  bool quiesce() noexcept override {
    ++trace->drains;
    if (!trace->allow_drain) return false;
    try {
      run_until(queue_.size());
    } catch (...) {
    }
    // A synchronous fake can prove cancellation of the rest of its queue.
    queue_.clear();
    done_ = 0;
    return true;
  }
  void check_health() const override {
    if (trace->cleanup_error) throw std::runtime_error("injected cleanup error");
  }
};

}  // namespace host_test
