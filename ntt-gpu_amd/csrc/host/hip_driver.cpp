// Device transport of the resident Session: implements the Driver seam
// (include/ntt/resident_driver.hpp; reference src/cuda/resident_adapter.inc:60-317)
// on top of the C ABI of libntt_hip.so. One context (device twiddle tables
// for the session basis) and its ordered stream per driver. Plain C++: no
// HIP header is needed on this side of the boundary.
#include <algorithm>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "ntt/resident_driver.hpp"
#include "ntt_hip.h"

namespace ntt::hip_transport {
using namespace ::ntt::proposal;
namespace transport = ::ntt::resident::detail;

[[noreturn]] void raise(int status, const char *what) {
  const std::string msg = std::string(what) + ": " + ntt_hip_last_error();
  if (status == NTT_HIP_EINVAL) throw ContractError(msg);
  throw std::runtime_error("HIP: " + msg);
}
inline void ok(int status, const char *what) {
  if (status != NTT_HIP_OK) raise(status, what);
}

// Bytes of device constants a session basis needs: forward and inverse
// {W, Shoup(W)} tables plus one per-prime record (ntt_hip_context_create).
inline size_t table_bytes(const PolynomialLayout &p) {
  return checked_product(p.towers.size(), checked_product(static_cast<size_t>(p.parameters->n()), 32) + 128);
}

struct Context {
  PolynomialLayout base;
  int device;
  ntt_hip_context *ctx = nullptr;
  bool cleanup_failed = false;
  Context(PolynomialLayout p, int ordinal) : base(std::move(p)), device(ordinal) {
    std::vector<uint64_t> q, psi;
    for (const TowerRef ref : base.towers) {
      q.push_back(base.parameters->tower(ref).q);
      psi.push_back(base.parameters->tower(ref).psi);
    }
    ok(ntt_hip_context_create(&ctx, device, base.parameters->n(), q.data(), psi.data(), q.size()), "context");
  }
  ~Context() {
    flush_pool();
    for (void *e : free_events)
      if (ntt_hip_event_destroy(device, e) != NTT_HIP_OK) cleanup_failed = true;
    if (ctx && ntt_hip_context_destroy(ctx) != NTT_HIP_OK) cleanup_failed = true;
  }

  // Device blocks and events are recycled. The Session releases a buffer or
  // an event only after the operation that used it is proven complete, so a
  // released block can serve the next output of the same size without a
  // hipMalloc/hipFree pair per operation (the reference allocates every
  // output afresh). At most kPoolBytes stay cached; an allocation that fails
  // flushes the cache and retries once.
  static constexpr size_t kPoolBytes = size_t{8} << 30;
  std::multimap<size_t, void *> free_blocks;  // bytes -> block
  size_t pooled_bytes = 0;
  std::vector<void *> free_events;

  void *take_block(size_t bytes) {
    const auto it = free_blocks.find(bytes);
    if (it != free_blocks.end()) {
      void *p = it->second;
      free_blocks.erase(it);
      pooled_bytes -= bytes;
      return p;
    }
    void *p = nullptr;
    if (ntt_hip_malloc(device, &p, bytes) != NTT_HIP_OK) {
      flush_pool();
      ok(ntt_hip_malloc(device, &p, bytes), "allocate");
    }
    return p;
  }
  void give_block(void *p, size_t bytes) {
    if (pooled_bytes + bytes <= kPoolBytes) {
      free_blocks.emplace(bytes, p);
      pooled_bytes += bytes;
    } else if (ntt_hip_free(device, p) != NTT_HIP_OK) {
      cleanup_failed = true;
    }
  }
  void flush_pool() {
    for (auto &[bytes, p] : free_blocks)
      if (ntt_hip_free(device, p) != NTT_HIP_OK) cleanup_failed = true;
    free_blocks.clear();
    pooled_bytes = 0;
  }
  void *take_event() {
    if (!free_events.empty()) {
      void *e = free_events.back();
      free_events.pop_back();
      return e;
    }
    void *e = nullptr;
    ok(ntt_hip_event_create(device, &e, 0), "event");
    return e;
  }
  void *stream() const { return ntt_hip_context_stream(ctx); }

  // Session::transform and pointwise copy their input into a fresh output and
  // then work on that output in place. The driver holds such a copy back until
  // its next call: when that call transforms the copy's destination, the copy
  // becomes the transform's out-of-place first pass (one pass over the data
  // instead of two); any other call, and the release of either buffer,
  // enqueues it first, so stream order is unchanged.
  struct PendingCopy {
    uint64_t *dst = nullptr;
    const uint64_t *src = nullptr;
    size_t n = 0;
  } pending;
  void flush_copy() {
    if (!pending.dst) return;
    const PendingCopy p = pending;
    pending = {};
    ok(ntt_hip_memcpy(device, p.dst, p.src, p.n * sizeof(uint64_t), NTT_HIP_D2D, stream()), "copy");
  }

  // Session::hybrid_accumulate issues one hybrid_mac per digit into the same
  // two outputs, back to back. The driver holds them back the same way and
  // enqueues the run as one pass over all its digits (b and a read and
  // written once instead of once per digit), on the first call that is not
  // the next digit of the run, or at the release of any buffer it reads.
  struct PendingMac {
    uint64_t *b = nullptr, *a = nullptr;
    size_t first = 0, towers = 0;
    bool first_digit = false;
    std::vector<const uint64_t *> digits, key_b, key_a;
  } mac;
  void flush_mac() {
    if (!mac.b) return;
    PendingMac m = std::move(mac);
    mac = {};
    ok(ntt_hip_hybrid_mac_digits(ctx, m.b, m.a, m.digits.data(), m.key_b.data(), m.key_a.data(), m.digits.size(),
                                 m.first, m.towers, m.first_digit, stream()),
       "hybrid_mac");
  }
  // Everything held back, in stream order (at most one of the two is pending).
  void flush_pending() {
    flush_copy();
    flush_mac();
  }
  bool pending_touches(const uint64_t *block, size_t elements) const {
    const auto in = [&](const uint64_t *p) { return p >= block && p < block + elements; };
    if (pending.dst && (in(pending.dst) || in(pending.src))) return true;
    if (!mac.b) return false;
    if (in(mac.b) || in(mac.a)) return true;
    for (size_t j = 0; j < mac.digits.size(); ++j)
      if (in(mac.digits[j]) || in(mac.key_b[j]) || in(mac.key_a[j])) return true;
    return false;
  }
  // Position of a contiguous layout inside the session basis.
  size_t first(const PolynomialLayout &p) const {
    validate(p);
    require(*p.parameters == *base.parameters && p.active_q <= base.active_q, "foreign resident HIP parameters");
    const auto it = std::find(base.towers.begin(), base.towers.end(), p.towers.front());
    const size_t at = static_cast<size_t>(it - base.towers.begin());
    require(at < base.towers.size() && p.towers.size() <= base.towers.size() - at &&
                std::equal(p.towers.begin(), p.towers.end(), it),
            "noncontiguous HIP plan view");
    return at;
  }
  std::vector<size_t> positions(const std::vector<TowerRef> &refs) const {
    std::vector<size_t> out;
    for (const TowerRef r : refs) {
      const auto it = std::find(base.towers.begin(), base.towers.end(), r);
      require(it != base.towers.end(), "tower outside the session basis");
      out.push_back(static_cast<size_t>(it - base.towers.begin()));
    }
    return out;
  }
  std::vector<size_t> p_positions() const {
    std::vector<TowerRef> refs;
    for (size_t j = 0; j < base.parameters->p().size(); ++j) refs.push_back({Family::P, j});
    return positions(refs);
  }
};

struct Buffer final : transport::Buffer {
  std::shared_ptr<Context> context;
  uint64_t *data = nullptr;
  size_t elements;
  Buffer(std::shared_ptr<Context> c, size_t n) : context(std::move(c)), elements(n) {
    if (context->cleanup_failed) throw std::runtime_error("resident HIP cleanup failed");
    data = static_cast<uint64_t *>(context->take_block(bytes()));
  }
  ~Buffer() override {
    if (context->pending_touches(data, elements)) {
      try {
        context->flush_pending();  // enqueued before the block can be reused
      } catch (...) {
        context->pending = {};
        context->mac = {};
        context->cleanup_failed = true;
      }
    }
    context->give_block(data, bytes());
  }
  size_t bytes() const { return std::max<size_t>(1, elements) * sizeof(uint64_t); }
};

struct Event final : transport::Event {
  std::shared_ptr<Context> context;
  void *event = nullptr;
  bool recorded = false;
  explicit Event(std::shared_ptr<Context> c) : context(std::move(c)) {
    if (context->cleanup_failed) throw std::runtime_error("resident HIP cleanup failed");
    event = context->take_event();
  }
  ~Event() override { context->free_events.push_back(event); }
};

class Driver final : public transport::Driver {
  std::shared_ptr<Context> c_;
  const Buffer &buf(const transport::Buffer &b) const {
    const auto &out = dynamic_cast<const Buffer &>(b);
    require(out.context == c_, "foreign HIP buffer");
    return out;
  }
  Event &evt(transport::Event &e) const {
    auto &out = dynamic_cast<Event &>(e);
    require(out.context == c_, "foreign HIP event");
    return out;
  }
  void healthy() const {
    if (c_->cleanup_failed) throw std::runtime_error("resident HIP cleanup failed");
  }
  void flush() { c_->flush_pending(); }

 public:
  Driver(PolynomialLayout p, int ordinal) : c_(std::make_shared<Context>(std::move(p), ordinal)) {}
  int ordinal() const override { return c_->device; }
  const PolynomialLayout &layout() const override { return c_->base; }
  size_t constant_bytes() const override { return table_bytes(c_->base); }
  void check_health() const override { healthy(); }
  std::unique_ptr<transport::Buffer> allocate(size_t n) override { return std::make_unique<Buffer>(c_, n); }
  std::unique_ptr<transport::Event> make_event() override { return std::make_unique<Event>(c_); }

  void upload(transport::Buffer &out, std::span<const uint64_t> src) override {
    const auto &a = buf(out);
    require(a.elements == src.size(), "HIP upload extent");
    healthy();
    flush();
    ok(ntt_hip_memcpy(c_->device, a.data, src.data(), src.size_bytes(), NTT_HIP_H2D, c_->stream()), "upload");
  }
  void copy(transport::Buffer &out, const transport::Buffer &in, size_t offset, size_t n) override {
    const auto &a = buf(out), &b = buf(in);
    require(a.elements == n && offset <= b.elements && n <= b.elements - offset, "HIP copy extent");
    healthy();
    flush();
    c_->pending = {a.data, b.data + offset, n};
  }
  void copy_range(transport::Buffer &out, size_t dst, const transport::Buffer &in, size_t src, size_t n) override {
    const auto &a = buf(out), &b = buf(in);
    require(dst <= a.elements && n <= a.elements - dst && src <= b.elements && n <= b.elements - src,
            "HIP range copy extent");
    healthy();
    flush();
    ok(ntt_hip_memcpy(c_->device, a.data + dst, b.data + src, n * sizeof(uint64_t), NTT_HIP_D2D, c_->stream()),
       "copy_range");
  }
  void transform(transport::Buffer &out, const PolynomialLayout &p, bool inverse) override {
    const auto &a = buf(out);
    require(a.elements == bytes(p) / sizeof(uint64_t), "HIP transform extent");
    healthy();
    const uint64_t *src = nullptr;
    if (c_->pending.dst == a.data && c_->pending.n == a.elements) {
      src = c_->pending.src;  // the held-back copy becomes the first pass's input
      c_->pending = {};
    } else {
      flush();
    }
    ok(ntt_hip_transform(c_->ctx, a.data, src, c_->first(p), p.towers.size(), inverse, c_->stream()), "transform");
  }
  void pointwise(transport::Buffer &out, const transport::Buffer &right, size_t offset,
                 const PolynomialLayout &p) override {
    const auto &a = buf(out), &b = buf(right);
    require(a.elements == bytes(p) / sizeof(uint64_t) && offset <= b.elements && a.elements <= b.elements - offset,
            "HIP pointwise extent");
    healthy();
    flush();
    ok(ntt_hip_pointwise(c_->ctx, a.data, b.data + offset, c_->first(p), p.towers.size(), c_->stream()),
       "pointwise");
  }
  void hybrid_mac(transport::Buffer &rb, transport::Buffer &ra, const transport::Buffer &digit,
                  const transport::Buffer &kb, const transport::Buffer &ka, const PolynomialLayout &layout,
                  bool first_digit) override {
    const auto &b = buf(rb), &a = buf(ra), &d = buf(digit), &k1 = buf(kb), &k2 = buf(ka);
    const size_t count = bytes(layout) / sizeof(uint64_t);
    require(b.elements == count && a.elements == count && d.elements == count && k1.elements == count &&
                k2.elements == count,
            "HIP Hybrid MAC extent");
    healthy();
    c_->flush_copy();
    const size_t first = c_->first(layout);
    auto &m = c_->mac;
    const bool next_digit = m.b == b.data && m.a == a.data && m.first == first && m.towers == layout.towers.size() &&
                            !first_digit;
    if (!next_digit) {
      c_->flush_mac();
      m.b = b.data;
      m.a = a.data;
      m.first = first;
      m.towers = layout.towers.size();
      m.first_digit = first_digit;
    }
    m.digits.push_back(d.data);
    m.key_b.push_back(k1.data);
    m.key_a.push_back(k2.data);
  }
  void hybrid_rotate(transport::Buffer &ob, transport::Buffer &oa, const transport::Buffer &eb,
                     const transport::Buffer &ea, const transport::Buffer &c0, size_t c0_offset,
                     const transport::Buffer &scales, const PolynomialLayout &layout, uint32_t automorphism) override {
    const auto &b = buf(ob), &a = buf(oa), &xb = buf(eb), &xa = buf(ea), &first = buf(c0), &table = buf(scales);
    const size_t n = layout.parameters->n(), elements = n * layout.towers.size();
    require(b.elements == elements && a.elements == elements && xb.elements == elements && xa.elements == elements &&
                c0_offset <= first.elements && layout.active_q * n <= first.elements - c0_offset &&
                table.elements == layout.active_q,
            "HIP Hybrid rotation extent");
    healthy();
    flush();
    const auto p = c_->p_positions();
    ok(ntt_hip_hybrid_rotate(c_->ctx, b.data, a.data, xb.data, xa.data, first.data + c0_offset, c_->first(layout),
                             layout.towers.size(), layout.active_q, p.data(), p.size(), automorphism, c_->stream()),
       "hybrid_rotate");
  }
  void switch_basis(transport::Buffer &out, const transport::Buffer &in, size_t offset,
                    const transport::Buffer &constants, const PolynomialLayout &source,
                    const PolynomialLayout &target) override {
    const auto &a = buf(out), &b = buf(in), &table = buf(constants);
    const size_t n = target.parameters->n(), s = source.towers.size(), t = target.towers.size();
    require(a.elements == n * t && offset <= b.elements && n * s <= b.elements - offset &&
                table.elements == 3 * s + t + s * t,
            "HIP basis extent");
    healthy();
    flush();
    // The context derives (and caches) its Montgomery-form plan for the same
    // source/target primes; the session's packed table is the shape check.
    const auto src = c_->positions(source.towers), dst = c_->positions(target.towers);
    ok(ntt_hip_switch_basis(c_->ctx, a.data, b.data + offset, src.data(), s, dst.data(), t, c_->stream()),
       "switch_basis");
  }
  void correct(transport::Buffer &out, const transport::Buffer &left, size_t left_offset,
               const transport::Buffer &right, size_t right_offset, const transport::Buffer &constants,
               const PolynomialLayout &p) override {
    const auto &a = buf(out), &b = buf(left), &c = buf(right), &table = buf(constants);
    require(a.elements == p.parameters->n() * p.towers.size() && left_offset <= b.elements &&
                a.elements <= b.elements - left_offset && right_offset <= c.elements &&
                a.elements <= c.elements - right_offset && table.elements == 2 * p.towers.size(),
            "HIP correction extent");
    healthy();
    flush();
    const auto pp = c_->p_positions();
    ok(ntt_hip_correct(c_->ctx, a.data, b.data + left_offset, c.data + right_offset, c_->first(p), p.towers.size(),
                       pp.data(), pp.size(), c_->stream()),
       "correct");
  }
  void download(std::span<uint64_t> target, const transport::Buffer &in, size_t offset) override {
    const auto &a = buf(in);
    require(offset <= a.elements && target.size() <= a.elements - offset, "HIP download extent");
    healthy();
    flush();
    ok(ntt_hip_memcpy(c_->device, target.data(), a.data + offset, target.size_bytes(), NTT_HIP_D2H, c_->stream()),
       "download");
  }
  void record(transport::Event &e) override {
    auto &out = evt(e);
    healthy();
    flush();  // the event covers every command issued before it
    ok(ntt_hip_event_record(c_->device, out.event, c_->stream()), "record");
    out.recorded = true;
  }
  bool ready(transport::Event &e) override {
    auto &out = evt(e);
    require(out.recorded, "unrecorded HIP event");
    healthy();
    const int s = ntt_hip_event_query(c_->device, out.event);
    if (s == NTT_HIP_NOT_READY) return false;
    ok(s, "query");
    return true;
  }
  void wait(transport::Event &e) override {
    auto &out = evt(e);
    require(out.recorded, "unrecorded HIP event");
    healthy();
    ok(ntt_hip_event_synchronize(c_->device, out.event), "wait");
  }
  bool quiesce() noexcept override {
    try {
      flush();
      return ntt_hip_stream_synchronize(c_->device, c_->stream()) == NTT_HIP_OK;
    } catch (...) {
      return false;
    }
  }
};

}  // namespace ntt::hip_transport

namespace ntt::resident {

// Preflight before any device call (reference resident_adapter.inc:306-317).
Session make_hip_session(const PolynomialLayout &p, int device, Limits limits) {
  proposal::validate(p);
  proposal::require(p.domain == proposal::Domain::Coefficient && p.order == proposal::Order::Natural &&
                        p.parameters->n() <= 65536 && p.towers.size() <= 65535 && device >= 0 && limits.operations,
                    "unsupported resident HIP session");
  const size_t constants = hip_transport::table_bytes(p);
  proposal::require(constants <= limits.device_bytes && constants <= limits.host_bytes,
                    "HIP plan exceeds resident setup budget");
  return Session(std::make_shared<hip_transport::Driver>(p, device), limits);
}

Session make_cuda_session(const PolynomialLayout &p, int device, Limits limits) {
  return make_hip_session(p, device, limits);
}

}  // namespace ntt::resident
