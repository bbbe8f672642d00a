// Resident operation chain (reference src/resident.cpp:1-662 semantics,
// include/ntt/resident.hpp interface). Every Session call validates on the
// host, reserves budget, pins its inputs and then enqueues on the driver's
// single ordered stream; outputs are fresh immutable allocations. The first
// transport error poisons the session; pins are released only after the
// driver proves quiescence, otherwise everything stays quarantined.
#include <algorithm>
#include <array>
#include <atomic>
#include <limits>
#include <map>
#include <unordered_map>

#include "ntt/resident_conversion.hpp"
#include "ntt/resident_driver.hpp"

namespace ntt::resident {
using namespace proposal;

namespace detail {

struct Budget {
  Limits limits;
  size_t device = 0, host = 0;
};

// RAII credit against the device or host byte budget.
struct Reservation {
  std::shared_ptr<Budget> budget;
  size_t count;
  bool host;
  Reservation(std::shared_ptr<Budget> b, size_t n, bool on_host) : budget(std::move(b)), count(n), host(on_host) {
    size_t &used = host ? budget->host : budget->device;
    const size_t limit = host ? budget->limits.host_bytes : budget->limits.device_bytes;
    require(used <= limit && n <= limit - used, "resident memory budget exhausted");
    used += n;
  }
  ~Reservation() { (host ? budget->host : budget->device) -= count; }
  Reservation(const Reservation &) = delete;
  Reservation &operator=(const Reservation &) = delete;
};

struct Allocation {
  Reservation reservation;  // declared first: released after the buffer
  AllocationRecord record;
  std::unique_ptr<Buffer> buffer;
  Allocation(std::shared_ptr<Budget> b, AllocationRecord r)
      : reservation(std::move(b), r.capacity_bytes, false), record(std::move(r)) {}
};

struct HostStorage {
  Reservation reservation;
  std::vector<uint64_t> values;
  HostStorage(std::shared_ptr<Budget> b, size_t count)
      : reservation(std::move(b), checked_product(count, sizeof(uint64_t)), true), values(count) {}
};

struct DeviceStorage {
  Reservation reservation;
  std::unique_ptr<Buffer> buffer;
  DeviceStorage(std::shared_ptr<Budget> b, size_t count)
      : reservation(std::move(b), checked_product(count, sizeof(uint64_t)), false) {}
};

struct Operation {
  CompletionRecord record;
  std::unique_ptr<Event> event;
  std::vector<std::shared_ptr<Allocation>> allocations;
  std::vector<std::shared_ptr<Operation>> dependencies;
  std::shared_ptr<HostStorage> staging;
  std::shared_ptr<DeviceStorage> constants;
  explicit Operation(CompletionRecord r) : record(r) {}
  void unpin() noexcept {
    event.reset();
    allocations.clear();
    dependencies.clear();
    staging.reset();
    constants.reset();
  }
};

static uint64_t next_session_id() {
  static std::atomic<uint64_t> serial{1};
  uint64_t value = serial.load();
  do {
    require(value != std::numeric_limits<uint64_t>::max(), "session identity exhausted");
  } while (!serial.compare_exchange_weak(value, value + 1));
  return value;
}

struct State : std::enable_shared_from_this<State> {
  std::shared_ptr<Driver> driver;
  PolynomialLayout base;
  Device device;
  Stream stream;
  std::shared_ptr<Budget> budget;
  Reservation constants;
  std::unordered_map<uint64_t, std::weak_ptr<Allocation>> registry;
  std::vector<std::shared_ptr<Operation>> operations;  // bounded journal
  std::shared_ptr<State> flight_lease;                 // self-pin while work may be queued
  std::exception_ptr error;
  std::exception_ptr drain_error = std::make_exception_ptr(std::runtime_error("resident quiescence unproven"));
  uint64_t allocation_id = 0;
  size_t uploads = 0, downloads = 0, copies = 0, parameter_uploads = 0;
  bool closed = false, quiescent = true;

  // Packed constant tables by moduli. Every switch_basis and correct still
  // stages and uploads its table (the driver sees the same calls), but each
  // distinct table is built once per session: building a 6 -> 24 basis plan
  // re-validates the primality of all 30 moduli and costs ~0.2 ms of host
  // time, which the device otherwise spends idle.
  static constexpr size_t kTableCache = 256;
  std::map<std::pair<std::vector<uint64_t>, std::vector<uint64_t>>, std::vector<uint64_t>> basis_tables;
  std::map<std::vector<uint64_t>, std::vector<uint64_t>> correction_tables;
  const std::vector<uint64_t> &basis_table(std::vector<uint64_t> from, std::vector<uint64_t> to) {
    auto key = std::make_pair(std::move(from), std::move(to));
    auto it = basis_tables.find(key);
    if (it == basis_tables.end()) {
      auto packed = pack_basis(ApproximateBasisPlan(key.first, key.second));
      if (basis_tables.size() >= kTableCache) basis_tables.clear();
      it = basis_tables.emplace(std::move(key), std::move(packed)).first;
    }
    return it->second;
  }
  // [q_t; P^-1 mod q_t] for the given Q moduli (reference resident.cpp:404-430).
  const std::vector<uint64_t> &correction_table(const ntt::proposal::Parameters &params, std::vector<uint64_t> q) {
    auto it = correction_tables.find(q);
    if (it == correction_tables.end()) {
      std::vector<uint64_t> packed = q;
      for (const uint64_t m : q) packed.push_back(inv_mod(special_prime_product(params, m), m));
      if (correction_tables.size() >= kTableCache) correction_tables.clear();
      it = correction_tables.emplace(std::move(q), std::move(packed)).first;
    }
    return it->second;
  }

  State(std::shared_ptr<Driver> d, Limits limits)
      : driver(std::move(d)),
        base(driver->layout()),
        device{driver->ordinal(), next_session_id()},
        stream{device, {1, 1}},
        budget(std::make_shared<Budget>(Budget{limits})),
        constants(budget, driver->constant_bytes(), false) {
    validate(base);
    validate(device);
    require(base.domain == Domain::Coefficient && base.order == Order::Natural, "session requires coefficient layout");
    require(base.parameters->n() <= 65536 && base.towers.size() <= 65535, "prototype extent unsupported");
    require(limits.operations > 0, "empty resident operation budget");
    driver->attach_once();
  }

  void open() const { require(!closed && !error, "resident session closed or poisoned"); }

  // p is a contiguous run of the session basis with the same parameters.
  void compatible(const PolynomialLayout &p) const {
    validate(p);
    require(*p.parameters == *base.parameters && p.active_q <= base.active_q, "foreign resident conversion parameters");
    const auto it = std::find(base.towers.begin(), base.towers.end(), p.towers.front());
    require(it != base.towers.end() && p.towers.size() <= static_cast<size_t>(base.towers.end() - it) &&
                std::equal(p.towers.begin(), p.towers.end(), it),
            "layout is not a contiguous subset of the session basis");
  }

  void conversion_input(const PolynomialLayout &p, bool up) const {
    compatible(p);
    require(base.towers == qp_basis(*base.parameters, base.active_q) && p.active_q == base.active_q &&
                p.active_q <= 40 && base.parameters->p().size() <= 40,
            "conversion requires a session with active Q then 1..40 P towers");
    require(p.towers == (up ? q_prefix(*p.parameters, p.active_q) : qp_basis(*p.parameters, p.active_q)),
            "conversion requires complete active Q or QP in original order");
    require(up || p.domain == Domain::Evaluation, "ModDown requires evaluation input");
  }

  void check(const Polynomial &p) const;

  std::shared_ptr<Allocation> reserve(const PolynomialLayout &layout) {
    for (auto it = registry.begin(); it != registry.end();) it = it->second.expired() ? registry.erase(it) : ++it;
    require(allocation_id != std::numeric_limits<uint64_t>::max(), "allocation identity exhausted");
    auto a = std::make_shared<Allocation>(budget, AllocationRecord{{++allocation_id, 1}, device, bytes(layout), layout});
    registry.emplace(a->record.token.id, a);
    return a;
  }

  // Journal the operation with every pin in place before any native work.
  std::shared_ptr<Operation> begin(std::vector<std::shared_ptr<Allocation>> pins,
                                   std::vector<std::shared_ptr<Operation>> sources, std::span<const Completion> waits,
                                   std::shared_ptr<HostStorage> staging = {});

  void finish(const std::shared_ptr<Operation> &op) { driver->record(*op->event); }

  bool drain() noexcept {
    if (!driver->quiesce()) return false;
    quiescent = true;
    for (auto &op : operations) {
      if (op->record.outcome == Outcome::Pending) op->record.outcome = error ? Outcome::Failed : Outcome::Succeeded;
      op->unpin();
    }
    flight_lease.reset();
    return true;
  }

  void fail(std::exception_ptr reason) noexcept {
    if (!error) error = reason;
    for (auto &op : operations)
      if (op->record.outcome == Outcome::Pending) op->record.outcome = Outcome::Failed;
    (void)drain();  // false keeps the journal, staging, buffers and self lease
  }

  // Completion of op retires it and everything submitted before it.
  void poll(const std::shared_ptr<Operation> &op, bool wait) {
    if (op->record.outcome == Outcome::Failed) std::rethrow_exception(error);
    if (op->record.outcome == Outcome::Succeeded) return;
    try {
      if (wait) driver->wait(*op->event);
      else if (!driver->ready(*op->event)) return;
      for (auto &prior : operations) {
        if (prior->record.operation.id > op->record.operation.id) break;
        if (prior->record.outcome == Outcome::Pending) prior->record.outcome = Outcome::Succeeded;
        prior->unpin();
      }
      if (op == operations.back()) {
        quiescent = true;
        flight_lease.reset();
      }
    } catch (...) {
      fail(std::current_exception());
      throw;
    }
  }

  size_t offset(const Polynomial &p) const;
  Polynomial output(const std::shared_ptr<Allocation> &a, const std::shared_ptr<Operation> &op);

  // Event, output buffer, enqueue and record under one failure guard.
  template <class Enqueue>
  Polynomial submit_output(const std::shared_ptr<Allocation> &a, const std::shared_ptr<Operation> &op,
                           Enqueue enqueue) {
    try {
      op->event = driver->make_event();
      a->buffer = driver->allocate(a->record.capacity_bytes / sizeof(uint64_t));
      require(bool(a->buffer) && bool(op->event), "driver returned null resource");
      enqueue(*a->buffer);
      finish(op);
      return output(a, op);
    } catch (...) {
      fail(std::current_exception());
      throw;
    }
  }

  // Shared body of operations that upload a small constant table first
  // (switch_basis, correct): event, outputs, table buffer, table upload.
  template <class Enqueue>
  void submit_with_constants(const std::shared_ptr<Operation> &op, const std::shared_ptr<DeviceStorage> &table,
                             std::initializer_list<std::shared_ptr<Allocation>> outputs, Enqueue enqueue) {
    try {
      op->event = driver->make_event();
      for (const auto &a : outputs) a->buffer = driver->allocate(a->record.capacity_bytes / sizeof(uint64_t));
      table->buffer = driver->allocate(op->staging->values.size());
      bool ok = bool(table->buffer) && bool(op->event);
      for (const auto &a : outputs) ok = ok && bool(a->buffer);
      require(ok, "driver returned null resource");
      driver->upload(*table->buffer, op->staging->values);
      ++parameter_uploads;
      enqueue(*table->buffer);
      finish(op);
    } catch (...) {
      fail(std::current_exception());
      throw;
    }
  }
};

}  // namespace detail

// ---------------------------------------------------------------- handles
Completion::Completion(std::shared_ptr<detail::State> s, std::shared_ptr<detail::Operation> o)
    : state_(std::move(s)), operation_(std::move(o)) {}

CompletionRecord Completion::query() const {
  require(state_ && operation_, "empty completion");
  try {
    state_->poll(operation_, false);
  } catch (...) {
    if (operation_->record.outcome != Outcome::Failed) throw;
  }
  return operation_->record;
}

void Completion::wait() const {
  require(state_ && operation_, "empty completion");
  state_->poll(operation_, true);
}

Polynomial::Polynomial(std::shared_ptr<detail::State> s, std::shared_ptr<detail::Allocation> a,
                       std::shared_ptr<detail::Operation> o, PolynomialLayout p, View v)
    : state_(std::move(s)), allocation_(std::move(a)), producer_(std::move(o)), layout_(std::move(p)), view_(v) {}

const PolynomialLayout &Polynomial::layout() const {
  require(bool(allocation_), "empty polynomial");
  return layout_;
}
View Polynomial::view() const {
  require(bool(allocation_), "empty polynomial");
  return view_;
}
Completion Polynomial::readiness() const { return Completion(state_, producer_); }

Download::Download(std::shared_ptr<detail::State> s, std::shared_ptr<detail::Operation> o,
                   std::shared_ptr<detail::HostStorage> h, PolynomialLayout p)
    : state_(std::move(s)), operation_(std::move(o)), storage_(std::move(h)), layout_(std::move(p)) {}

Completion Download::completion() const { return Completion(state_, operation_); }

std::span<const uint64_t> Download::wait_values() const {
  require(state_ && operation_ && storage_, "empty download");
  completion().wait();
  if (!validated_) {
    try {
      validate_host_values(layout_, storage_->values);
      validated_ = true;
    } catch (...) {
      operation_->record.outcome = Outcome::Failed;
      state_->fail(std::current_exception());
      throw;
    }
  }
  return storage_->values;
}

// ---------------------------------------------------------------- State helpers needing complete types
namespace detail {

void State::check(const Polynomial &p) const {
  require(p.state_.get() == this && p.allocation_ && p.producer_, "foreign or empty polynomial");
  const auto it = registry.find(p.view_.allocation.id);
  require(it != registry.end() && it->second.lock() == p.allocation_, "unregistered allocation");
  require_layout(validate_view(p.view_, p.allocation_->record, device), p.layout_);
  require(p.producer_->record.outcome != Outcome::Failed, "failed producer");
}

std::shared_ptr<Operation> State::begin(std::vector<std::shared_ptr<Allocation>> pins,
                                        std::vector<std::shared_ptr<Operation>> sources,
                                        std::span<const Completion> waits, std::shared_ptr<HostStorage> staging) {
  open();
  require(operations.size() < budget->limits.operations, "resident operation budget exhausted");
  auto op = std::make_shared<Operation>(
      CompletionRecord{{operations.size() + 1, 1}, stream, Outcome::Pending});
  Execution execution{stream, op->record.operation.id, {}};
  for (const Completion &w : waits) {
    require(w.state_.get() == this && w.operation_, "foreign or empty dependency");
    execution.waits.push_back(w.operation_->record);
    sources.push_back(w.operation_);
  }
  for (const auto &source : sources) validate_dependency(source->record, execution);
  op->allocations = std::move(pins);
  op->dependencies = std::move(sources);
  op->staging = std::move(staging);
  operations.push_back(op);
  flight_lease = shared_from_this();
  quiescent = false;
  return op;
}

size_t State::offset(const Polynomial &p) const {
  return byte_offset(p.view_, p.allocation_->record) / sizeof(uint64_t);
}

Polynomial State::output(const std::shared_ptr<Allocation> &a, const std::shared_ptr<Operation> &op) {
  const auto &rec = a->record;
  return Polynomial(shared_from_this(), a, op, rec.layout, {rec.token, 0, rec.layout.towers.size(), rec.layout.active_q});
}

}  // namespace detail

// ---------------------------------------------------------------- Session
Session::Session(std::shared_ptr<detail::Driver> driver, Limits limits) {
  require(bool(driver), "missing resident driver");
  state_ = std::make_shared<detail::State>(std::move(driver), limits);
}

Session::~Session() noexcept {
  if (state_) try {
      close();
    } catch (...) {
    }
}

detail::State &Session::open_state() {
  require(bool(state_), "moved session");
  state_->open();
  return *state_;
}

Device Session::device() const {
  require(bool(state_), "moved session");
  return state_->device;
}

Stats Session::stats() const {
  require(bool(state_), "moved session");
  const auto &s = *state_;
  size_t pending = 0;
  for (const auto &op : s.operations) pending += op->record.outcome == Outcome::Pending;
  return {s.operations.size(), pending,     s.budget->device, s.budget->host, s.uploads, s.downloads,
          s.copies,            s.parameter_uploads, s.closed, bool(s.error), s.quiescent};
}

std::vector<CompletionRecord> Session::history() const {
  require(bool(state_), "moved session");
  std::vector<CompletionRecord> out;
  for (const auto &op : state_->operations) out.push_back(op->record);
  return out;
}

Polynomial Session::upload(std::span<const uint64_t> values, std::span<const Completion> waits) {
  auto &s = open_state();
  validate_host_values(s.base, values);
  return upload_checked(s.base, values, waits);
}

Polynomial Session::upload_evaluation(std::span<const uint64_t> values, std::span<const Completion> waits) {
  auto &s = open_state();
  const auto layout = transform_layout(s.base, false);
  validate_host_values(layout, values);
  return upload_checked(layout, values, waits);
}

Polynomial Session::upload(const PolynomialLayout &p, std::span<const uint64_t> values,
                           std::span<const Completion> waits) {
  auto &s = open_state();
  s.compatible(p);
  validate_host_values(p, values);
  return upload_checked(p, values, waits);
}

Polynomial Session::upload_checked(const PolynomialLayout &layout, std::span<const uint64_t> values,
                                   std::span<const Completion> waits) {
  auto &s = *state_;
  // The session owns a staging copy: the caller's memory is never borrowed.
  auto staging = std::make_shared<detail::HostStorage>(s.budget, values.size());
  std::copy(values.begin(), values.end(), staging->values.begin());
  auto a = s.reserve(layout);
  auto op = s.begin({a}, {}, waits, staging);
  return s.submit_output(a, op, [&](detail::Buffer &out) {
    s.driver->upload(out, staging->values);
    ++s.uploads;
  });
}

Polynomial Session::transform(const Polynomial &input, bool inverse, std::span<const Completion> waits) {
  auto &s = open_state();
  s.check(input);
  auto a = s.reserve(transform_layout(input.layout_, inverse));
  auto op = s.begin({input.allocation_, a}, {input.producer_}, waits);
  return s.submit_output(a, op, [&](detail::Buffer &out) {
    s.driver->copy(out, *input.allocation_->buffer, s.offset(input), bytes(input.layout_) / sizeof(uint64_t));
    ++s.copies;
    s.driver->transform(out, a->record.layout, inverse);
  });
}

Polynomial Session::forward(const Polynomial &p, std::span<const Completion> waits) { return transform(p, false, waits); }
Polynomial Session::inverse(const Polynomial &p, std::span<const Completion> waits) { return transform(p, true, waits); }

Polynomial Session::pointwise(const Polynomial &left, const Polynomial &right, std::span<const Completion> waits) {
  auto &s = open_state();
  s.check(left);
  s.check(right);
  auto a = s.reserve(pointwise_layout(left.layout_, right.layout_));
  auto op = s.begin({left.allocation_, right.allocation_, a}, {left.producer_, right.producer_}, waits);
  return s.submit_output(a, op, [&](detail::Buffer &out) {
    s.driver->copy(out, *left.allocation_->buffer, s.offset(left), bytes(left.layout_) / sizeof(uint64_t));
    ++s.copies;
    s.driver->pointwise(out, *right.allocation_->buffer, s.offset(right), a->record.layout);
  });
}

Polynomial Session::switch_basis(const Polynomial &input, const PolynomialLayout &target,
                                 std::span<const Completion> waits) {
  auto &s = open_state();
  s.check(input);
  s.compatible(input.layout_);
  s.compatible(target);
  require(input.layout_.domain == Domain::Coefficient && target.domain == Domain::Coefficient &&
              input.layout_.active_q == target.active_q,
          "basis switching requires coefficient layouts at the same active level");
  std::vector<uint64_t> from, to;
  for (const TowerRef t : input.layout_.towers) from.push_back(target.parameters->tower(t).q);
  for (const TowerRef t : target.towers) to.push_back(target.parameters->tower(t).q);
  const auto &packed = s.basis_table(std::move(from), std::move(to));
  auto staging = std::make_shared<detail::HostStorage>(s.budget, packed.size());
  std::copy(packed.begin(), packed.end(), staging->values.begin());
  auto table = std::make_shared<detail::DeviceStorage>(s.budget, packed.size());
  auto a = s.reserve(target);
  auto op = s.begin({input.allocation_, a}, {input.producer_}, waits, staging);
  op->constants = table;  // pinned before the first possible enqueue
  s.submit_with_constants(op, table, {a}, [&](detail::Buffer &constants) {
    s.driver->switch_basis(*a->buffer, *input.allocation_->buffer, s.offset(input), constants, input.layout_, target);
  });
  return s.output(a, op);
}

Polynomial Session::join(std::span<const Polynomial> parts) {
  auto &s = open_state();
  require(!parts.empty(), "empty join");
  PolynomialLayout layout = parts.front().layout_;
  layout.towers.clear();
  std::vector<std::shared_ptr<detail::Allocation>> pins;
  std::vector<std::shared_ptr<detail::Operation>> producers;
  for (const Polynomial &part : parts) {
    s.check(part);
    require(*layout.parameters == *part.layout_.parameters && layout.active_q == part.layout_.active_q &&
                layout.domain == part.layout_.domain && layout.order == part.layout_.order,
            "incompatible join layouts");
    layout.towers.insert(layout.towers.end(), part.layout_.towers.begin(), part.layout_.towers.end());
    pins.push_back(part.allocation_);
    producers.push_back(part.producer_);
  }
  s.compatible(layout);
  auto a = s.reserve(layout);
  pins.push_back(a);
  auto op = s.begin(std::move(pins), std::move(producers), {});
  return s.submit_output(a, op, [&](detail::Buffer &out) {
    size_t at = 0;
    for (const Polynomial &part : parts) {
      const size_t count = bytes(part.layout_) / sizeof(uint64_t);
      s.driver->copy_range(out, at, *part.allocation_->buffer, s.offset(part), count);
      at += count;
      ++s.copies;
    }
  });
}

Polynomial Session::correct(const Polynomial &left, const Polynomial &right) {
  auto &s = open_state();
  s.check(left);
  s.check(right);
  const auto layout = pointwise_layout(left.layout_, right.layout_);
  std::vector<uint64_t> moduli;
  for (const TowerRef t : layout.towers) moduli.push_back(layout.parameters->tower(t).q);
  const auto &packed = s.correction_table(*layout.parameters, std::move(moduli));
  auto staging = std::make_shared<detail::HostStorage>(s.budget, packed.size());
  std::copy(packed.begin(), packed.end(), staging->values.begin());
  auto table = std::make_shared<detail::DeviceStorage>(s.budget, packed.size());
  auto a = s.reserve(layout);
  auto op = s.begin({left.allocation_, right.allocation_, a}, {left.producer_, right.producer_}, {}, staging);
  op->constants = table;
  s.submit_with_constants(op, table, {a}, [&](detail::Buffer &constants) {
    s.driver->correct(*a->buffer, *left.allocation_->buffer, s.offset(left), *right.allocation_->buffer,
                      s.offset(right), constants, layout);
  });
  return s.output(a, op);
}

Polynomial Session::mod_up(const Polynomial &input, std::span<const Completion> waits) {
  auto &s = open_state();
  s.check(input);
  s.conversion_input(input.layout_, true);
  PolynomialLayout p_layout = s.base;
  p_layout.towers.erase(p_layout.towers.begin(), p_layout.towers.begin() + s.base.active_q);
  Polynomial coefficients, evaluations;
  if (input.layout_.domain == Domain::Coefficient) {
    coefficients = input;
    evaluations = forward(input, waits);
  } else {
    evaluations = input;
    coefficients = inverse(input, waits);
  }
  const Polynomial p_coefficients = switch_basis(coefficients, p_layout);
  // Built from named handles: g++ < 13 leaks already-constructed elements of
  // an aggregate whose initializer throws (GCC bug 66139).
  const Polynomial p_evaluations = forward(p_coefficients);
  const std::array<Polynomial, 2> parts{evaluations, p_evaluations};
  return join(parts);
}

Polynomial Session::mod_down(const Polynomial &input, std::span<const Completion> waits) {
  auto &s = open_state();
  s.check(input);
  s.conversion_input(input.layout_, false);
  const size_t L = input.layout_.active_q;
  const Polynomial q = subview(input, 0, L, L);
  const Polynomial p = subview(input, L, input.layout_.parameters->p().size(), L);
  const auto q_coefficient = transform_layout(q.layout_, true);
  const Polynomial correction = forward(switch_basis(inverse(p, waits), q_coefficient));
  return correct(q, correction);
}

std::vector<Polynomial> Session::hybrid_digits(const Polynomial &input, std::span<const Completion> waits) {
  auto &s = open_state();
  s.check(input);
  s.conversion_input(input.layout_, true);
  require(input.layout_.domain == Domain::Evaluation, "Hybrid digits require evaluation Q input");
  const size_t L = input.layout_.active_q, alpha = s.base.parameters->alpha();
  std::vector<Polynomial> digits;
  digits.reserve(1 + (L - 1) / alpha);
  for (size_t first = 0; first < L; first += alpha) {
    const size_t count = std::min(alpha, L - first);
    const Polynomial evaluations = subview(input, first, count, L);
    const Polynomial coefficients = inverse(evaluations, waits);
    std::vector<Polynomial> parts;
    // Complement targets in basis order, at most 40 per conversion plan.
    auto extend = [&](size_t begin, size_t end) {
      while (begin < end) {
        const size_t next = begin + std::min<size_t>(40, end - begin);
        PolynomialLayout target = s.base;
        target.towers.assign(s.base.towers.begin() + begin, s.base.towers.begin() + next);
        parts.push_back(forward(switch_basis(coefficients, target)));
        begin = next;
      }
    };
    extend(0, first);
    parts.push_back(evaluations);
    extend(first + count, s.base.towers.size());
    digits.push_back(join(parts));
  }
  return digits;
}

HybridKey Session::upload_hybrid_key(const PolynomialLayout &full_key_layout, std::span<const HybridKeyDigit> digits) {
  auto &s = open_state();
  const auto active = transform_layout(s.base, false);
  s.conversion_input(active, false);
  validate(full_key_layout);
  const Parameters &params = *s.base.parameters;
  const size_t M = params.q().size(), alpha = params.alpha();
  require(*full_key_layout.parameters == params && full_key_layout.active_q == M &&
              full_key_layout.towers == qp_basis(params, M) && full_key_layout.domain == Domain::Evaluation,
          "Hybrid keys require full-context evaluation QP");
  require(digits.size() == 1 + (M - 1) / alpha, "wrong full-context Hybrid key digit count");
  for (const HybridKeyDigit &d : digits) {  // every row is checked before any submission
    validate_host_values(full_key_layout, d.b);
    validate_host_values(full_key_layout, d.a);
  }
  const auto rows = key_indices(params, active.active_q);
  const size_t active_digits = 1 + (active.active_q - 1) / alpha, n = params.n();
  std::vector<uint64_t> packed(rows.size() * n);
  auto upload_rows = [&](std::span<const uint64_t> values) {
    for (size_t t = 0; t < rows.size(); ++t) std::copy_n(values.begin() + rows[t] * n, n, packed.begin() + t * n);
    return upload_checked(active, packed, {});
  };
  HybridKey key;
  for (const HybridKeyDigit &d : digits.first(active_digits)) {
    key.b_.push_back(upload_rows(d.b));
    key.a_.push_back(upload_rows(d.a));
  }
  return key;
}

HybridContributions Session::hybrid_accumulate(std::span<const Polynomial> digits, const HybridKey &key,
                                               std::span<const Completion> waits) {
  auto &s = open_state();
  const auto layout = transform_layout(s.base, false);
  s.conversion_input(layout, false);
  const size_t active_digits = 1 + (layout.active_q - 1) / layout.parameters->alpha();
  require(digits.size() == active_digits && key.b_.size() == active_digits && key.a_.size() == active_digits,
          "wrong active Hybrid digit count");
  std::vector<std::shared_ptr<detail::Allocation>> pins;
  std::vector<std::shared_ptr<detail::Operation>> producers;
  for (size_t d = 0; d < active_digits; ++d)
    for (const Polynomial *part : {&digits[d], &key.b_[d], &key.a_[d]}) {
      s.check(*part);
      require_layout(part->layout_, layout);
      require(s.offset(*part) == 0, "Hybrid MAC requires full active QP buffers");
      pins.push_back(part->allocation_);
      producers.push_back(part->producer_);
    }
  auto b = s.reserve(layout), a = s.reserve(layout);
  pins.insert(pins.end(), {b, a});
  auto op = s.begin(std::move(pins), std::move(producers), waits);
  try {
    op->event = s.driver->make_event();
    const size_t elements = bytes(layout) / sizeof(uint64_t);
    b->buffer = s.driver->allocate(elements);
    a->buffer = s.driver->allocate(elements);
    require(bool(b->buffer) && bool(a->buffer) && bool(op->event), "driver returned null resource");
    for (size_t d = 0; d < active_digits; ++d)
      s.driver->hybrid_mac(*b->buffer, *a->buffer, *digits[d].allocation_->buffer, *key.b_[d].allocation_->buffer,
                           *key.a_[d].allocation_->buffer, layout, d == 0);
    s.finish(op);
    return {s.output(b, op), s.output(a, op)};
  } catch (...) {
    s.fail(std::current_exception());
    throw;
  }
}

HybridContributions Session::hybrid_rotate_extended(const HybridContributions &extended, const Polynomial &c0,
                                                    uint32_t automorphism) {
  auto &s = open_state();
  s.check(extended.b);
  s.check(extended.a);
  s.check(c0);
  const auto &layout = extended.b.layout_;
  s.conversion_input(layout, false);
  require_layout(extended.a.layout_, layout);
  s.conversion_input(c0.layout_, true);
  require(c0.layout_.domain == Domain::Evaluation && s.offset(extended.b) == 0 && s.offset(extended.a) == 0,
          "Hybrid rotation requires evaluation Q and full evaluation QP buffers");
  require((automorphism & 1) && automorphism < 2 * layout.parameters->n(), "invalid negacyclic automorphism index");
  auto staging = std::make_shared<detail::HostStorage>(s.budget, layout.active_q);
  for (size_t t = 0; t < layout.active_q; ++t)
    staging->values[t] = detail::special_prime_product(*layout.parameters, layout.parameters->q()[t].q);
  auto table = std::make_shared<detail::DeviceStorage>(s.budget, layout.active_q);
  auto b = s.reserve(layout), a = s.reserve(layout);
  auto op = s.begin({extended.b.allocation_, extended.a.allocation_, c0.allocation_, b, a},
                    {extended.b.producer_, extended.a.producer_, c0.producer_}, {}, staging);
  op->constants = table;
  s.submit_with_constants(op, table, {b, a}, [&](detail::Buffer &scales) {
    s.driver->hybrid_rotate(*b->buffer, *a->buffer, *extended.b.allocation_->buffer, *extended.a.allocation_->buffer,
                            *c0.allocation_->buffer, s.offset(c0), scales, layout, automorphism);
  });
  return {s.output(b, op), s.output(a, op)};
}

Polynomial Session::subview(const Polynomial &p, size_t first, size_t count, size_t active_q) {
  auto &s = open_state();
  s.check(p);
  require(first <= p.view_.tower_count && count <= p.view_.tower_count - first, "subview exceeds parent extent");
  require(active_q <= p.view_.active_q, "subview cannot add Q towers");
  const View v{p.view_.allocation, p.view_.first_tower + first, count, active_q};
  auto layout = validate_view(v, p.allocation_->record, s.device);
  return Polynomial(state_, p.allocation_, p.producer_, std::move(layout), v);
}

Download Session::download(const Polynomial &p, std::span<const Completion> waits) {
  auto &s = open_state();
  s.check(p);
  auto staging = std::make_shared<detail::HostStorage>(s.budget, bytes(p.layout_) / sizeof(uint64_t));
  auto op = s.begin({p.allocation_}, {p.producer_}, waits, staging);
  try {
    op->event = s.driver->make_event();
    require(bool(op->event), "driver returned null event");
    s.driver->download(staging->values, *p.allocation_->buffer, s.offset(p));
    ++s.downloads;
    s.finish(op);
    return Download(state_, op, staging, p.layout_);
  } catch (...) {
    s.fail(std::current_exception());
    throw;
  }
}

Download Session::multiply(std::span<const uint64_t> a, std::span<const uint64_t> b) {
  auto &s = open_state();
  validate_host_values(s.base, a);  // both inputs checked before the first enqueue
  validate_host_values(s.base, b);
  const Polynomial da = upload_checked(s.base, a, {}), db = upload_checked(s.base, b, {});
  const Polynomial product = pointwise(forward(da), forward(db));
  return download(inverse(product));
}

void Session::close() {
  require(bool(state_), "moved session");
  auto &s = *state_;
  s.closed = true;
  if (!s.quiescent && !s.drain()) s.fail(s.drain_error);
  for (auto &[id, weak] : s.registry) {
    (void)id;
    if (auto a = weak.lock()) a->record.live = false;
  }
  try {
    s.driver->check_health();
  } catch (...) {
    s.fail(std::current_exception());
  }
  if (s.error) std::rethrow_exception(s.error);
}

}  // namespace ntt::resident
