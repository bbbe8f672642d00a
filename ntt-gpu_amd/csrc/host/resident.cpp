// Resident Session runtime over the Driver seam (public interface
// include/ntt/resident.hpp; the semantics a caller observes follow the
// reference's src/resident.cpp:1-662 and docs/resident-operation-chain.md).
//
// Structure of this implementation:
//   Ledger      device/host byte budgets; a Charge is an RAII debit.
//   Allocation  one device output: its budget charge, record and buffer.
//   HostStorage host staging (uploads, downloads, constant tables).
//   Operation   one journal entry: its completion record, its event and a
//               type-erased pin list (inputs, producers, staging, constants)
//               that keeps everything the queued work may touch alive.
//   State       the journal (append-only, bounded by Limits::operations), a
//               settled prefix (operations proven complete), the first error
//               and the driver. The single ordered stream makes completion
//               monotone: when entry k is proven done, entries 1..k are too.
// Every public operation validates on the host, charges its budgets and
// records its journal entry before the first driver call, then issues the
// driver calls in one fixed order (event, output buffers, constant buffer,
// constant upload, the work itself, event record). The first exception from
// the driver poisons the session: pending entries become Failed and their
// pins are released only if the driver proves it is quiescent; otherwise
// they stay pinned (quarantined) for the life of the process.
#include <algorithm>
#include <array>
#include <atomic>
#include <limits>
#include <map>

#include "ntt/resident_conversion.hpp"
#include "ntt/resident_driver.hpp"

namespace ntt::resident {
using namespace proposal;

namespace detail {

// ---------------------------------------------------------------- budgets
enum class Pool { Device, Host };

struct Ledger {
  Limits limits;
  size_t used[2] = {0, 0};
  size_t cap(Pool p) const { return p == Pool::Device ? limits.device_bytes : limits.host_bytes; }
};

class Charge {
  std::shared_ptr<Ledger> ledger_;
  size_t bytes_;
  Pool pool_;

 public:
  Charge(std::shared_ptr<Ledger> ledger, size_t bytes, Pool pool)
      : ledger_(std::move(ledger)), bytes_(bytes), pool_(pool) {
    size_t &used = ledger_->used[int(pool_)];
    const size_t cap = ledger_->cap(pool_);
    require(used <= cap && bytes_ <= cap - used,
            pool_ == Pool::Device ? "device byte limit of the session exceeded"
                                  : "host byte limit of the session exceeded");
    used += bytes_;
  }
  ~Charge() { ledger_->used[int(pool_)] -= bytes_; }
  Charge(const Charge &) = delete;
  Charge &operator=(const Charge &) = delete;
};

// ---------------------------------------------------------------- resources
struct Allocation {
  Charge charge;  // first member: released after the buffer below
  AllocationRecord record;
  std::unique_ptr<Buffer> buffer;
  Allocation(std::shared_ptr<Ledger> ledger, AllocationRecord r)
      : charge(std::move(ledger), r.capacity_bytes, Pool::Device), record(std::move(r)) {}
  size_t words() const { return record.capacity_bytes / sizeof(uint64_t); }
};

struct HostStorage {
  Charge charge;
  std::vector<uint64_t> values;
  HostStorage(std::shared_ptr<Ledger> ledger, std::span<const uint64_t> from)
      : charge(std::move(ledger), checked_product(from.size(), sizeof(uint64_t)), Pool::Host),
        values(from.begin(), from.end()) {}
  HostStorage(std::shared_ptr<Ledger> ledger, size_t count)
      : charge(std::move(ledger), checked_product(count, sizeof(uint64_t)), Pool::Host), values(count) {}
};

// A per-operation device table (basis-conversion plan, correction or
// rotation scales), uploaded from a HostStorage in the same operation.
struct ConstantBlock {
  Charge charge;
  std::unique_ptr<Buffer> buffer;
  ConstantBlock(std::shared_ptr<Ledger> ledger, size_t words)
      : charge(std::move(ledger), checked_product(words, sizeof(uint64_t)), Pool::Device) {}
};

struct Operation {
  CompletionRecord record;
  std::unique_ptr<Event> event;
  std::vector<std::shared_ptr<const void>> pins;
  explicit Operation(CompletionRecord r) : record(r) {}
  void release() noexcept {
    event.reset();
    pins.clear();
  }
};

uint64_t fresh_session_id() {
  static std::atomic<uint64_t> counter{0};
  const uint64_t id = counter.fetch_add(1, std::memory_order_relaxed) + 1;
  require(id != 0 && id != std::numeric_limits<uint64_t>::max(), "session identifiers exhausted");
  return id;
}

// What one enqueue needs besides its body: inputs to pin and order after,
// fresh outputs, optional host staging and an optional device table filled
// from that staging.
struct Request {
  std::vector<const Polynomial *> inputs;
  std::vector<std::shared_ptr<Allocation>> outputs;
  std::shared_ptr<HostStorage> staging;
  std::shared_ptr<ConstantBlock> table;
  std::span<const Completion> waits;
};

struct State : std::enable_shared_from_this<State> {
  std::shared_ptr<Driver> driver;
  PolynomialLayout base;
  Device device{};
  Stream stream{};
  std::shared_ptr<Ledger> ledger;
  std::unique_ptr<Charge> table_charge;  // the driver's own constant tables
  std::vector<std::shared_ptr<Operation>> journal;
  size_t settled = 0;             // journal[0, settled) is proven complete (or failed and released)
  std::shared_ptr<State> anchor;  // self reference while work may be queued
  std::exception_ptr first_error;
  uint64_t last_token = 0;
  size_t uploads = 0, downloads = 0, copies = 0, table_uploads = 0;
  bool closed = false;

  // Each distinct conversion/correction table is packed once per session
  // (building a 6 -> 24 plan re-validates 30 moduli, ~0.2 ms of host time);
  // every operation still stages and uploads its own copy.
  std::map<std::pair<std::vector<uint64_t>, std::vector<uint64_t>>, std::vector<uint64_t>> plan_memo;
  std::map<std::vector<uint64_t>, std::vector<uint64_t>> correction_memo;

  State(std::shared_ptr<Driver> d, Limits limits) : driver(std::move(d)) {
    base = driver->layout();
    validate(base);
    require(base.domain == Domain::Coefficient && base.order == Order::Natural,
            "a session basis is given in coefficient form, natural order");
    require(base.parameters->n() <= 65536, "ring dimension above 65536 is not supported");
    require(base.towers.size() <= 65535, "more than 65535 towers is not supported");
    require(limits.operations > 0, "the operation limit must be positive");
    device = {driver->ordinal(), fresh_session_id()};
    validate(device);
    stream = {device, {1, 1}};
    ledger = std::make_shared<Ledger>(Ledger{limits});
    table_charge = std::make_unique<Charge>(ledger, driver->constant_bytes(), Pool::Device);
    driver->attach_once();
  }

  bool quiescent() const { return settled == journal.size(); }
  void require_open() const {
    require(!closed, "session is closed");
    require(!first_error, "session is poisoned by an earlier failure");
  }

  // ---- layouts
  // A contiguous run of session towers at no higher level than the session.
  void within_basis(const PolynomialLayout &p) const {
    validate(p);
    require(*p.parameters == *base.parameters, "layout belongs to other parameters");
    require(p.active_q <= base.active_q, "layout is above the session level");
    const auto start = std::find(base.towers.begin(), base.towers.end(), p.towers.front());
    const bool run = start != base.towers.end() &&
                     p.towers.size() <= size_t(base.towers.end() - start) &&
                     std::equal(p.towers.begin(), p.towers.end(), start);
    require(run, "towers are not a contiguous run of the session basis");
  }
  // Conversion sessions hold active Q then P; inputs are the whole Q prefix
  // (towards QP) or the whole QP basis (back to Q, evaluation form).
  void conversion_layout(const PolynomialLayout &p, bool towards_qp) const {
    within_basis(p);
    const Parameters &par = *base.parameters;
    require(base.towers == qp_basis(par, base.active_q), "session basis is not active Q followed by P");
    require(p.active_q == base.active_q && p.active_q <= 40 && par.p().size() <= 40,
            "conversion needs the session level and at most 40 Q and 40 P towers");
    require(p.towers == (towards_qp ? q_prefix(par, p.active_q) : qp_basis(par, p.active_q)),
            towards_qp ? "expected the complete active Q basis" : "expected the complete active QP basis");
    require(towards_qp || p.domain == Domain::Evaluation, "reduction to Q takes evaluation input");
  }
  PolynomialLayout basis_slice(size_t first, size_t end) const {
    PolynomialLayout out = base;
    out.towers.assign(base.towers.begin() + first, base.towers.begin() + end);
    return out;
  }

  // ---- handles
  void own(const Polynomial &p) const {
    require(p.state_.get() == this, "polynomial belongs to another session");
    require(p.allocation_ && p.producer_, "empty polynomial");
    require_layout(validate_view(p.view_, p.allocation_->record, device), p.layout_);
    require(p.producer_->record.outcome != Outcome::Failed, "polynomial was produced by a failed operation");
  }
  size_t word_offset(const Polynomial &p) const { return byte_offset(p.view_, p.allocation_->record) / 8; }
  static size_t word_count(const Polynomial &p) { return bytes(p.layout_) / 8; }
  const Buffer &buffer_of(const Polynomial &p) const { return *p.allocation_->buffer; }

  std::shared_ptr<Allocation> new_output(const PolynomialLayout &layout) {
    require(last_token != std::numeric_limits<uint64_t>::max(), "allocation identifiers exhausted");
    return std::make_shared<Allocation>(ledger, AllocationRecord{{++last_token, 1}, device, bytes(layout), layout});
  }
  Polynomial handle(const std::shared_ptr<Allocation> &a, const std::shared_ptr<Operation> &op) {
    const View v{a->record.token, 0, a->record.layout.towers.size(), a->record.layout.active_q};
    return Polynomial(shared_from_this(), a, op, a->record.layout, v);
  }

  // ---- the one submission routine
  template <class Body>
  std::shared_ptr<Operation> submit(Request &rq, Body &&body) {
    require_open();
    require(journal.size() < ledger->limits.operations, "operation limit of the session reached");
    auto op = std::make_shared<Operation>(CompletionRecord{{journal.size() + 1, 1}, stream, Outcome::Pending});
    Execution exec{stream, op->record.operation.id, {}};
    std::vector<std::shared_ptr<Operation>> after;
    for (const Completion &w : rq.waits) {
      require(w.state_.get() == this && w.operation_, "wait on a completion of another session");
      exec.waits.push_back(w.operation_->record);
      after.push_back(w.operation_);
    }
    for (const Polynomial *in : rq.inputs) after.push_back(in->producer_);
    for (const auto &dep : after) validate_dependency(dep->record, exec);
    // every pin and the journal entry exist before the first driver call
    op->pins.reserve(after.size() + rq.inputs.size() + rq.outputs.size() + 2);
    for (auto &dep : after) op->pins.push_back(std::move(dep));
    for (const Polynomial *in : rq.inputs) op->pins.push_back(in->allocation_);
    for (const auto &out : rq.outputs) op->pins.push_back(out);
    if (rq.staging) op->pins.push_back(rq.staging);
    if (rq.table) op->pins.push_back(rq.table);
    journal.push_back(op);
    anchor = shared_from_this();
    try {
      op->event = driver->make_event();
      require(bool(op->event), "driver returned no event");
      for (const auto &out : rq.outputs) {
        out->buffer = driver->allocate(out->words());
        require(bool(out->buffer), "driver returned no buffer");
      }
      if (rq.table) {
        rq.table->buffer = driver->allocate(rq.staging->values.size());
        require(bool(rq.table->buffer), "driver returned no buffer");
        driver->upload(*rq.table->buffer, rq.staging->values);
        ++table_uploads;
      }
      body();
      driver->record(*op->event);
    } catch (...) {
      poison(std::current_exception());
      throw;
    }
    return op;
  }

  // ---- completion
  // Release every entry up to and including index `last` (0-based).
  void settle_through(size_t last) noexcept {
    for (; settled <= last && settled < journal.size(); ++settled) {
      Operation &e = *journal[settled];
      if (e.record.outcome == Outcome::Pending) e.record.outcome = Outcome::Succeeded;
      e.release();
    }
    if (quiescent()) anchor.reset();
  }
  // Drain after a failure or at close: only a quiescent driver lets pins go.
  bool drain() noexcept {
    if (!driver->quiesce()) return false;
    for (; settled < journal.size(); ++settled) {
      Operation &e = *journal[settled];
      if (e.record.outcome == Outcome::Pending) e.record.outcome = first_error ? Outcome::Failed : Outcome::Succeeded;
      e.release();
    }
    anchor.reset();
    return true;
  }
  void poison(std::exception_ptr why) noexcept {
    if (!first_error) first_error = why;
    for (size_t i = settled; i < journal.size(); ++i)
      if (journal[i]->record.outcome == Outcome::Pending) journal[i]->record.outcome = Outcome::Failed;
    (void)drain();  // false: everything stays pinned (quarantine)
  }
  void await(const std::shared_ptr<Operation> &op, bool block) {
    if (op->record.outcome == Outcome::Failed) std::rethrow_exception(first_error);
    const size_t index = op->record.operation.id - 1;
    if (op->record.outcome == Outcome::Succeeded || index < settled) return;
    try {
      if (block) driver->wait(*op->event);
      else if (!driver->ready(*op->event)) return;
      settle_through(index);
    } catch (...) {
      poison(std::current_exception());
      throw;
    }
  }

  // ---- constant tables
  const std::vector<uint64_t> &plan_table(std::vector<uint64_t> from, std::vector<uint64_t> to) {
    auto key = std::make_pair(std::move(from), std::move(to));
    auto hit = plan_memo.find(key);
    if (hit != plan_memo.end()) return hit->second;
    if (plan_memo.size() >= 256) plan_memo.clear();
    std::vector<uint64_t> packed = pack_basis(ApproximateBasisPlan(key.first, key.second));
    return plan_memo.emplace(std::move(key), std::move(packed)).first->second;
  }
  // [q_t ...; P^-1 mod q_t ...]
  const std::vector<uint64_t> &correction_table(std::vector<uint64_t> moduli) {
    auto hit = correction_memo.find(moduli);
    if (hit != correction_memo.end()) return hit->second;
    if (correction_memo.size() >= 256) correction_memo.clear();
    std::vector<uint64_t> packed = moduli;
    for (const uint64_t m : moduli) packed.push_back(inv_mod(special_prime_product(*base.parameters, m), m));
    return correction_memo.emplace(std::move(moduli), std::move(packed)).first->second;
  }
  std::vector<uint64_t> moduli_of(const PolynomialLayout &p) const {
    std::vector<uint64_t> out;
    out.reserve(p.towers.size());
    for (const TowerRef t : p.towers) out.push_back(p.parameters->tower(t).q);
    return out;
  }
};

}  // namespace detail

// ---------------------------------------------------------------- handles
Completion::Completion(std::shared_ptr<detail::State> s, std::shared_ptr<detail::Operation> o)
    : state_(std::move(s)), operation_(std::move(o)) {}

CompletionRecord Completion::query() const {
  require(state_ && operation_, "empty completion");
  try {
    state_->await(operation_, false);
  } catch (...) {
    // a failed operation reports its record; any other error propagates
    if (operation_->record.outcome != Outcome::Failed) throw;
  }
  return operation_->record;
}

void Completion::wait() const {
  require(state_ && operation_, "empty completion");
  state_->await(operation_, true);
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
    } catch (...) {
      // a non-canonical result is a transport failure of this operation
      operation_->record.outcome = Outcome::Failed;
      state_->poison(std::current_exception());
      throw;
    }
    validated_ = true;
  }
  return storage_->values;
}

// ---------------------------------------------------------------- session
Session::Session(std::shared_ptr<detail::Driver> driver, Limits limits) {
  require(bool(driver), "a session needs a driver");
  state_ = std::make_shared<detail::State>(std::move(driver), limits);
}

Session::~Session() noexcept {
  if (!state_) return;
  try {
    close();
  } catch (...) {
  }
}

detail::State &Session::open_state() {
  require(bool(state_), "session was moved from");
  state_->require_open();
  return *state_;
}

Device Session::device() const {
  require(bool(state_), "session was moved from");
  return state_->device;
}

Stats Session::stats() const {
  require(bool(state_), "session was moved from");
  const detail::State &s = *state_;
  size_t pending = 0;
  for (size_t i = s.settled; i < s.journal.size(); ++i) pending += s.journal[i]->record.outcome == Outcome::Pending;
  return {s.journal.size(), pending, s.ledger->used[0], s.ledger->used[1], s.uploads, s.downloads, s.copies,
          s.table_uploads, s.closed, bool(s.first_error), s.quiescent()};
}

std::vector<CompletionRecord> Session::history() const {
  require(bool(state_), "session was moved from");
  std::vector<CompletionRecord> out;
  out.reserve(state_->journal.size());
  for (const auto &e : state_->journal) out.push_back(e->record);
  return out;
}

Polynomial Session::upload(std::span<const uint64_t> values, std::span<const Completion> waits) {
  detail::State &s = open_state();
  validate_host_values(s.base, values);
  return upload_checked(s.base, values, waits);
}

Polynomial Session::upload(const PolynomialLayout &layout, std::span<const uint64_t> values,
                           std::span<const Completion> waits) {
  detail::State &s = open_state();
  s.within_basis(layout);
  validate_host_values(layout, values);
  return upload_checked(layout, values, waits);
}

Polynomial Session::upload_evaluation(std::span<const uint64_t> values, std::span<const Completion> waits) {
  detail::State &s = open_state();
  const PolynomialLayout layout = transform_layout(s.base, false);
  validate_host_values(layout, values);
  return upload_checked(layout, values, waits);
}

Polynomial Session::upload_checked(const PolynomialLayout &layout, std::span<const uint64_t> values,
                                   std::span<const Completion> waits) {
  detail::State &s = *state_;
  detail::Request rq;
  rq.staging = std::make_shared<detail::HostStorage>(s.ledger, values);
  rq.outputs = {s.new_output(layout)};
  rq.waits = waits;
  auto op = s.submit(rq, [&] {
    s.driver->upload(*rq.outputs[0]->buffer, rq.staging->values);
    ++s.uploads;
  });
  return s.handle(rq.outputs[0], op);
}

Polynomial Session::transform(const Polynomial &input, bool inverse, std::span<const Completion> waits) {
  detail::State &s = open_state();
  s.own(input);
  const PolynomialLayout layout = transform_layout(input.layout_, inverse);
  detail::Request rq;
  rq.inputs = {&input};
  rq.outputs = {s.new_output(layout)};
  rq.waits = waits;
  auto op = s.submit(rq, [&] {
    // outputs are fresh: copy the input, then transform the copy in place
    detail::Buffer &out = *rq.outputs[0]->buffer;
    s.driver->copy(out, s.buffer_of(input), s.word_offset(input), s.word_count(input));
    ++s.copies;
    s.driver->transform(out, layout, inverse);
  });
  return s.handle(rq.outputs[0], op);
}

Polynomial Session::forward(const Polynomial &p, std::span<const Completion> waits) { return transform(p, false, waits); }
Polynomial Session::inverse(const Polynomial &p, std::span<const Completion> waits) { return transform(p, true, waits); }

Polynomial Session::pointwise(const Polynomial &left, const Polynomial &right, std::span<const Completion> waits) {
  detail::State &s = open_state();
  s.own(left);
  s.own(right);
  const PolynomialLayout layout = pointwise_layout(left.layout_, right.layout_);
  detail::Request rq;
  rq.inputs = {&left, &right};
  rq.outputs = {s.new_output(layout)};
  rq.waits = waits;
  auto op = s.submit(rq, [&] {
    detail::Buffer &out = *rq.outputs[0]->buffer;
    s.driver->copy(out, s.buffer_of(left), s.word_offset(left), s.word_count(left));
    ++s.copies;
    s.driver->pointwise(out, s.buffer_of(right), s.word_offset(right), layout);
  });
  return s.handle(rq.outputs[0], op);
}

Polynomial Session::switch_basis(const Polynomial &input, const PolynomialLayout &target,
                                 std::span<const Completion> waits) {
  detail::State &s = open_state();
  s.own(input);
  s.within_basis(input.layout_);
  s.within_basis(target);
  require(input.layout_.domain == Domain::Coefficient && target.domain == Domain::Coefficient,
          "basis switching works on coefficient layouts");
  require(input.layout_.active_q == target.active_q, "basis switching keeps the level");
  const std::vector<uint64_t> &table = s.plan_table(s.moduli_of(input.layout_), s.moduli_of(target));
  detail::Request rq;
  rq.inputs = {&input};
  rq.staging = std::make_shared<detail::HostStorage>(s.ledger, table);
  rq.table = std::make_shared<detail::ConstantBlock>(s.ledger, table.size());
  rq.outputs = {s.new_output(target)};
  rq.waits = waits;
  auto op = s.submit(rq, [&] {
    s.driver->switch_basis(*rq.outputs[0]->buffer, s.buffer_of(input), s.word_offset(input), *rq.table->buffer,
                           input.layout_, target);
  });
  return s.handle(rq.outputs[0], op);
}

// Concatenation of whole-tower parts into one fresh allocation (the digits of
// Hybrid key switching and the QP output of ModUp).
Polynomial Session::join(std::span<const Polynomial> parts) {
  detail::State &s = open_state();
  require(!parts.empty(), "nothing to join");
  PolynomialLayout layout = parts[0].layout_;
  layout.towers.clear();
  detail::Request rq;
  for (const Polynomial &part : parts) {
    s.own(part);
    const PolynomialLayout &l = part.layout_;
    require(*l.parameters == *layout.parameters && l.active_q == layout.active_q && l.domain == layout.domain &&
                l.order == layout.order,
            "joined parts differ in parameters, level or form");
    layout.towers.insert(layout.towers.end(), l.towers.begin(), l.towers.end());
    rq.inputs.push_back(&part);
  }
  s.within_basis(layout);
  rq.outputs = {s.new_output(layout)};
  auto op = s.submit(rq, [&] {
    size_t at = 0;
    for (const Polynomial &part : parts) {
      const size_t words = s.word_count(part);
      s.driver->copy_range(*rq.outputs[0]->buffer, at, s.buffer_of(part), s.word_offset(part), words);
      at += words;
      ++s.copies;
    }
  });
  return s.handle(rq.outputs[0], op);
}

// (left - right) * P^-1 per Q tower (ModDown's last step).
Polynomial Session::correct(const Polynomial &left, const Polynomial &right) {
  detail::State &s = open_state();
  s.own(left);
  s.own(right);
  const PolynomialLayout layout = pointwise_layout(left.layout_, right.layout_);
  const std::vector<uint64_t> &table = s.correction_table(s.moduli_of(layout));
  detail::Request rq;
  rq.inputs = {&left, &right};
  rq.staging = std::make_shared<detail::HostStorage>(s.ledger, table);
  rq.table = std::make_shared<detail::ConstantBlock>(s.ledger, table.size());
  rq.outputs = {s.new_output(layout)};
  auto op = s.submit(rq, [&] {
    s.driver->correct(*rq.outputs[0]->buffer, s.buffer_of(left), s.word_offset(left), s.buffer_of(right),
                      s.word_offset(right), *rq.table->buffer, layout);
  });
  return s.handle(rq.outputs[0], op);
}

Polynomial Session::mod_up(const Polynomial &input, std::span<const Completion> waits) {
  detail::State &s = open_state();
  s.own(input);
  s.conversion_layout(input.layout_, true);
  const size_t L = s.base.active_q;
  const bool coefficient_input = input.layout_.domain == Domain::Coefficient;
  // the Q part of the result in evaluation form, the conversion source in coefficients
  const Polynomial q_eval = coefficient_input ? forward(input, waits) : input;
  const Polynomial q_coeff = coefficient_input ? input : inverse(input, waits);
  const Polynomial p_coeff = switch_basis(q_coeff, s.basis_slice(L, s.base.towers.size()));
  const Polynomial p_eval = forward(p_coeff);
  const std::array<Polynomial, 2> halves{q_eval, p_eval};
  return join(halves);
}

Polynomial Session::mod_down(const Polynomial &input, std::span<const Completion> waits) {
  detail::State &s = open_state();
  s.own(input);
  s.conversion_layout(input.layout_, false);
  const size_t L = input.layout_.active_q, K = input.layout_.parameters->p().size();
  const Polynomial q_part = subview(input, 0, L, L);
  const Polynomial p_part = subview(input, L, K, L);
  const Polynomial p_coeff = inverse(p_part, waits);
  const Polynomial lifted = switch_basis(p_coeff, transform_layout(q_part.layout_, true));
  const Polynomial lifted_eval = forward(lifted);
  return correct(q_part, lifted_eval);
}

std::vector<Polynomial> Session::hybrid_digits(const Polynomial &input, std::span<const Completion> waits) {
  detail::State &s = open_state();
  s.own(input);
  s.conversion_layout(input.layout_, true);
  require(input.layout_.domain == Domain::Evaluation, "Hybrid digits take evaluation Q input");
  const size_t L = input.layout_.active_q, alpha = s.base.parameters->alpha(), T = s.base.towers.size();
  std::vector<Polynomial> digits;
  for (size_t first = 0; first < L; first += alpha) {
    const size_t end = std::min(L, first + alpha);
    const Polynomial own_eval = subview(input, first, end - first, L);
    const Polynomial own_coeff = inverse(own_eval, waits);
    std::vector<Polynomial> parts;
    // towers outside the digit, converted in runs of at most 40 targets
    auto extend = [&](size_t from, size_t to) {
      for (size_t lo = from; lo < to; lo += 40) {
        const size_t hi = std::min(to, lo + 40);
        const Polynomial converted = switch_basis(own_coeff, s.basis_slice(lo, hi));
        parts.push_back(forward(converted));
      }
    };
    extend(0, first);
    parts.push_back(own_eval);
    extend(end, T);
    digits.push_back(join(parts));
  }
  return digits;
}

HybridKey Session::upload_hybrid_key(const PolynomialLayout &full_key_layout, std::span<const HybridKeyDigit> digits) {
  detail::State &s = open_state();
  const PolynomialLayout active = transform_layout(s.base, false);
  s.conversion_layout(active, false);
  validate(full_key_layout);
  const Parameters &par = *s.base.parameters;
  const size_t M = par.q().size(), alpha = par.alpha();
  require(*full_key_layout.parameters == par && full_key_layout.active_q == M &&
              full_key_layout.towers == qp_basis(par, M) && full_key_layout.domain == Domain::Evaluation,
          "a Hybrid key is given in evaluation form over the full QP basis");
  require(digits.size() == (M + alpha - 1) / alpha, "a Hybrid key has one digit per alpha full-context Q towers");
  for (const HybridKeyDigit &d : digits) {
    validate_host_values(full_key_layout, d.b);
    validate_host_values(full_key_layout, d.a);
  }
  // rows of the active towers: i for Q, i + (M - L) for P
  const std::vector<size_t> rows = key_indices(par, active.active_q);
  const size_t n = par.n(), used = (active.active_q + alpha - 1) / alpha;
  std::vector<uint64_t> gathered(rows.size() * n);
  auto channel = [&](std::span<const uint64_t> full) {
    for (size_t t = 0; t < rows.size(); ++t)
      std::copy_n(full.begin() + rows[t] * n, n, gathered.begin() + t * n);
    return upload_checked(active, gathered, {});
  };
  HybridKey key;
  for (size_t d = 0; d < used; ++d) {
    key.b_.push_back(channel(digits[d].b));
    key.a_.push_back(channel(digits[d].a));
  }
  return key;
}

HybridContributions Session::hybrid_accumulate(std::span<const Polynomial> digits, const HybridKey &key,
                                               std::span<const Completion> waits) {
  detail::State &s = open_state();
  const PolynomialLayout layout = transform_layout(s.base, false);
  s.conversion_layout(layout, false);
  const size_t count = (layout.active_q + layout.parameters->alpha() - 1) / layout.parameters->alpha();
  require(digits.size() == count && key.b_.size() == count && key.a_.size() == count,
          "digit and key counts must match the session level");
  detail::Request rq;
  for (size_t d = 0; d < count; ++d)
    for (const Polynomial *part : {&digits[d], &key.b_[d], &key.a_[d]}) {
      s.own(*part);
      require_layout(part->layout_, layout);
      require(s.word_offset(*part) == 0, "the key inner product takes whole QP allocations");
      rq.inputs.push_back(part);
    }
  rq.outputs = {s.new_output(layout), s.new_output(layout)};
  rq.waits = waits;
  auto op = s.submit(rq, [&] {
    for (size_t d = 0; d < count; ++d)
      s.driver->hybrid_mac(*rq.outputs[0]->buffer, *rq.outputs[1]->buffer, s.buffer_of(digits[d]),
                           s.buffer_of(key.b_[d]), s.buffer_of(key.a_[d]), layout, d == 0);
  });
  return {s.handle(rq.outputs[0], op), s.handle(rq.outputs[1], op)};
}

HybridContributions Session::hybrid_rotate_extended(const HybridContributions &extended, const Polynomial &c0,
                                                    uint32_t automorphism) {
  detail::State &s = open_state();
  s.own(extended.b);
  s.own(extended.a);
  s.own(c0);
  const PolynomialLayout &layout = extended.b.layout_;
  s.conversion_layout(layout, false);
  require_layout(extended.a.layout_, layout);
  s.conversion_layout(c0.layout_, true);
  require(c0.layout_.domain == Domain::Evaluation, "c0 is given in evaluation form");
  require(s.word_offset(extended.b) == 0 && s.word_offset(extended.a) == 0,
          "the rotation takes whole QP allocations");
  require((automorphism & 1) && automorphism < 2 * layout.parameters->n(), "invalid negacyclic automorphism index");
  // P mod q_t for every active Q tower (the scale of c0)
  std::vector<uint64_t> scales(layout.active_q);
  for (size_t t = 0; t < layout.active_q; ++t)
    scales[t] = detail::special_prime_product(*layout.parameters, layout.parameters->q()[t].q);
  detail::Request rq;
  rq.inputs = {&extended.b, &extended.a, &c0};
  rq.staging = std::make_shared<detail::HostStorage>(s.ledger, scales);
  rq.table = std::make_shared<detail::ConstantBlock>(s.ledger, scales.size());
  rq.outputs = {s.new_output(layout), s.new_output(layout)};
  auto op = s.submit(rq, [&] {
    s.driver->hybrid_rotate(*rq.outputs[0]->buffer, *rq.outputs[1]->buffer, s.buffer_of(extended.b),
                            s.buffer_of(extended.a), s.buffer_of(c0), s.word_offset(c0), *rq.table->buffer, layout,
                            automorphism);
  });
  return {s.handle(rq.outputs[0], op), s.handle(rq.outputs[1], op)};
}

Polynomial Session::subview(const Polynomial &p, size_t first_tower, size_t count, size_t active_q) {
  detail::State &s = open_state();
  s.own(p);
  require(first_tower <= p.view_.tower_count && count <= p.view_.tower_count - first_tower,
          "subview reaches past its parent");
  require(active_q <= p.view_.active_q, "a subview cannot raise the level");
  const View v{p.view_.allocation, p.view_.first_tower + first_tower, count, active_q};
  PolynomialLayout layout = validate_view(v, p.allocation_->record, s.device);
  return Polynomial(state_, p.allocation_, p.producer_, std::move(layout), v);
}

Download Session::download(const Polynomial &p, std::span<const Completion> waits) {
  detail::State &s = open_state();
  s.own(p);
  detail::Request rq;
  rq.inputs = {&p};
  rq.staging = std::make_shared<detail::HostStorage>(s.ledger, s.word_count(p));
  rq.waits = waits;
  auto op = s.submit(rq, [&] {
    s.driver->download(rq.staging->values, s.buffer_of(p), s.word_offset(p));
    ++s.downloads;
  });
  return Download(state_, op, rq.staging, p.layout_);
}

Download Session::multiply(std::span<const uint64_t> a, std::span<const uint64_t> b) {
  detail::State &s = open_state();
  // both operands are checked before anything is enqueued
  validate_host_values(s.base, a);
  validate_host_values(s.base, b);
  const Polynomial da = upload_checked(s.base, a, {});
  const Polynomial db = upload_checked(s.base, b, {});
  const Polynomial ea = forward(da), eb = forward(db);
  const Polynomial product = pointwise(ea, eb);
  return download(inverse(product));
}

void Session::close() {
  require(bool(state_), "session was moved from");
  detail::State &s = *state_;
  s.closed = true;
  if (!s.quiescent() && !s.drain())
    s.poison(std::make_exception_ptr(std::runtime_error("queued work could not be proven complete")));
  try {
    s.driver->check_health();
  } catch (...) {
    s.poison(std::current_exception());
  }
  if (s.first_error) std::rethrow_exception(s.first_error);
}

}  // namespace ntt::resident
