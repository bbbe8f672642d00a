#pragma once
// Resident operation chain (reference include/ntt/resident.hpp:1-144
// interface): immutable polynomial handles, one ordered stream per session,
// session-owned host staging, budgets and a bounded operation log.
// make_cuda_session is the reference's factory name; in this build it
// creates the HIP transport (same as make_hip_session).
#include <cstdint>
#include <memory>
#include <span>
#include <vector>

#include "ntt/backend_contract.hpp"

namespace ntt::resident {
using proposal::CompletionRecord;
using proposal::Device;
using proposal::Outcome;
using proposal::PolynomialLayout;
using proposal::View;
namespace detail {
struct State;
struct Allocation;
struct Operation;
struct HostStorage;
class Driver;
}  // namespace detail

struct Limits {
  size_t device_bytes = 256 * 1024 * 1024;
  size_t host_bytes = 128 * 1024 * 1024;
  size_t operations = 1024;
};
struct Stats {
  size_t submitted, pending, device_bytes, host_bytes;
  size_t uploads, downloads, device_copies;
  size_t parameter_uploads;
  bool closed, poisoned, quiescent;
};

class Completion {
  std::shared_ptr<detail::State> state_;
  std::shared_ptr<detail::Operation> operation_;
  Completion(std::shared_ptr<detail::State>, std::shared_ptr<detail::Operation>);
  friend class Session;
  friend class Polynomial;
  friend class Download;
  friend struct detail::State;

 public:
  Completion() = default;
  CompletionRecord query() const;
  void wait() const;
};

class Polynomial {
  std::shared_ptr<detail::State> state_;
  std::shared_ptr<detail::Allocation> allocation_;
  std::shared_ptr<detail::Operation> producer_;
  PolynomialLayout layout_{};
  View view_{};
  Polynomial(std::shared_ptr<detail::State>, std::shared_ptr<detail::Allocation>, std::shared_ptr<detail::Operation>,
             PolynomialLayout, View);
  friend class Session;
  friend struct detail::State;

 public:
  Polynomial() = default;
  const PolynomialLayout &layout() const;
  View view() const;
  Completion readiness() const;
};

class Download {
  std::shared_ptr<detail::State> state_;
  std::shared_ptr<detail::Operation> operation_;
  std::shared_ptr<detail::HostStorage> storage_;
  PolynomialLayout layout_{};
  mutable bool validated_ = false;
  Download(std::shared_ptr<detail::State>, std::shared_ptr<detail::Operation>, std::shared_ptr<detail::HostStorage>,
           PolynomialLayout);
  friend class Session;

 public:
  Download() = default;
  Completion completion() const;
  // Exposed only after successful completion and canonical validation.
  std::span<const uint64_t> wait_values() const;
};

struct HybridKeyDigit {
  std::span<const uint64_t> b, a;
};
class HybridKey {
  std::vector<Polynomial> b_, a_;
  friend class Session;
};
struct HybridContributions {
  Polynomial b, a;
};

class Session {
  std::shared_ptr<detail::State> state_;
  detail::State &open_state();
  Polynomial upload_checked(const PolynomialLayout &, std::span<const uint64_t>, std::span<const Completion>);
  Polynomial transform(const Polynomial &, bool inverse, std::span<const Completion>);
  Polynomial join(std::span<const Polynomial>);
  Polynomial correct(const Polynomial &, const Polynomial &);

 public:
  // Driver injection is the transport seam (tests inject a host queue);
  // make_hip_session / make_cuda_session create the device transport.
  explicit Session(std::shared_ptr<detail::Driver>, Limits = {});
  ~Session() noexcept;
  Session(const Session &) = delete;
  Session &operator=(const Session &) = delete;
  Session(Session &&) noexcept = default;
  Session &operator=(Session &&) = delete;
  Device device() const;
  Stats stats() const;
  std::vector<CompletionRecord> history() const;
  Polynomial upload(std::span<const uint64_t>, std::span<const Completion> waits = {});
  Polynomial upload(const PolynomialLayout &, std::span<const uint64_t>, std::span<const Completion> waits = {});
  Polynomial upload_evaluation(std::span<const uint64_t>, std::span<const Completion> waits = {});
  Polynomial forward(const Polynomial &, std::span<const Completion> waits = {});
  Polynomial inverse(const Polynomial &, std::span<const Completion> waits = {});
  Polynomial pointwise(const Polynomial &, const Polynomial &, std::span<const Completion> waits = {});
  // Coefficient-domain unreduced approximate CRT; source/target disjoint.
  Polynomial switch_basis(const Polynomial &, const PolynomialLayout &target, std::span<const Completion> waits = {});
  // CKKS, reduced-noise off. Session basis: active Q followed by P.
  Polynomial mod_up(const Polynomial &, std::span<const Completion> waits = {});
  Polynomial mod_down(const Polynomial &, std::span<const Completion> waits = {});
  std::vector<Polynomial> hybrid_digits(const Polynomial &, std::span<const Completion> waits = {});
  HybridKey upload_hybrid_key(const PolynomialLayout &full_key_layout, std::span<const HybridKeyDigit>);
  HybridContributions hybrid_accumulate(std::span<const Polynomial> digits, const HybridKey &,
                                        std::span<const Completion> waits = {});
  HybridContributions hybrid_rotate_extended(const HybridContributions &, const Polynomial &c0,
                                             uint32_t automorphism);
  Polynomial subview(const Polynomial &, size_t first_tower, size_t count, size_t active_q);
  Download download(const Polynomial &, std::span<const Completion> waits = {});
  Download multiply(std::span<const uint64_t>, std::span<const uint64_t>);
  // Stops admission and drains; reports the session's first error.
  void close();
};

Session make_hip_session(const PolynomialLayout &coefficient_layout, int device = 0, Limits = {});
Session make_cuda_session(const PolynomialLayout &coefficient_layout, int device = 0, Limits = {});

}  // namespace ntt::resident
