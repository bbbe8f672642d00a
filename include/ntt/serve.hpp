#pragma once
// Continuous-batching request server for the CKKS hot path on one GPU.
// No counterpart in the reference (it runs one synchronous graph per call):
// requests of mixed ring dimension N, level L and operation are admitted as
// stream slots and device memory free up, grouped with compatible pending
// requests into one batched launch, and overlap on the device. Evaluation
// keys and twiddle tables stay resident across requests.
#include <chrono>
#include <cstdint>
#include <memory>
#include <span>
#include <vector>

namespace ntt::serve {

enum class Op : int { KeySwitch = 0, Forward = 1, Inverse = 2, ModUp = 3, ModDown = 4 };

// A key-switching parameter set: M Q towers then K P towers of ring
// dimension n, digit size alpha, and its full-context evaluation key
// (ceil(M/alpha) digits of (M+K) towers each, evaluation form).
struct ParamSet {
  uint64_t n = 0;
  std::vector<uint64_t> q, q_psi, p, p_psi;
  size_t alpha = 1;
  std::vector<uint64_t> key_b, key_a;
};

struct Options {
  int device = 0;
  size_t hbm_budget_bytes = size_t{32} << 30;  // everything the server holds on the device: resident plans and keys,
                                               // request buffers and workspaces in flight
  size_t streams = 4;                          // concurrent batches (hardware queues)
  size_t max_batch = 32;                       // requests per launch
};

struct Result {
  std::vector<uint64_t> out, out2;  // KeySwitch: B then A contributions; otherwise out only
  double latency_ms = 0;            // submit -> result visible on the host (device requests: in HBM)
  size_t batch_size = 0;            // requests in the launch that served it
};

struct Stats {
  size_t submitted = 0, completed = 0, batches = 0, max_in_flight = 0;
  // device bytes held (slot arenas + resident plans) now and at the peak
  size_t device_bytes_in_flight = 0, peak_device_bytes = 0;
  size_t resident_bytes = 0;  // contexts, keys and per-ring pools kept across requests
  size_t plans_resident = 0, plans_evicted = 0;
};

class Server {
 public:
  explicit Server(Options = {});
  ~Server();  // drains every admitted request, then stops
  Server(const Server &) = delete;
  Server &operator=(const Server &) = delete;

  uint32_t add_params(const ParamSet &);
  // Input words: KeySwitch/Forward/Inverse/ModUp: level*n (ModUp/KeySwitch:
  // evaluation Q); ModDown: (level + K)*n evaluation QP. Copied at submit.
  uint64_t submit(uint32_t params, Op op, size_t level, std::span<const uint64_t> input);
  // Device-resident request: `in`, `out` (and `out2`, the A contribution of a
  // key switch) are device buffers owned by the caller, so ciphertexts stay in
  // HBM across requests and nothing crosses the host; a batch of one runs on
  // them directly, larger batches gather into the slot's arena and scatter
  // back device to device. Inputs must hold canonical residues (the host does
  // not read them). wait() returns an empty Result::out.
  uint64_t submit_device(uint32_t params, Op op, size_t level, const uint64_t *in, uint64_t *out, uint64_t *out2);
  bool ready(uint64_t ticket) const;
  Result wait(uint64_t ticket);  // blocks; each ticket can be collected once
  Stats stats() const;
  size_t output_words(uint32_t params, Op op, size_t level) const;

 private:
  struct Impl;
  std::unique_ptr<Impl> impl_;
};

}  // namespace ntt::serve
