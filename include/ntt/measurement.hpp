#pragma once
// Timing protocol of the reference's JSONL benches (reference
// include/ntt/measurement.hpp:13-95, benchmark_mode.hpp:8-33,
// benchmark_inputs.hpp:8-22): inputs reset before every warmup and sample
// outside the timed interval; per-sample and summary lines as JSON (median,
// p5, p95 by linear interpolation); the fixed operand generator (MMIX LCG
// from seed 0x7251d9, left and right operands interleaved).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <iomanip>
#include <numeric>
#include <optional>
#include <ostream>
#include <stdexcept>
#include <string>
#include <string_view>
#include <utility>
#include <vector>

namespace ntt::measurement {

enum class BenchmarkMode { PublicWithCpu, PublicGpuOnly, ResidentEvents };

struct BenchmarkProtocol {
  BenchmarkMode mode;
  std::string_view command, timing;
  size_t operation_count;
};

// The command names and timing labels are the reference's CLI contract; the
// events mode is timed with HIP events here.
inline constexpr BenchmarkProtocol kBenchmarkProtocols[] = {
    {BenchmarkMode::PublicWithCpu, "bench-jsonl", "host_wall_public_call", 6},
    {BenchmarkMode::PublicGpuOnly, "gpu-bench-jsonl", "host_wall_gpu_public_call", 5},
    {BenchmarkMode::ResidentEvents, "events-jsonl", "hip_event_resident_and_pinned_copy", 6},
};

inline const BenchmarkProtocol &benchmark_protocol(BenchmarkMode mode) {
  for (const auto &p : kBenchmarkProtocols)
    if (p.mode == mode) return p;
  throw std::invalid_argument("unknown benchmark mode");
}

inline std::optional<BenchmarkMode> benchmark_mode(std::string_view command) {
  for (const auto &p : kBenchmarkProtocols)
    if (p.command == command) return p.mode;
  return std::nullopt;
}

inline void validate_counts(size_t warmups, size_t samples) {
  if (warmups < 1 || warmups > 10000 || samples < 2 || samples > 100000)
    throw std::invalid_argument("require 1..10000 warmups and 2..100000 samples");
}

inline double checked_sample(double ms) {
  if (!std::isfinite(ms) || ms < 0) throw std::runtime_error("invalid timing sample");
  return ms;
}

// Host wall clock around `operation`; `prepare` runs before every call, untimed.
template <class Prepare, class Operation>
std::vector<double> collect(size_t warmups, size_t samples, Prepare prepare, Operation operation) {
  validate_counts(warmups, samples);
  std::vector<double> out;
  out.reserve(samples);
  for (size_t i = 0; i < warmups + samples; ++i) {
    prepare();
    const auto t0 = std::chrono::steady_clock::now();
    operation();
    const auto t1 = std::chrono::steady_clock::now();
    if (i >= warmups) out.push_back(checked_sample(std::chrono::duration<double, std::milli>(t1 - t0).count()));
  }
  return out;
}

// External clock: `operation` returns its own elapsed milliseconds (device
// events); `verify` runs after every call, outside the timed interval.
template <class Prepare, class Operation, class Verify>
std::vector<double> collect_elapsed(size_t warmups, size_t samples, Prepare prepare, Operation operation,
                                    Verify verify) {
  validate_counts(warmups, samples);
  std::vector<double> out;
  out.reserve(samples);
  for (size_t i = 0; i < warmups + samples; ++i) {
    prepare();
    const double ms = checked_sample(operation());
    verify();
    if (i >= warmups) out.push_back(ms);
  }
  return out;
}

inline double percentile(std::vector<double> v, double fraction) {
  if (v.empty() || !(fraction >= 0 && fraction <= 1)) throw std::invalid_argument("invalid percentile input");
  for (double x : v) checked_sample(x);
  std::sort(v.begin(), v.end());
  const double pos = fraction * static_cast<double>(v.size() - 1);
  const size_t lo = static_cast<size_t>(pos), hi = std::min(lo + 1, v.size() - 1);
  return v[lo] + (v[hi] - v[lo]) * (pos - static_cast<double>(lo));
}

inline void json_string(std::ostream &out, std::string_view s) {
  static const char hex[] = "0123456789abcdef";
  out << '"';
  for (unsigned char c : s) {
    if (c == '"' || c == '\\')
      out << '\\' << c;
    else if (c < 0x20)
      out << "\\u00" << hex[c >> 4] << hex[c & 15];
    else
      out << c;
  }
  out << '"';
}

inline void write_samples(std::ostream &out, std::string_view operation, const std::vector<double> &samples) {
  out << std::setprecision(17);
  for (size_t i = 0; i < samples.size(); ++i) {
    out << "{\"type\":\"sample\",\"operation\":";
    json_string(out, operation);
    out << ",\"iteration\":" << i << ",\"milliseconds\":" << samples[i] << "}\n";
  }
  const double mean = std::accumulate(samples.begin(), samples.end(), 0.0) / static_cast<double>(samples.size());
  out << "{\"type\":\"summary\",\"operation\":";
  json_string(out, operation);
  out << ",\"count\":" << samples.size() << ",\"mean_ms\":" << mean << ",\"median_ms\":" << percentile(samples, 0.5)
      << ",\"p5_ms\":" << percentile(samples, 0.05) << ",\"p95_ms\":" << percentile(samples, 0.95) << "}\n";
}

inline constexpr uint64_t kInputSeed = 0x7251d9;

// Left and right operands, limb-major, drawn alternately from one LCG stream.
inline std::pair<std::vector<uint64_t>, std::vector<uint64_t>> benchmark_inputs(uint64_t n,
                                                                                const std::vector<uint64_t> &q) {
  std::vector<uint64_t> left(q.size() * n), right(q.size() * n);
  uint64_t s = kInputSeed;
  auto next = [&] { return s = s * 6364136223846793005ULL + 1442695040888963407ULL; };
  for (size_t i = 0; i < left.size(); ++i) {
    const uint64_t m = q[i / n];
    left[i] = next() % m;
    right[i] = next() % m;
  }
  return {std::move(left), std::move(right)};
}

}  // namespace ntt::measurement
