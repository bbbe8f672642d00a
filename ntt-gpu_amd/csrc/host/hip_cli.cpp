// ntt_hip_cli: the synchronous GPU transform behind the reference CUDA CLI's
// command contract (reference src/cuda_cli.cpp:1-256, bench bodies
// src/cuda/ntt_kernels.cu:491-821):
//
//   ntt_hip_cli <vector-file> transform|identity   forward limbs as "LIMB k" blocks / round trip
//   ntt_hip_cli info                               device description
//   ntt_hip_cli verify <vector-file> [...]         GPU == CPU reference per file + a generated 2^16 case
//   ntt_hip_cli bench [min_exp max_exp limbs]      markdown table (46-bit primes)
//   ntt_hip_cli bench-jsonl|gpu-bench-jsonl|events-jsonl <parameter-vector> [warmups samples]
//
// JSONL modes print a workload line, per-sample and summary lines per
// operation, and a completion line; every timed result is checked bit-exactly
// against the CPU reference NTT before it is reported.
#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "ntt/cuda_ntt.hpp"
#include "ntt/measurement.hpp"
#include "ntt/ntt.hpp"
#include "ntt/param_gen.hpp"
#include "ntt/vector_io.hpp"
#include "ntt_hip.h"

namespace {

using ntt::hip::HipNttTransform;
namespace ms = ntt::measurement;

constexpr const char *kSchedule =
    "two-pass LDS transform: 2^16 = 256 x 256 (column pass + row pass, 8 stages each in registers/LDS), "
    "wide-lazy Shoup butterflies (q < 2^61), gfx950";

void ok(int status, const char *what) {
  if (status == NTT_HIP_OK) return;
  const std::string msg = std::string("HIP: ") + what + ": " + ntt_hip_last_error();
  if (status == NTT_HIP_EINVAL) throw std::invalid_argument(msg);
  throw std::runtime_error(msg);
}

int failures = 0;
void check(bool condition, const std::string &what) {
  if (!condition) {
    ++failures;
    std::cerr << "FAIL: " << what << "\n";
  }
}

uint64_t lcg_next(uint64_t &s) { return s = s * 6364136223846793005ULL + 1442695040888963407ULL; }

std::vector<uint64_t> limb(const std::vector<uint64_t> &v, uint64_t n, size_t k) {
  return {v.begin() + k * n, v.begin() + (k + 1) * n};
}

struct Device {
  int index = 0, count = 0, cus = 0, driver = 0, runtime = 0;
  uint64_t memory = 0;
  std::string name, arch;
};

Device describe() {
  Device d;
  ok(ntt_hip_device_count(&d.count), "device count");
  if (d.count == 0) throw std::runtime_error("no visible HIP device");
  char name[256], arch[128];
  ok(ntt_hip_device_name(d.index, name, sizeof name), "device name");
  ok(ntt_hip_device_properties(d.index, arch, sizeof arch, &d.memory, &d.cus, &d.driver, &d.runtime),
     "device properties");
  d.name = name;
  d.arch = arch;
  return d;
}

int cmd_info() {
  const Device d = describe();
  std::cout << "visible devices: " << d.count << "\nselected device: " << d.index << "\nname: " << d.name
            << "\ntransform schedule: " << kSchedule << "\narchitecture: " << d.arch
            << "\ncompute units: " << d.cus << "\nmemory bytes: " << d.memory << "\ndriver version: " << d.driver
            << "\nruntime version: " << d.runtime << '\n';
  return 0;
}

int cmd_transform(const ntt::VectorFile &f) {
  HipNttTransform t(f.n, f.q, f.psi);
  std::vector<uint64_t> v = ntt::stage_limbs(f);
  t.forward(v);
  std::string out;
  for (size_t k = 0; k < f.q.size(); ++k) {
    out += "LIMB " + std::to_string(k) + "\n";
    for (uint64_t i = 0; i < f.n; ++i) out += std::to_string(v[k * f.n + i]) + "\n";
  }
  std::cout << out;
  return 0;
}

int cmd_identity(const ntt::VectorFile &f) {
  HipNttTransform t(f.n, f.q, f.psi);
  const std::vector<uint64_t> in = ntt::stage_limbs(f);
  std::vector<uint64_t> v = in;
  t.forward(v);
  t.inverse(v);
  for (size_t i = 0; i < v.size(); ++i)
    if (v[i] != in[i]) {
      std::cerr << "IDENTITY FAIL limb " << i / f.n << " coeff " << i % f.n << ": " << v[i] << " != " << in[i] << "\n";
      return 1;
    }
  std::cout << "IDENTITY OK\n";
  return 0;
}

// GPU forward == CPU reference forward per limb, round trip restores the
// input, GPU multiply == reference negacyclic product per limb.
void verify_case(const std::string &label, uint64_t n, const std::vector<uint64_t> &in, HipNttTransform &t) {
  const size_t limbs = t.limbs();
  std::vector<uint64_t> v = in;
  t.forward(v);
  for (size_t k = 0; k < limbs; ++k) {
    std::vector<uint64_t> want = limb(in, n, k);
    t.ref(k).forward(want);
    check(limb(v, n, k) == want, label + ": gpu forward == ref forward, limb " + std::to_string(k));
  }
  t.inverse(v);
  check(v == in, label + ": gpu roundtrip identity");
  std::vector<uint64_t> right(limbs * n);
  for (size_t k = 0; k < limbs; ++k) {
    uint64_t s = 0x9E3779B97F4A7C15ULL ^ t.q(k);
    for (uint64_t i = 0; i < n; ++i) right[k * n + i] = lcg_next(s) % t.q(k);
  }
  const std::vector<uint64_t> product = t.multiply(in, right);
  for (size_t k = 0; k < limbs; ++k)
    check(limb(product, n, k) == ntt::ntt_multiply(t.ref(k), limb(in, n, k), limb(right, n, k)),
          label + ": gpu multiply == ref multiply, limb " + std::to_string(k));
}

int cmd_verify(int argc, char **argv) {
  for (int a = 2; a < argc; ++a) {
    const ntt::VectorFile f = ntt::read_vector_file(argv[a]);
    HipNttTransform t(f.n, f.q, f.psi);
    verify_case(argv[a], f.n, ntt::stage_limbs(f), t);
    std::cout << "VERIFIED: " << argv[a] << " (n=" << f.n << ", limbs=" << f.q.size() << ")\n";
  }
  constexpr uint64_t n = 65536, limbs = 3;
  const std::vector<uint64_t> q = ntt::find_primes(n, limbs);
  std::vector<uint64_t> psi, in(limbs * n);
  for (uint64_t m : q) psi.push_back(ntt::find_psi(m, n));
  for (size_t k = 0; k < limbs; ++k) {
    uint64_t s = 0x9E3779B97F4A7C15ULL ^ q[k];
    for (uint64_t i = 0; i < n; ++i) in[k * n + i] = lcg_next(s) % q[k];
  }
  HipNttTransform t(n, q, psi);
  verify_case("generated-n65536", n, in, t);
  std::cout << "VERIFIED: generated n=" << n << ", limbs=" << limbs << "\n";
  if (failures) return 1;
  std::cout << "VERIFIED OK\n";
  return 0;
}

double wall_ms(std::chrono::steady_clock::time_point a) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - a).count();
}

int cmd_bench(int argc, char **argv) {
  uint32_t lo = 16, hi = 16;
  uint64_t limbs = 3;
  if (argc >= 4) lo = std::stoul(argv[2]), hi = std::stoul(argv[3]);
  if (argc >= 5) limbs = std::stoull(argv[4]);
  if (lo < 1 || hi > 16 || lo > hi || limbs == 0 || limbs > 65535)
    throw std::invalid_argument("bench requires 1<=min<=max<=16 and 1<=limbs<=65535");
  const Device d = describe();
  std::cout << "device: " << d.name << " (" << d.arch << ")\ntransform schedule: " << kSchedule
            << "\nbaseline parameters: 46-bit primes; host validation and copies included\n"
            << "| n | limbs | ref fwd ms | gpu fwd ms | gpu inv ms | gpu mul ms |\n|---|---|---|---|---|---|\n";
  constexpr int kWarm = 2, kTimed = 10;
  for (uint32_t e = lo; e <= hi; ++e) {
    const uint64_t n = uint64_t{1} << e;
    const std::vector<uint64_t> q = ntt::find_primes(n, limbs);
    std::vector<uint64_t> psi;
    for (uint64_t m : q) psi.push_back(ntt::find_psi(m, n));
    HipNttTransform t(n, q, psi);
    std::vector<uint64_t> left(limbs * n), right(limbs * n);
    uint64_t s = 0x9E3779B97F4A7C15ULL;
    for (size_t i = 0; i < left.size(); ++i) {
      left[i] = lcg_next(s) % q[i / n];
      right[i] = lcg_next(s) % q[i / n];
    }
    std::vector<std::vector<uint64_t>> cpu(limbs);
    for (size_t k = 0; k < limbs; ++k) cpu[k] = limb(left, n, k);
    double ref = 0, fwd = 0, inv = 0, mul = 0;
    for (int r = 0; r < kWarm + kTimed; ++r) {
      const auto t0 = std::chrono::steady_clock::now();
      for (size_t k = 0; k < limbs; ++k) t.ref(k).forward(cpu[k]);
      if (r >= kWarm) ref += wall_ms(t0);
    }
    std::vector<uint64_t> w = left;
    for (int r = 0; r < kWarm; ++r) t.forward(w);
    for (int r = 0; r < kTimed; ++r) {
      w = left;
      const auto t0 = std::chrono::steady_clock::now();
      t.forward(w);
      fwd += wall_ms(t0);
    }
    for (int r = 0; r < kTimed; ++r) {
      const auto t0 = std::chrono::steady_clock::now();
      t.inverse(w);
      inv += wall_ms(t0);
    }
    for (int r = 0; r < kWarm; ++r) t.multiply(left, right);
    for (int r = 0; r < kTimed; ++r) {
      const auto t0 = std::chrono::steady_clock::now();
      t.multiply(left, right);
      mul += wall_ms(t0);
    }
    std::cout << "| 2^" << e << " (" << n << ") | " << limbs << " | " << ref / kTimed << " | " << fwd / kTimed << " | "
              << inv / kTimed << " | " << mul / kTimed << " |\n";
    std::cout.flush();
  }
  return 0;
}

// Device-resident operations timed with HIP events on the context's stream;
// inputs are uploaded and results checked outside the event interval.
void events_mode(uint64_t n, const std::vector<uint64_t> &q, const std::vector<uint64_t> &psi, size_t warmups,
                 size_t samples, const std::vector<uint64_t> &left, const std::vector<uint64_t> &evaluation,
                 const std::vector<uint64_t> &right_evaluation, const std::vector<uint64_t> &product,
                 const std::vector<uint64_t> &coefficients, const std::vector<uint64_t> &right) {
  const int dev = 0;
  const size_t count = q.size() * n, bytes = count * sizeof(uint64_t);
  struct Owned {
    ntt_hip_context *ctx = nullptr;
    void *a = nullptr, *b = nullptr, *pinned = nullptr, *start = nullptr, *stop = nullptr;
    ~Owned() {
      if (start) ntt_hip_event_destroy(0, start);
      if (stop) ntt_hip_event_destroy(0, stop);
      if (a) ntt_hip_free(0, a);
      if (b) ntt_hip_free(0, b);
      if (pinned) ntt_hip_host_free(pinned);
      if (ctx) ntt_hip_context_destroy(ctx);
    }
  } o;
  ok(ntt_hip_context_create(&o.ctx, dev, n, q.data(), psi.data(), q.size()), "context");
  ok(ntt_hip_malloc(dev, &o.a, bytes), "allocate");
  ok(ntt_hip_malloc(dev, &o.b, bytes), "allocate");
  ok(ntt_hip_host_alloc(&o.pinned, bytes), "pinned allocate");
  ok(ntt_hip_event_create(dev, &o.start, 1), "event");
  ok(ntt_hip_event_create(dev, &o.stop, 1), "event");
  void *st = ntt_hip_context_stream(o.ctx);
  auto *A = static_cast<uint64_t *>(o.a), *B = static_cast<uint64_t *>(o.b);
  auto *P = static_cast<uint64_t *>(o.pinned);
  std::vector<uint64_t> seen(count);
  const std::vector<uint64_t> zeros(count, 0);  // lives until every copy from it has completed
  auto upload = [&](uint64_t *dst, const std::vector<uint64_t> &v) {
    ok(ntt_hip_memcpy(dev, dst, v.data(), bytes, NTT_HIP_H2D, st), "upload");
  };
  auto expect_a = [&](const std::vector<uint64_t> &want, const char *what) {
    ok(ntt_hip_memcpy(dev, seen.data(), A, bytes, NTT_HIP_D2H, st), "download");
    ok(ntt_hip_stream_synchronize(dev, st), "synchronize");
    if (seen != want) throw std::runtime_error(std::string("event benchmark result mismatch: ") + what);
  };
  auto measure = [&](const char *name, auto prepare, auto operation, auto verify) {
    const auto v = ms::collect_elapsed(
        warmups, samples,
        [&] {
          prepare();
          ok(ntt_hip_stream_synchronize(dev, st), "synchronize");
        },
        [&] {
          ok(ntt_hip_event_record(dev, o.start, st), "record");
          operation();
          ok(ntt_hip_event_record(dev, o.stop, st), "record");
          ok(ntt_hip_event_synchronize(dev, o.stop), "event wait");
          float t = 0;
          ok(ntt_hip_event_elapsed(&t, o.start, o.stop), "elapsed");
          return static_cast<double>(t);
        },
        verify);
    ms::write_samples(std::cout, name, v);
  };
  const size_t towers = q.size();
  auto fwd = [&](uint64_t *x) { ok(ntt_hip_transform(o.ctx, x, nullptr, 0, towers, 0, st), "forward"); };
  auto inv = [&](uint64_t *x) { ok(ntt_hip_transform(o.ctx, x, nullptr, 0, towers, 1, st), "inverse"); };
  auto pw = [&] { ok(ntt_hip_pointwise(o.ctx, A, B, 0, towers, st), "pointwise"); };
  measure("gpu_forward", [&] { upload(A, left); }, [&] { fwd(A); }, [&] { expect_a(evaluation, "gpu_forward"); });
  measure("gpu_inverse", [&] { upload(A, evaluation); }, [&] { inv(A); }, [&] { expect_a(left, "gpu_inverse"); });
  measure(
      "gpu_pointwise", [&] { upload(A, evaluation), upload(B, right_evaluation); }, pw,
      [&] { expect_a(product, "gpu_pointwise"); });
  measure(
      "gpu_fused_multiply", [&] { upload(A, left), upload(B, right); },
      [&] {
        fwd(A);
        fwd(B);
        pw();
        inv(A);
      },
      [&] { expect_a(coefficients, "gpu_fused_multiply"); });
  measure(
      "h2d_pinned",
      [&] {
        std::copy(left.begin(), left.end(), P);
        upload(A, zeros);
      },
      [&] { ok(ntt_hip_memcpy(dev, A, P, bytes, NTT_HIP_H2D, st), "pinned upload"); },
      [&] { expect_a(left, "h2d_pinned"); });
  measure(
      "d2h_pinned",
      [&] {
        upload(A, left);
        std::fill(P, P + count, 0);
      },
      [&] { ok(ntt_hip_memcpy(dev, P, A, bytes, NTT_HIP_D2H, st), "pinned download"); },
      [&] {
        if (!std::equal(left.begin(), left.end(), P)) throw std::runtime_error("pinned D2H benchmark result mismatch");
      });
}

int cmd_jsonl(ms::BenchmarkMode mode, const ntt::VectorFile &f, size_t warmups, size_t samples) {
  const auto &protocol = ms::benchmark_protocol(mode);
  ms::validate_counts(warmups, samples);
  const uint64_t n = f.n;
  const std::vector<uint64_t> &q = f.q, &psi = f.psi;
  HipNttTransform t(n, q, psi);  // validates the parameters before any device work
  const Device d = describe();
  const auto [left, right] = ms::benchmark_inputs(n, q);
  const size_t count = left.size();
  auto cpu_forward = [&](std::vector<uint64_t> &v) {
    for (size_t k = 0; k < q.size(); ++k) {
      std::vector<uint64_t> l = limb(v, n, k);
      t.ref(k).forward(l);
      std::copy(l.begin(), l.end(), v.begin() + k * n);
    }
  };
  std::vector<uint64_t> evaluation = left, right_evaluation = right;
  cpu_forward(evaluation);
  cpu_forward(right_evaluation);
  std::vector<uint64_t> product(count), coefficients;
  for (size_t i = 0; i < count; ++i) product[i] = ntt::mul_mod(evaluation[i], right_evaluation[i], q[i / n]);
  coefficients = product;
  for (size_t k = 0; k < q.size(); ++k) {
    std::vector<uint64_t> l = limb(coefficients, n, k);
    t.ref(k).inverse(l);
    std::copy(l.begin(), l.end(), coefficients.begin() + k * n);
  }
  std::ostream &out = std::cout;
  out << "{\"type\":\"workload\",\"schema\":1,\"n\":" << n << ",\"towers\":" << q.size() << ",\"warmups\":" << warmups
      << ",\"samples\":" << samples << ",\"timing\":\"" << protocol.timing
      << "\",\"input_policy\":\"fixed_seed_reset_outside_timer\",\"seed\":" << ms::kInputSeed
      << ",\"setup_included\":false,\"device\":";
  ms::json_string(out, d.name);
  out << ",\"transform_schedule\":";
  ms::json_string(out, kSchedule);
  out << ",\"architecture\":";
  ms::json_string(out, d.arch);
  out << ",\"device_bytes\":" << d.memory << ",\"driver_version\":" << d.driver
      << ",\"runtime_version\":" << d.runtime << ",\"moduli\":[";
  for (size_t k = 0; k < q.size(); ++k) out << (k ? "," : "") << q[k];
  out << "],\"roots\":[";
  for (size_t k = 0; k < psi.size(); ++k) out << (k ? "," : "") << psi[k];
  out << "]}\n";
  if (mode == ms::BenchmarkMode::ResidentEvents) {
    events_mode(n, q, psi, warmups, samples, left, evaluation, right_evaluation, product, coefficients, right);
    out << "{\"type\":\"complete\",\"operations\":" << protocol.operation_count << "}\n";
    return 0;
  }
  std::vector<uint64_t> work, other, prod;
  auto measure = [&](const char *name, auto prepare, auto operation, const std::vector<uint64_t> &want,
                     const std::vector<uint64_t> &got) {
    const auto v = ms::collect(warmups, samples, prepare, operation);
    if (got != want) throw std::runtime_error(std::string("benchmark result mismatch: ") + name);
    ms::write_samples(out, name, v);
  };
  if (mode == ms::BenchmarkMode::PublicWithCpu) {
    std::vector<std::vector<uint64_t>> cpu(q.size());
    const auto v = ms::collect(
        warmups, samples,
        [&] {
          for (size_t k = 0; k < q.size(); ++k) cpu[k] = limb(left, n, k);
        },
        [&] {
          for (size_t k = 0; k < q.size(); ++k) t.ref(k).forward(cpu[k]);
        });
    for (size_t k = 0; k < q.size(); ++k)
      if (cpu[k] != limb(evaluation, n, k)) throw std::runtime_error("CPU benchmark result mismatch");
    ms::write_samples(out, "cpu_reference_forward", v);
  }
  measure("gpu_forward", [&] { work = left; }, [&] { t.forward(work); }, evaluation, work);
  measure("gpu_inverse", [&] { work = evaluation; }, [&] { t.inverse(work); }, left, work);
  measure("gpu_pointwise", [&] { work = evaluation; }, [&] { t.pointwise_multiply(work, right_evaluation); }, product,
          work);
  measure(
      "gpu_unfused_multiply", [&] { work = left, other = right; },
      [&] {
        t.forward(work);
        t.forward(other);
        t.pointwise_multiply(work, other);
        t.inverse(work);
      },
      coefficients, work);
  measure("gpu_fused_multiply", [&] { std::vector<uint64_t>().swap(prod); }, [&] { prod = t.multiply(left, right); },
          coefficients, prod);
  out << "{\"type\":\"complete\",\"operations\":" << protocol.operation_count << "}\n";
  return 0;
}

const char *kUsage =
    "usage: ntt_hip_cli <vector-file> <transform|identity>\n"
    "       ntt_hip_cli info\n"
    "       ntt_hip_cli verify <vector-file> [...]\n"
    "       ntt_hip_cli bench [min_exp max_exp limbs]\n"
    "       ntt_hip_cli bench-jsonl|gpu-bench-jsonl|events-jsonl <parameter-vector> [warmups samples]\n";

}  // namespace

int main(int argc, char **argv) {
  try {
    const std::string first = argc > 1 ? argv[1] : "";
    if (const auto mode = ms::benchmark_mode(first)) {
      if (argc != 3 && argc != 5)
        throw std::invalid_argument(
            "usage: bench-jsonl|gpu-bench-jsonl|events-jsonl <parameter-vector> [warmups samples]");
      return cmd_jsonl(*mode, ntt::read_vector_file(argv[2]), argc == 5 ? std::stoul(argv[3]) : 5,
                       argc == 5 ? std::stoul(argv[4]) : 100);
    }
    if (argc == 2 && first == "info") return cmd_info();
    if (argc >= 2 && first == "verify") {
      if (argc < 3) {
        std::cerr << "usage: ntt_hip_cli verify <vector-file> [...]\n";
        return 2;
      }
      return cmd_verify(argc, argv);
    }
    if (argc >= 2 && first == "bench") return cmd_bench(argc, argv);
    if (argc == 3) {
      const ntt::VectorFile f = ntt::read_vector_file(argv[1]);
      const std::string command = argv[2];
      if (command == "transform") return cmd_transform(f);
      if (command == "identity") return cmd_identity(f);
      std::cerr << "unknown command: " << command << "\n";
      return 2;
    }
    std::cerr << kUsage;
    return 2;
  } catch (const std::exception &e) {
    std::cerr << "error: " << e.what() << "\n";
    return 1;
  }
}
