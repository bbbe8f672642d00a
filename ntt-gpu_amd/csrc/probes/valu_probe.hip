// VALU throughput probe for gfx950: the integer multiply forms the kernels are
// built from, and the NTT's own wide-lazy butterflies (modarith.hpp ct8/gs8)
// in dependent chains with enough independent chains per thread and waves per
// CU to saturate issue. The butterfly rates are the VALU roof of the NTT:
// a transform of N = 2^16 points is 16 stages x 32768 butterflies. The
// 32-bit forms (ct32/gs32) are the roof of towers below 2^30.
//
// Usage: valu_probe [--json]   (one line per probe; exits non-zero on a HIP error)
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>

#include "../modarith.hpp"

using namespace nttamd;

#define PROBE_CHECK(x)                                                                    \
  do {                                                                                    \
    hipError_t e_ = (x);                                                                  \
    if (e_ != hipSuccess) {                                                               \
      std::fprintf(stderr, "HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__); \
      return 1;                                                                           \
    }                                                                                     \
  } while (0)

constexpr int kIter = 2048;  // dependent steps per chain
constexpr int kChains = 8;   // independent chains per thread

__global__ void p_mad_u64_u32(uint64_t *out, uint32_t seed) {
  uint32_t a[kChains];
  uint64_t acc[kChains];
  for (int i = 0; i < kChains; ++i) a[i] = seed + threadIdx.x * 7 + i, acc[i] = i;
  for (int it = 0; it < kIter; ++it)
#pragma unroll
    for (int i = 0; i < kChains; ++i) acc[i] = (uint64_t)a[i] * (uint32_t)(acc[i] >> 32) + acc[i];
  uint64_t s = 0;
  for (int i = 0; i < kChains; ++i) s ^= acc[i];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// two instructions per step (v_mul_hi_u32, v_or_b32)
__global__ void p_mul_hi_u32(uint64_t *out, uint32_t seed) {
  uint32_t a[kChains], acc[kChains];
  for (int i = 0; i < kChains; ++i) a[i] = seed + threadIdx.x * 7 + i, acc[i] = i + 3;
  for (int it = 0; it < kIter; ++it)
#pragma unroll
    for (int i = 0; i < kChains; ++i) acc[i] = __umulhi(acc[i], a[i]) | 1u;
  uint64_t s = 0;
  for (int i = 0; i < kChains; ++i) s ^= acc[i];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ void p_shoup_exact(uint64_t *out, uint64_t q, uint64_t w, uint64_t wp) {
  uint64_t acc[kChains];
  for (int i = 0; i < kChains; ++i) acc[i] = (threadIdx.x * 131 + i) % q;
  for (int it = 0; it < kIter / 4; ++it)
#pragma unroll
    for (int i = 0; i < kChains; ++i) acc[i] = shoup(acc[i], w, wp, q);
  uint64_t s = 0;
  for (int i = 0; i < kChains; ++i) s ^= acc[i];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// forward butterflies in chains: (x, y) -> ct8 -> (x, y), residues stay in [0, 8q)
__global__ void p_ct8(uint64_t *out, uint64_t q, uint64_t w, uint64_t wp, uint64_t nq, uint64_t nq2, uint64_t nq4) {
  const Lazy8 m = lazy8_of(q, nq, nq2, nq4);
  uint64_t x[kChains], y[kChains];
  for (int i = 0; i < kChains; ++i) x[i] = (threadIdx.x * 131 + i) % q, y[i] = (threadIdx.x * 17 + 5 * i) % q;
  for (int it = 0; it < kIter / 4; ++it)
#pragma unroll
    for (int i = 0; i < kChains; ++i) ct8(x[i], y[i], w, wp, m);
  uint64_t s = 0;
  for (int i = 0; i < kChains; ++i) s ^= x[i] ^ y[i];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ void p_gs8(uint64_t *out, uint64_t q, uint64_t w, uint64_t wp, uint64_t nq, uint64_t nq2, uint64_t nq4) {
  const Lazy8 m = lazy8_of(q, nq, nq2, nq4);
  uint64_t x[kChains], y[kChains];
  for (int i = 0; i < kChains; ++i) x[i] = (threadIdx.x * 131 + i) % q, y[i] = (threadIdx.x * 17 + 5 * i) % q;
  for (int it = 0; it < kIter / 4; ++it)
#pragma unroll
    for (int i = 0; i < kChains; ++i) gs8(x[i], y[i], w, wp, m);
  uint64_t s = 0;
  for (int i = 0; i < kChains; ++i) s ^= x[i] ^ y[i];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// narrow towers (q < 2^30): 32-bit butterflies, forward residues in [0, 4q), inverse in [0, 2q)
__global__ void p_ct32(uint64_t *out, uint32_t q, uint32_t w, uint32_t wp) {
  const Narrow m = narrow_of(q);
  uint32_t x[kChains], y[kChains];
  for (int i = 0; i < kChains; ++i) x[i] = (threadIdx.x * 131 + i) % q, y[i] = (threadIdx.x * 17 + 5 * i) % q;
  for (int it = 0; it < kIter / 4; ++it)
#pragma unroll
    for (int i = 0; i < kChains; ++i) ct32(x[i], y[i], w, wp, m);
  uint64_t s = 0;
  for (int i = 0; i < kChains; ++i) s ^= x[i] ^ y[i];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ void p_gs32(uint64_t *out, uint32_t q, uint32_t w, uint32_t wp) {
  const Narrow m = narrow_of(q);
  uint32_t x[kChains], y[kChains];
  for (int i = 0; i < kChains; ++i) x[i] = (threadIdx.x * 131 + i) % q, y[i] = (threadIdx.x * 17 + 5 * i) % q;
  for (int it = 0; it < kIter / 4; ++it)
#pragma unroll
    for (int i = 0; i < kChains; ++i) gs32(x[i], y[i], w, wp, m);
  uint64_t s = 0;
  for (int i = 0; i < kChains; ++i) s ^= x[i] ^ y[i];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

template <class Launch>
float time_ms(Launch launch) {
  hipEvent_t a, b;
  (void)hipEventCreate(&a);
  (void)hipEventCreate(&b);
  launch();
  (void)hipDeviceSynchronize();
  (void)hipEventRecord(a);
  for (int r = 0; r < 5; ++r) launch();
  (void)hipEventRecord(b);
  (void)hipEventSynchronize(b);
  float ms = 0;
  (void)hipEventElapsedTime(&ms, a, b);
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return ms / 5;
}

int main(int argc, char **argv) {
  const bool json = argc > 1 && std::strcmp(argv[1], "--json") == 0;
  hipDeviceProp_t p;
  PROBE_CHECK(hipGetDeviceProperties(&p, 0));
  const int blocks = p.multiProcessorCount * 16, threads = 256;
  uint64_t *buf;
  PROBE_CHECK(hipMalloc(&buf, (size_t)blocks * threads * 8));
  const uint64_t q = 1152921504606584833ull;  // 60-bit NTT prime (2^60 - 2^18 + 1 class)
  const uint64_t w = 123456789012345ull % q;
  const uint64_t wp = (uint64_t)(((unsigned __int128)w << 64) / q);
  const uint32_t q32 = 1073479681u;  // 30-bit NTT prime (class 1 mod 2^17)
  const uint32_t w32 = 123456789u % q32, wp32 = (uint32_t)(((uint64_t)w32 << 32) / q32);
  const double lanes = (double)blocks * threads;
  struct Row {
    const char *name, *unit;
    float ms;
    double ops;
  } rows[] = {
      {"v_mad_u64_u32", "instr", time_ms([&] { p_mad_u64_u32<<<blocks, threads>>>(buf, 1); }), lanes * kIter * kChains},
      {"v_mul_hi_u32+v_or_b32", "step", time_ms([&] { p_mul_hi_u32<<<blocks, threads>>>(buf, 1); }),
       lanes * kIter * kChains},
      {"shoup_mulmod_exact", "mulmod", time_ms([&] { p_shoup_exact<<<blocks, threads>>>(buf, q, w, wp); }),
       lanes * kIter / 4 * kChains},
      {"ct8_forward_butterfly", "butterfly", time_ms([&] { p_ct8<<<blocks, threads>>>(buf, q, w, wp, 0 - q, 0 - 2 * q, 0 - 4 * q); }),
       lanes * kIter / 4 * kChains},
      {"gs8_inverse_butterfly", "butterfly", time_ms([&] { p_gs8<<<blocks, threads>>>(buf, q, w, wp, 0 - q, 0 - 2 * q, 0 - 4 * q); }),
       lanes * kIter / 4 * kChains},
      {"ct32_forward_butterfly", "butterfly", time_ms([&] { p_ct32<<<blocks, threads>>>(buf, q32, w32, wp32); }),
       lanes * kIter / 4 * kChains},
      {"gs32_inverse_butterfly", "butterfly", time_ms([&] { p_gs32<<<blocks, threads>>>(buf, q32, w32, wp32); }),
       lanes * kIter / 4 * kChains},
  };
  PROBE_CHECK(hipGetLastError());
  PROBE_CHECK(hipFree(buf));
  const double clk_hz = p.clockRate * 1e3, cus = p.multiProcessorCount;
  if (!json)
    std::printf("device %s (%s), %d CUs, clock %.0f MHz\n", p.name, p.gcnArchName, p.multiProcessorCount,
                clk_hz / 1e6);
  for (const Row &r : rows) {
    const double per_s = r.ops / (r.ms * 1e-3);
    if (json)
      std::printf("{\"probe\": \"%s\", \"unit\": \"%s\", \"per_s\": %.6e, \"per_clk_per_cu\": %.3f, \"cus\": %d, "
                  "\"clock_mhz\": %.0f}\n",
                  r.name, r.unit, per_s, per_s / clk_hz / cus, p.multiProcessorCount, clk_hz / 1e6);
    else
      std::printf("%-24s %10.4f ms  %12.4e %s/s  %8.3f per clock per CU\n", r.name, r.ms, per_s, r.unit,
                  per_s / clk_hz / cus);
  }
  return 0;
}
