// C ABI of the library: contexts with device tables, cached conversion plans,
// the Driver-seam operations and the fused batched Hybrid key switch.
// Declarations and reference citations: include/ntt_hip.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <functional>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/ntt_hip.h"
#include "conv_device.hpp"

using namespace nttamd;
using u128 = unsigned __int128;

namespace {

thread_local std::string g_last_error;

struct ContractError : std::invalid_argument {
  using std::invalid_argument::invalid_argument;
};
struct RuntimeError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

void require(bool ok, const char *why) {
  if (!ok) throw ContractError(why);
}

#define HIP_CHECK(call)                                                                          \
  do {                                                                                           \
    const hipError_t hip_status_ = (call);                                                       \
    if (hip_status_ != hipSuccess)                                                               \
      throw RuntimeError(std::string("HIP: ") + #call + ": " + hipGetErrorString(hip_status_)); \
  } while (0)

template <class F>
int guarded(F &&f) {
  try {
    return f();
  } catch (const ContractError &e) {
    g_last_error = e.what();
    return NTT_HIP_EINVAL;
  } catch (const std::exception &e) {
    g_last_error = e.what();
    return NTT_HIP_ERUNTIME;
  } catch (...) {
    g_last_error = "unknown failure";
    return NTT_HIP_ERUNTIME;
  }
}

// Restores the caller's current device (reference src/cuda/ntt_kernels.cu:57-75).
class DeviceGuard {
  int previous_ = -1;
  bool changed_ = false;

 public:
  explicit DeviceGuard(int device) {
    HIP_CHECK(hipGetDevice(&previous_));
    if (previous_ != device) {
      HIP_CHECK(hipSetDevice(device));
      changed_ = true;
    }
  }
  ~DeviceGuard() {
    if (changed_) (void)hipSetDevice(previous_);
  }
  DeviceGuard(const DeviceGuard &) = delete;
  DeviceGuard &operator=(const DeviceGuard &) = delete;
};

// ---------------------------------------------------------------- host arithmetic
uint64_t mulmod(uint64_t a, uint64_t b, uint64_t q) { return (uint64_t)((u128)a * b % q); }
uint64_t powmod(uint64_t a, uint64_t e, uint64_t q) {
  uint64_t r = 1 % q;
  a %= q;
  while (e) {
    if (e & 1) r = mulmod(r, a, q);
    a = mulmod(a, a, q);
    e >>= 1;
  }
  return r;
}
uint64_t invmod(uint64_t x, uint64_t q) { return powmod(x % q, q - 2, q); }
uint64_t shoup_of(uint64_t w, uint64_t q) { return (uint64_t)(((u128)w << 64) / q); }
uint64_t neg_inv64(uint64_t q) {  // -q^-1 mod 2^64, q odd (Newton)
  uint64_t x = q;  // correct to 3 bits
  for (int i = 0; i < 6; ++i) x *= 2 - q * x;
  return ~x + 1;
}
uint64_t mont_of(uint64_t x, uint64_t q) { return (uint64_t)(((u128)(x % q) << 64) % q); }
uint32_t bitrev(uint32_t x, uint32_t lg) {
  uint32_t r = 0;
  for (uint32_t i = 0; i < lg; ++i) {
    r = (r << 1) | (x & 1);
    x >>= 1;
  }
  return r;
}
bool is_prime(uint64_t n) {
  if (n < 2) return false;
  static const uint64_t bases[12] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
  for (uint64_t p : bases)
    if (n % p == 0) return n == p;
  uint64_t d = n - 1, s = 0;
  while ((d & 1) == 0) {
    d >>= 1;
    ++s;
  }
  for (uint64_t a : bases) {
    uint64_t x = powmod(a, d, n);
    if (x == 1 || x == n - 1) continue;
    bool witness = true;
    for (uint64_t r = 0; r + 1 < s; ++r) {
      x = mulmod(x, x, n);
      if (x == n - 1) {
        witness = false;
        break;
      }
    }
    if (witness) return false;
  }
  return true;
}

// ---------------------------------------------------------------- device blobs
struct DeviceBlob {
  int device = 0;
  void *ptr = nullptr;
  size_t bytes = 0;
  DeviceBlob() = default;
  DeviceBlob(int dev, size_t n) : device(dev), bytes(n) {
    DeviceGuard g(dev);
    HIP_CHECK(hipMalloc(&ptr, n ? n : 1));
  }
  ~DeviceBlob() {
    if (ptr) {
      try {
        DeviceGuard g(device);
        (void)hipFree(ptr);
      } catch (...) {
      }
    }
  }
  DeviceBlob(const DeviceBlob &) = delete;
  DeviceBlob &operator=(const DeviceBlob &) = delete;
};

struct BasisPlanDev {
  std::unique_ptr<DeviceBlob> blob;
  const BasisConst *dev = nullptr;
  uint32_t s = 0, t = 0;
  bool per_term = false;
};

}  // namespace

struct ntt_hip_context {
  int device = 0;
  uint32_t n = 0, log_n = 0;
  std::vector<uint64_t> q, psi;
  int arith = kLazy;  // NTT arithmetic mode for every tower of the basis (with the largest forward headroom)
  uint32_t ntt_chunk = 0;      // towers per NTT pass pair (0 = all / default), set by the autotuner
  // NTT shape chosen by the autotuner (defaults: the largest forward headroom
  // the moduli allow, 8 residues per thread in the row pass, no occupancy cap)
  int fwd_headroom = -1;       // -1: from the moduli; 0..kMaxHeadroom: at most that
  bool row16 = false;          // 16 residues per thread in the N = 2^16 row pass
  uint32_t wg_cap = 0;         // workgroups per CU for the N = 2^16 passes (0 = no cap)
  // A modulus below 2^24 in the basis: the quotient estimate that headroom
  // classes above 0 use to reduce (reduce2_est) needs 24-bit moduli, so
  // every launch of such a context runs class 0 (a reduction per stage).
  bool small_q = false;
  int ntt_mode() const {
    int a = arith;
    if ((a & kArithMask) == kLazy && (fwd_headroom >= 0 || small_q)) {
      const int h = small_q ? 0 : std::min(fwd_headroom, (a >> kArithShift) & 7);
      a = kLazy | (h << kArithShift);
    }
    return nttamd::ntt_mode(a, row16, wg_cap);
  }
  // The mode of a launch over towers [first, first + count) only: the
  // headroom class of their own largest modulus (at most the tuned
  // fwd_headroom), so narrower towers skip reductions a wider tower of the
  // basis would need, forward and inverse.
  int ntt_mode_for_q(uint64_t max_q) const {
    int a = arith;
    if ((a & kArithMask) == kLazy) {
      int h = small_q ? 0 : headroom_for(max_q);
      if (fwd_headroom >= 0) h = std::min(h, fwd_headroom);
      a = kLazy | (h << kArithShift);
    }
    return nttamd::ntt_mode(a, row16, wg_cap);
  }
  int ntt_mode_for(size_t first, size_t count) const {
    return count ? ntt_mode_for_q(*std::max_element(q.begin() + first, q.begin() + first + count)) : ntt_mode();
  }
  // Standalone transforms (ntt_hip_transform*) with no tuned chunk run the
  // towers in groups whose pass-A output (the hand-off pass B reads back)
  // fits the 256 MB Infinity Cache: measured at N = 2^16 on 2048 towers,
  // 448-512 towers per group make the forward transform 17-27% faster than
  // one group. The fused key switch keeps one group per stage (its stages
  // gained nothing from grouping).
  uint32_t transform_chunk() const {
    if (ntt_chunk) return ntt_chunk;
    if (log_n <= 12) return 0;  // one-kernel transforms: no hand-off between passes
    const uint64_t towers = (uint64_t{224} << 20) / (uint64_t{8} * n);
    return towers >= 64 && towers < 65536 ? (uint32_t)towers : 0;
  }
  uint32_t conv_targets = 0;   // basis-conversion targets per workgroup (0 = from the grid size)
  hipStream_t stream = nullptr;
  std::unique_ptr<DeviceBlob> tables;  // twiddles of every tower
  std::unique_ptr<DeviceBlob> primes;  // PrimeConst[towers]
  std::vector<PrimeConst> host_primes;
  std::mutex mu;
  std::map<std::pair<std::vector<size_t>, std::vector<size_t>>, std::shared_ptr<BasisPlanDev>> basis_cache;
  std::map<std::pair<std::vector<size_t>, std::vector<size_t>>, std::unique_ptr<DeviceBlob>> scale_cache;

  const PrimeConst *dprimes() const { return (const PrimeConst *)primes->ptr; }
  hipStream_t pick(void *s) const { return s ? (hipStream_t)s : stream; }
  void check_range(size_t first, size_t towers) const {
    require(towers > 0 && first < q.size() && towers <= q.size() - first, "tower range outside the context basis");
  }
  ~ntt_hip_context() {
    try {
      DeviceGuard g(device);
      if (stream) {
        (void)hipStreamSynchronize(stream);
        (void)hipStreamDestroy(stream);
      }
    } catch (...) {
    }
  }
};

namespace {

// Approximate CRT plan for context towers src -> dst (reference
// include/ntt/basis_conversion.hpp:32-114): Bhat_i^-1 mod b_i with Shoup
// companion, and (Bhat_i mod m) * 2^64 mod m for Montgomery accumulation.
std::shared_ptr<BasisPlanDev> basis_plan(ntt_hip_context &c, const std::vector<size_t> &src,
                                         const std::vector<size_t> &dst) {
  std::lock_guard<std::mutex> lock(c.mu);
  auto key = std::make_pair(src, dst);
  auto it = c.basis_cache.find(key);
  if (it != c.basis_cache.end()) return it->second;
  const size_t s = src.size(), t = dst.size();
  require(s >= 1 && s <= 40 && t >= 1 && t <= 40, "basis switch requires 1..40 source and target primes");
  std::vector<uint64_t> all;
  for (size_t i : src) {
    require(i < c.q.size(), "source tower outside the context basis");
    all.push_back(c.q[i]);
  }
  for (size_t i : dst) {
    require(i < c.q.size(), "target tower outside the context basis");
    all.push_back(c.q[i]);
  }
  std::sort(all.begin(), all.end());
  require(std::adjacent_find(all.begin(), all.end()) == all.end(), "basis switch requires distinct disjoint primes");
  std::vector<SourceConst> sc(s);
  std::vector<TargetConst> tc(t);
  std::vector<uint64_t> hats(s * t);
  u128 sum_src = 0;
  for (size_t k = 0; k < s; ++k) {
    const uint64_t b = c.q[src[k]];
    sum_src += b;
    uint64_t hat = 1;
    for (size_t o = 0; o < s; ++o)
      if (o != k) hat = mulmod(hat, c.q[src[o]] % b, b);
    sc[k].q = b;
    sc[k].inv = invmod(hat, b);
    sc[k].inv_shoup = shoup_of(sc[k].inv, b);
    for (size_t m = 0; m < t; ++m) {
      const uint64_t qm = c.q[dst[m]];
      uint64_t h = 1;
      for (size_t o = 0; o < s; ++o)
        if (o != k) h = mulmod(h, c.q[src[o]] % qm, qm);
      hats[k * t + m] = mont_of(h, qm);
    }
  }
  for (size_t m = 0; m < t; ++m) {
    tc[m].q = c.q[dst[m]];
    tc[m].qinv_neg = neg_inv64(tc[m].q);
  }
  auto plan = std::make_shared<BasisPlanDev>();
  plan->s = (uint32_t)s;
  plan->t = (uint32_t)t;
  // exact single reduction needs sum_i y_i * hat < m * 2^64, i.e. sum b_i < 2^64
  plan->per_term = sum_src >= ((u128)1 << 64);
  static_assert(sizeof(BasisConst) <= 64, "plan header");
  const size_t off_src = 64, off_dst = off_src + s * sizeof(SourceConst), off_hat = off_dst + t * sizeof(TargetConst);
  const size_t off_limb = off_hat + s * t * sizeof(uint64_t);
  const size_t bytes = off_limb + t * s * 6 * sizeof(uint32_t);
  plan->blob = std::make_unique<DeviceBlob>(c.device, bytes);
  char *base = (char *)plan->blob->ptr;
  BasisConst head{(uint32_t)s, (uint32_t)t, (const SourceConst *)(base + off_src),
                  (const TargetConst *)(base + off_dst), (const uint64_t *)(base + off_hat),
                  (const uint32_t *)(base + off_limb)};
  // register kernel's copy: target-major; for the source's 21-bit limb j the
  // constant hat * 2^(21 j) mod m, as {low, high} 32-bit halves
  std::vector<uint32_t> limbs(t * s * 6);
  for (size_t m = 0; m < t; ++m) {
    const uint64_t qm = tc[m].q;
    for (size_t k = 0; k < s; ++k) {
      uint64_t h = hats[k * t + m];
      for (int j = 0; j < 3; ++j) {
        limbs[((m * s + k) * 3 + j) * 2] = (uint32_t)h;
        limbs[((m * s + k) * 3 + j) * 2 + 1] = (uint32_t)(h >> 32);
        h = mulmod(h, (uint64_t{1} << 21) % qm, qm);
      }
    }
  }
  std::vector<char> host(bytes);
  std::memcpy(host.data(), &head, sizeof(head));
  std::memcpy(host.data() + off_src, sc.data(), s * sizeof(SourceConst));
  std::memcpy(host.data() + off_dst, tc.data(), t * sizeof(TargetConst));
  std::memcpy(host.data() + off_hat, hats.data(), s * t * sizeof(uint64_t));
  std::memcpy(host.data() + off_limb, limbs.data(), limbs.size() * sizeof(uint32_t));
  {
    DeviceGuard g(c.device);
    HIP_CHECK(hipMemcpy(base, host.data(), bytes, hipMemcpyHostToDevice));
  }
  plan->dev = (const BasisConst *)base;
  c.basis_cache.emplace(key, plan);
  return plan;
}

// Per-tower {c, Shoup(c)} constants cached by (towers, P towers); mode 0:
// P^-1 mod q (correction), mode 1: P mod q (rotation).
const ulonglong2 *scale_table(ntt_hip_context &c, size_t first, size_t towers, const std::vector<size_t> &p,
                              int mode) {
  std::lock_guard<std::mutex> lock(c.mu);
  std::vector<size_t> key_t = {first, towers, (size_t)mode};
  auto key = std::make_pair(key_t, p);
  auto it = c.scale_cache.find(key);
  if (it != c.scale_cache.end()) return (const ulonglong2 *)it->second->ptr;
  std::vector<ulonglong2> host(towers);
  for (size_t t = 0; t < towers; ++t) {
    const uint64_t q = c.q[first + t];
    uint64_t prod = 1;
    for (size_t j : p) prod = mulmod(prod, c.q[j] % q, q);
    const uint64_t v = mode == 0 ? invmod(prod, q) : prod;
    host[t].x = v;
    host[t].y = shoup_of(v, q);
  }
  auto blob = std::make_unique<DeviceBlob>(c.device, towers * sizeof(ulonglong2));
  {
    DeviceGuard g(c.device);
    HIP_CHECK(hipMemcpy(blob->ptr, host.data(), towers * sizeof(ulonglong2), hipMemcpyHostToDevice));
  }
  const ulonglong2 *out = (const ulonglong2 *)blob->ptr;
  c.scale_cache.emplace(key, std::move(blob));
  return out;
}

void validate_towers(uint64_t n, const uint64_t *q, const uint64_t *psi, size_t towers) {
  require(n >= 2 && n <= 65536 && (n & (n - 1)) == 0, "requires power-of-two N in [2,65536]");
  require(q && psi && towers > 0 && towers <= 65535, "one (q, psi) pair per tower; 1..65535 towers");
  for (size_t k = 0; k < towers; ++k) {
    require(q[k] > n && q[k] < (uint64_t{1} << 63) && q[k] % (2 * n) == 1 && is_prime(q[k]),
            "requires prime q < 2^63 with 2N dividing q-1");
    require(psi[k] != 0 && psi[k] < q[k] && powmod(psi[k], n, q[k]) == q[k] - 1,
            "requires a canonical primitive 2N-th root");
  }
}

TowerMap affine(uint64_t *base, const uint64_t *src, uint64_t n, size_t towers, size_t first, size_t poly_stride,
                size_t src_poly_stride) {
  TowerMap m{};
  m.base = base;
  m.src = src;
  m.jobs = nullptr;
  m.poly_stride = poly_stride;
  m.src_poly_stride = src_poly_stride;
  m.tower_stride = n;
  m.towers_per_poly = (uint32_t)towers;
  m.prime_offset = (uint32_t)first;
  return m;
}

// Forward-transform map whose last pass writes (sub - NTT(x)) * scale[t]
// (the ModDown correction) to out0 (polynomials < split) / out1 (the rest).
TowerMap with_correction(TowerMap m, const uint64_t *sub, uint64_t sub_poly_stride, uint64_t *out0, uint64_t *out1,
                         uint32_t split, uint64_t out_poly_stride, const ulonglong2 *scale) {
  m.epi.sub = sub;
  m.epi.sub_poly_stride = sub_poly_stride;
  m.epi.out[0] = out0;
  m.epi.out[1] = out1;
  m.epi.split = split;
  m.epi.out_poly_stride = out_poly_stride;
  m.epi.scale = scale;
  return m;
}

void hip_ok(hipError_t e, const char *what) {
  if (e != hipSuccess) throw RuntimeError(std::string("HIP: ") + what + ": " + hipGetErrorString(e));
}

}  // namespace

// ================================================================ key switch plan
struct ntt_hip_keyswitch {
  ntt_hip_context *ctx = nullptr;
  uint32_t L = 0, K = 0, alpha = 0, beta = 0;
  size_t max_batch = 0;
  std::vector<std::shared_ptr<BasisPlanDev>> up;  // per digit: digit towers -> complement
  std::shared_ptr<BasisPlanDev> down;             // P -> Q
  std::shared_ptr<BasisPlanDev> full_up;          // all active Q -> P (ModUp)
  const ulonglong2 *pinv = nullptr;               // P^-1 mod q_i, i < L
  const ulonglong2 *pmod = nullptr;               // P mod q_i, i < L (rotation: B + P c0)
  std::unique_ptr<DeviceBlob> tower_prime;        // identity map 0..L+K
  std::unique_ptr<DeviceBlob> jobs;               // complement NTT jobs for max_batch
  size_t jobs_per_poly = 0;
  std::vector<uint32_t> tower_job_begin;          // [T + 1]: jobs of tower t are [begin[t], begin[t + 1])
  // ModUp complement transforms and the key inner product interleaved by
  // groups of this many towers (0: all towers at once), so the MAC reads a
  // group's digits while they are still in the Infinity Cache. Measured at
  // N = 2^16, L = 24, dnum = 4, batch 64: groups of 2 made the unfused path
  // 1.25% faster and the fused one 0.5% slower.
  uint32_t mac_group = kDefaultMacGroup;
  static constexpr uint32_t kDefaultMacGroup = 0;
  // ModUp's row pass fused with the key inner product (launch_modup_mac_rows)
  // where it applies: 0 off (row pass, then the MAC reads the digits back),
  // 1 with 128-bit sums, 2 with 64-bit sums of half-reduced products. Same
  // workload: 8605 KeySwitch/s with 64-bit sums, 7812 unfused, 7559 with
  // 128-bit sums (their 64 accumulator registers spill at 4 waves per SIMD).
  uint32_t fuse_mac = kDefaultFuseMac;
  static constexpr uint32_t kDefaultFuseMac = 2;
  // log2 of the 64-bit sums' digit bound (beta <= 2^mac_sum_log, 2^(mac_sum_log
  // + 1) max q <= 2^64), or -1 when 64-bit sums do not fit.
  int mac_sum_log = -1;
  std::unique_ptr<DeviceBlob> key;                // beta*(L+K)*n * 2 (b then a), Montgomery form
  bool key_loaded = false;
  bool mac_per_term = false;
  // Batches split into `lanes` sub-batches on concurrent streams (fork/join
  // with events), so one lane's VALU-bound conversion overlaps another's
  // memory-bound transform passes. Streams and events are created on first use.
  uint32_t lanes = kDefaultLanes;
  std::vector<hipStream_t> side;
  std::vector<hipEvent_t> joins;
  hipEvent_t fork = nullptr;
  static constexpr uint32_t kDefaultLanes = 2;
  static constexpr size_t kMinLaneBatch = 4;  // smaller sub-batches are not split

  ~ntt_hip_keyswitch() {
    try {
      DeviceGuard g(ctx->device);
      for (hipStream_t st : side) (void)hipStreamDestroy(st);
      for (hipEvent_t e : joins) (void)hipEventDestroy(e);
      if (fork) (void)hipEventDestroy(fork);
    } catch (...) {
    }
  }
  void ensure_lanes(uint32_t count) {
    if (!fork) HIP_CHECK(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
    while (side.size() + 1 < count) {
      hipStream_t st = nullptr;
      hipEvent_t e = nullptr;
      HIP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));  // default priority
      if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
        (void)hipStreamDestroy(st);
        throw RuntimeError("HIP: lane event creation failed");
      }
      side.push_back(st);  // streams and join events stay paired
      joins.push_back(e);
    }
  }

  size_t n() const { return ctx->n; }
  // workspace layout (elements per component)
  // Per-component strides of the workspace buffers, padded by 64 KiB: launches
  // run tower-major (the same tower of consecutive components side by side),
  // and unpadded strides (multiples of 512 KiB) put those concurrent accesses
  // on the same HBM channels (measured 6% slower NTT passes).
  static constexpr size_t kPad = 8192;
  size_t coeff_words() const { return (size_t)L * n() + kPad; }
  size_t digit_words() const { return (size_t)beta * (L + K) * n() + kPad; }
  size_t z_words() const { return (size_t)(L + K) * n() + kPad; }
  size_t pc_words() const { return (size_t)K * n() + kPad; }
  size_t corr_words() const { return (size_t)L * n() + kPad; }
  size_t per_component() const {
    return coeff_words() + digit_words() + 2 * z_words() + 2 * pc_words() + 2 * corr_words();
  }
};

extern "C" {

const char *ntt_hip_last_error(void) { return g_last_error.c_str(); }
int ntt_hip_abi_version(void) { return 1; }

int ntt_hip_device_count(int *count) {
  return guarded([&] {
    require(count != nullptr, "null count");
    HIP_CHECK(hipGetDeviceCount(count));
    return NTT_HIP_OK;
  });
}

int ntt_hip_device_name(int device, char *name, size_t capacity) {
  return guarded([&] {
    require(name && capacity > 0, "null name buffer");
    hipDeviceProp_t p;
    HIP_CHECK(hipGetDeviceProperties(&p, device));
    // some drivers leave the marketing name empty: then the ordinal stands in for it
    char fallback[32];
    std::snprintf(fallback, sizeof fallback, "HIP device %d", device);
    std::snprintf(name, capacity, "%s (%s, %d CUs)", p.name[0] ? p.name : fallback, p.gcnArchName,
                  p.multiProcessorCount);
    return NTT_HIP_OK;
  });
}

int ntt_hip_device_properties(int device, char *arch, size_t arch_capacity, uint64_t *memory_bytes,
                              int *compute_units, int *driver_version, int *runtime_version) {
  return guarded([&] {
    require(arch && arch_capacity > 0 && memory_bytes && compute_units && driver_version && runtime_version,
            "null argument");
    hipDeviceProp_t p;
    HIP_CHECK(hipGetDeviceProperties(&p, device));
    HIP_CHECK(hipDriverGetVersion(driver_version));
    HIP_CHECK(hipRuntimeGetVersion(runtime_version));
    std::snprintf(arch, arch_capacity, "%s", p.gcnArchName);
    *memory_bytes = p.totalGlobalMem;
    *compute_units = p.multiProcessorCount;
    return NTT_HIP_OK;
  });
}

int ntt_hip_malloc(int device, void **ptr, size_t bytes) {
  return guarded([&] {
    require(ptr != nullptr && bytes > 0, "invalid allocation request");
    DeviceGuard g(device);
    HIP_CHECK(hipMalloc(ptr, bytes));
    return NTT_HIP_OK;
  });
}

int ntt_hip_free(int device, void *ptr) {
  return guarded([&] {
    DeviceGuard g(device);
    HIP_CHECK(hipFree(ptr));
    return NTT_HIP_OK;
  });
}

int ntt_hip_host_alloc(void **ptr, size_t bytes) {
  return guarded([&] {
    require(ptr != nullptr && bytes > 0, "invalid allocation request");
    HIP_CHECK(hipHostMalloc(ptr, bytes, hipHostMallocDefault));
    return NTT_HIP_OK;
  });
}

int ntt_hip_host_free(void *ptr) {
  return guarded([&] {
    HIP_CHECK(hipHostFree(ptr));
    return NTT_HIP_OK;
  });
}

int ntt_hip_memcpy(int device, void *dst, const void *src, size_t bytes, int kind, void *stream) {
  return guarded([&] {
    require(dst && src, "null copy pointer");
    require(kind == NTT_HIP_H2D || kind == NTT_HIP_D2H || kind == NTT_HIP_D2D, "invalid copy kind");
    if (bytes == 0) return NTT_HIP_OK;
    const hipMemcpyKind k = kind == NTT_HIP_H2D   ? hipMemcpyHostToDevice
                            : kind == NTT_HIP_D2H ? hipMemcpyDeviceToHost
                                                  : hipMemcpyDeviceToDevice;
    DeviceGuard g(device);
    HIP_CHECK(hipMemcpyAsync(dst, src, bytes, k, (hipStream_t)stream));
    return NTT_HIP_OK;
  });
}

int ntt_hip_stream_create(int device, void **stream) {
  return guarded([&] {
    require(stream != nullptr, "null stream out");
    DeviceGuard g(device);
    hipStream_t s;
    HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
    return NTT_HIP_OK;
  });
}

int ntt_hip_stream_destroy(int device, void *stream) {
  return guarded([&] {
    DeviceGuard g(device);
    HIP_CHECK(hipStreamDestroy((hipStream_t)stream));
    return NTT_HIP_OK;
  });
}

int ntt_hip_stream_synchronize(int device, void *stream) {
  return guarded([&] {
    DeviceGuard g(device);
    HIP_CHECK(hipStreamSynchronize((hipStream_t)stream));
    return NTT_HIP_OK;
  });
}

int ntt_hip_event_create(int device, void **event, int timing) {
  return guarded([&] {
    require(event != nullptr, "null event out");
    DeviceGuard g(device);
    hipEvent_t e;
    HIP_CHECK(hipEventCreateWithFlags(&e, timing ? hipEventDefault : hipEventDisableTiming));
    *event = e;
    return NTT_HIP_OK;
  });
}

int ntt_hip_event_destroy(int device, void *event) {
  return guarded([&] {
    DeviceGuard g(device);
    HIP_CHECK(hipEventDestroy((hipEvent_t)event));
    return NTT_HIP_OK;
  });
}

int ntt_hip_event_record(int device, void *event, void *stream) {
  return guarded([&] {
    DeviceGuard g(device);
    HIP_CHECK(hipEventRecord((hipEvent_t)event, (hipStream_t)stream));
    return NTT_HIP_OK;
  });
}

int ntt_hip_event_query(int device, void *event) {
  return guarded([&] {
    DeviceGuard g(device);
    const hipError_t s = hipEventQuery((hipEvent_t)event);
    if (s == hipErrorNotReady) return NTT_HIP_NOT_READY;
    HIP_CHECK(s);
    return NTT_HIP_OK;
  });
}

int ntt_hip_event_synchronize(int device, void *event) {
  return guarded([&] {
    DeviceGuard g(device);
    HIP_CHECK(hipEventSynchronize((hipEvent_t)event));
    return NTT_HIP_OK;
  });
}

int ntt_hip_event_elapsed(float *ms, void *start, void *stop) {
  return guarded([&] {
    require(ms != nullptr, "null elapsed out");
    HIP_CHECK(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
    return NTT_HIP_OK;
  });
}

// ---------------------------------------------------------------- context
int ntt_hip_context_create(ntt_hip_context **out, int device, uint64_t n, const uint64_t *q, const uint64_t *psi,
                           size_t towers) {
  return guarded([&] {
    require(out != nullptr, "null context out");
    require(device >= 0, "invalid device ordinal");
    validate_towers(n, q, psi, towers);
    auto c = std::make_unique<ntt_hip_context>();
    c->device = device;
    c->n = (uint32_t)n;
    c->log_n = 0;
    while ((1u << c->log_n) < n) ++c->log_n;
    c->q.assign(q, q + towers);
    c->psi.assign(psi, psi + towers);
    // 32-bit butterflies when every q < 2^30; wide-lazy (residues below 8q < 2^64) when every q < 2^61
    c->arith = arith_for(*std::max_element(c->q.begin(), c->q.end()));
    c->small_q = *std::min_element(c->q.begin(), c->q.end()) < (uint64_t{1} << 24);
    // Host tables: fwd[k] = {psi^bitrev(k), Shoup}, inv[k] = {psi^-bitrev(k), Shoup}
    // (reference transform_stage.hpp:19-30), one thread per group of towers.
    std::vector<ulonglong2> host(towers * 2 * (size_t)n);
    auto build = [&](size_t t0, size_t t1) {
      for (size_t t = t0; t < t1; ++t) {
        const uint64_t qt = c->q[t];
        for (int dir = 0; dir < 2; ++dir) {
          const uint64_t root = dir == 0 ? c->psi[t] : invmod(c->psi[t], qt);
          ulonglong2 *tab = host.data() + (t * 2 + dir) * (size_t)n;
          uint64_t power = 1;
          for (uint32_t e = 0; e < n; ++e) {
            ulonglong2 &w = tab[bitrev(e, c->log_n)];
            w.x = power;
            w.y = shoup_of(power, qt);
            power = mulmod(power, root, qt);
          }
        }
      }
    };
    const size_t workers = std::min<size_t>(towers, std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
    std::vector<std::thread> pool;
    for (size_t w = 0; w < workers; ++w)
      pool.emplace_back(build, towers * w / workers, towers * (w + 1) / workers);
    for (auto &th : pool) th.join();
    DeviceGuard g(device);
    HIP_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    c->tables = std::make_unique<DeviceBlob>(device, host.size() * sizeof(ulonglong2));
    HIP_CHECK(hipMemcpy(c->tables->ptr, host.data(), host.size() * sizeof(ulonglong2), hipMemcpyHostToDevice));
    const ulonglong2 *dt = (const ulonglong2 *)c->tables->ptr;
    c->host_primes.resize(towers);
    for (size_t t = 0; t < towers; ++t) {
      PrimeConst &p = c->host_primes[t];
      const uint64_t qt = c->q[t];
      p.q = qt;
      p.two_q = 2 * qt;
      p.nq = 0 - qt;
      p.nq2 = 0 - 2 * qt;
      p.nq4 = 0 - 4 * qt;
      p.qinv_neg = neg_inv64(qt);
      p.ninv = invmod(n % qt, qt);
      p.ninv_shoup = shoup_of(p.ninv, qt);
      const uint64_t w1 = host[(t * 2 + 1) * (size_t)n + 1].x;  // psi^-bitrev(1)
      p.last_w = mulmod(w1, p.ninv, qt);
      p.last_wp = shoup_of(p.last_w, qt);
      p.r2 = (uint64_t)(((u128)mont_of(1, qt) << 64) % qt);  // 2^128 mod q
      const u128 mu = ~(u128)0 / qt;  // = floor(2^128 / q): q is odd, so it does not divide 2^128
      p.mu_lo = (uint64_t)mu;
      p.mu_hi = (uint64_t)(mu >> 64);
      const int bits = 64 - __builtin_clzll(qt);
      p.est_shift = bits >= 24 ? (uint32_t)(bits - 24) : 0;
      p.est_mul = bits >= 24 ? (uint32_t)(((u128)1 << (bits + 31)) / qt) : 0;
      p.fwd = dt + (t * 2) * (size_t)n;
      p.inv = dt + (t * 2 + 1) * (size_t)n;
    }
    c->primes = std::make_unique<DeviceBlob>(device, towers * sizeof(PrimeConst));
    HIP_CHECK(hipMemcpy(c->primes->ptr, c->host_primes.data(), towers * sizeof(PrimeConst), hipMemcpyHostToDevice));
    *out = c.release();
    return NTT_HIP_OK;
  });
}

int ntt_hip_context_destroy(ntt_hip_context *ctx) {
  return guarded([&] {
    delete ctx;
    return NTT_HIP_OK;
  });
}

void *ntt_hip_context_stream(ntt_hip_context *ctx) { return ctx ? (void *)ctx->stream : nullptr; }
int ntt_hip_context_device(ntt_hip_context *ctx) { return ctx ? ctx->device : -1; }
size_t ntt_hip_context_table_bytes(ntt_hip_context *ctx) {
  return ctx ? ctx->tables->bytes + ctx->primes->bytes : 0;
}

int ntt_hip_context_device_bytes(ntt_hip_context *ctx, size_t *bytes) {
  return guarded([&] {
    require(ctx && bytes, "null argument");
    std::lock_guard<std::mutex> lock(ctx->mu);
    size_t total = ctx->tables->bytes + ctx->primes->bytes;
    for (const auto &kv : ctx->basis_cache)
      if (kv.second && kv.second->blob) total += kv.second->blob->bytes;
    for (const auto &kv : ctx->scale_cache)
      if (kv.second) total += kv.second->bytes;
    *bytes = total;
    return NTT_HIP_OK;
  });
}

int ntt_hip_transform_jobs(ntt_hip_context *ctx, uint64_t *data, const uint64_t *src, const ntt_hip_tower_job *jobs,
                           size_t count, void *jobs_workspace, int inverse, void *stream) {
  return guarded([&] {
    require(ctx && data && jobs && jobs_workspace, "null argument");
    require(count >= 1 && count <= 65535, "1..65535 tower jobs per launch");
    static_assert(sizeof(ntt_hip_tower_job) == sizeof(TowerJob), "job record layout");
    for (size_t j = 0; j < count; ++j)
      require(jobs[j].tower < ctx->q.size(), "job tower outside the context basis");
    const hipStream_t s = ctx->pick(stream);
    DeviceGuard g(ctx->device);
    // the table is staged from the caller's host array on the launch stream
    HIP_CHECK(hipMemcpyAsync(jobs_workspace, jobs, count * sizeof(TowerJob), hipMemcpyHostToDevice, s));
    TowerMap m{};
    m.base = data;
    m.src = src;
    m.jobs = (const TowerJob *)jobs_workspace;
    m.poly_stride = 0;
    m.src_poly_stride = 0;
    m.tower_stride = ctx->n;
    m.towers_per_poly = (uint32_t)count;
    uint64_t max_q = 0;
    for (size_t j = 0; j < count; ++j) max_q = std::max(max_q, ctx->q[jobs[j].tower]);
    hip_ok(launch_ntt(m, ctx->dprimes(), ctx->log_n, (uint32_t)count, inverse != 0, ctx->ntt_mode_for_q(max_q), s,
                      ctx->transform_chunk()),
           "ntt job launch");
    return NTT_HIP_OK;
  });
}

int ntt_hip_transform_mixed_workspace_bytes(size_t count, size_t *bytes) {
  return guarded([&] {
    require(bytes, "null argument");
    require(count >= 1 && count <= (size_t{1} << 20), "1..2^20 mixed jobs per launch");
    *bytes = count * sizeof(MixedJob);
    return NTT_HIP_OK;
  });
}

int ntt_hip_transform_mixed(const ntt_hip_mixed_job *jobs, size_t count, void *jobs_workspace, int inverse,
                            void *stream) {
  return guarded([&] {
    require(jobs && jobs_workspace, "null argument");
    require(count >= 1 && count <= (size_t{1} << 20), "1..2^20 mixed jobs per launch");
    const ntt_hip_context *first = jobs[0].ctx;
    require(first != nullptr, "null context");
    // the table, one run per arithmetic (32-bit towers, then wide-lazy ones)
    std::vector<MixedJob> table;
    table.reserve(count);
    uint32_t runs[2][2] = {{0, 0}, {0, 0}};  // [narrow, lazy] x {jobs, blocks}
    for (int pass = 0; pass < 2; ++pass) {
      const int want = pass == 0 ? kNarrow : kLazy;
      for (size_t j = 0; j < count; ++j) {
        const ntt_hip_context *c = jobs[j].ctx;
        require(c != nullptr && jobs[j].data != nullptr, "null context or buffer");
        require(c->device == first->device, "mixed jobs on one device");
        require(c->log_n >= 13 && c->log_n <= 16, "mixed launches take N = 2^13 .. 2^16");
        require(jobs[j].tower < c->q.size(), "job tower outside its context basis");
        const int a = c->arith & kArithMask;
        require(a == kLazy || a == kNarrow, "mixed launches take moduli below 2^61");
        if (a != want) continue;
        MixedJob m{};
        m.data = jobs[j].data;
        m.src = jobs[j].src ? jobs[j].src : jobs[j].data;
        m.prime = c->dprimes() + jobs[j].tower;
        m.log_n = c->log_n;
        m.first = runs[pass][1];
        runs[pass][0] += 1;
        runs[pass][1] += (uint32_t)(c->n / 4096);
        table.push_back(m);
      }
    }
    const hipStream_t s = first->pick(stream);
    DeviceGuard g(first->device);
    HIP_CHECK(hipMemcpyAsync(jobs_workspace, table.data(), table.size() * sizeof(MixedJob), hipMemcpyHostToDevice, s));
    const MixedJob *dev = (const MixedJob *)jobs_workspace;
    if (runs[0][0])
      hip_ok(launch_ntt_mixed(dev, runs[0][0], runs[0][1], inverse != 0, kNarrow, s), "mixed ntt launch");
    if (runs[1][0])
      hip_ok(launch_ntt_mixed(dev + runs[0][0], runs[1][0], runs[1][1], inverse != 0, kLazy, s), "mixed ntt launch");
    return NTT_HIP_OK;
  });
}

int ntt_hip_transform(ntt_hip_context *ctx, uint64_t *data, const uint64_t *src, size_t first, size_t towers,
                      int inverse, void *stream) {
  return ntt_hip_transform_batch(ctx, data, src, 1, 0, 0, first, towers, inverse, stream);
}

int ntt_hip_transform_batch(ntt_hip_context *ctx, uint64_t *data, const uint64_t *src, size_t polys,
                            size_t poly_stride, size_t src_poly_stride, size_t first, size_t towers, int inverse,
                            void *stream) {
  return guarded([&] {
    require(ctx && data, "null context or buffer");
    ctx->check_range(first, towers);
    require(polys >= 1 && polys * towers <= 65535, "1..65535 tower transforms per launch");
    require(polys == 1 || (poly_stride >= towers * ctx->n && (!src || src_poly_stride >= towers * ctx->n)),
            "polynomial stride shorter than its towers");
    DeviceGuard g(ctx->device);
    const TowerMap m = affine(data, src, ctx->n, towers, first, poly_stride, src_poly_stride);
    hip_ok(launch_ntt(m, ctx->dprimes(), ctx->log_n, (uint32_t)(polys * towers), inverse != 0,
                      ctx->ntt_mode_for(first, towers), ctx->pick(stream), ctx->transform_chunk()),
           "ntt launch");
    return NTT_HIP_OK;
  });
}

int ntt_hip_pointwise(ntt_hip_context *ctx, uint64_t *out, const uint64_t *right, size_t first, size_t towers,
                      void *stream) {
  return guarded([&] {
    require(ctx && out && right, "null context or buffer");
    ctx->check_range(first, towers);
    DeviceGuard g(ctx->device);
    hip_ok(launch_pointwise(out, right, ctx->dprimes(), (uint32_t)first, ctx->n, (uint32_t)towers, ctx->pick(stream)),
           "pointwise launch");
    return NTT_HIP_OK;
  });
}

int ntt_hip_switch_basis(ntt_hip_context *ctx, uint64_t *out, const uint64_t *in, const size_t *src, size_t s,
                         const size_t *dst, size_t t, void *stream) {
  return guarded([&] {
    require(ctx && out && in && src && dst, "null argument");
    auto plan = basis_plan(*ctx, std::vector<size_t>(src, src + s), std::vector<size_t>(dst, dst + t));
    BasisLaunch a{};
    a.plan = plan->dev;
    a.in = in;
    a.out = out;
    a.n = ctx->n;
    a.targets_per_block = (uint32_t)std::min<size_t>(t, ctx->conv_targets);
    a.plan_host_s = plan->s;
    a.plan_host_t = plan->t;
    a.skip_at = 0xffffffffu;
    a.skip_len = 0;
    DeviceGuard g(ctx->device);
    hip_ok(launch_basis_switch(a, plan->per_term, 1, ctx->pick(stream)), "basis switch launch");
    return NTT_HIP_OK;
  });
}

int ntt_hip_correct(ntt_hip_context *ctx, uint64_t *out, const uint64_t *left, const uint64_t *right, size_t first,
                    size_t towers, const size_t *p, size_t k, void *stream) {
  return guarded([&] {
    require(ctx && out && left && right && p && k > 0, "null argument");
    ctx->check_range(first, towers);
    std::vector<size_t> pv(p, p + k);
    for (size_t j : pv) require(j < ctx->q.size() && (j < first || j >= first + towers), "invalid P tower");
    CorrectLaunch c{};
    c.out = out;
    c.left = left;
    c.right = right;
    c.primes = ctx->dprimes();
    c.scale = scale_table(*ctx, first, towers, pv, 0);
    c.first_prime = (uint32_t)first;
    c.n = ctx->n;
    c.towers = (uint32_t)towers;
    DeviceGuard g(ctx->device);
    hip_ok(launch_correct(c, 1, ctx->pick(stream)), "correction launch");
    return NTT_HIP_OK;
  });
}

int ntt_hip_hybrid_mac(ntt_hip_context *ctx, uint64_t *b, uint64_t *a, const uint64_t *digit, const uint64_t *key_b,
                       const uint64_t *key_a, size_t first, size_t towers, int first_digit, void *stream) {
  return guarded([&] {
    require(ctx && b && a && digit && key_b && key_a, "null argument");
    ctx->check_range(first, towers);
    DeviceGuard g(ctx->device);
    hip_ok(launch_mac_digit(b, a, digit, key_b, key_a, ctx->dprimes(), (uint32_t)first, ctx->n, (uint32_t)towers,
                            first_digit != 0, ctx->pick(stream)),
           "hybrid mac launch");
    return NTT_HIP_OK;
  });
}

int ntt_hip_hybrid_mac_digits(ntt_hip_context *ctx, uint64_t *b, uint64_t *a, const uint64_t *const *digits,
                              const uint64_t *const *key_b, const uint64_t *const *key_a, size_t count, size_t first,
                              size_t towers, int first_digit, void *stream) {
  return guarded([&] {
    require(ctx && b && a && digits && key_b && key_a, "null argument");
    require(count >= 1, "no digits");
    for (size_t j = 0; j < count; ++j) require(digits[j] && key_b[j] && key_a[j], "null digit or key row");
    ctx->check_range(first, towers);
    DeviceGuard g(ctx->device);
    // Launches of at most kMacDigitsPerLaunch digits; each after the first accumulates.
    for (size_t j0 = 0; j0 < count; j0 += kMacDigitsPerLaunch) {
      MacDigitsLaunch m{};
      m.b = b;
      m.a = a;
      m.count = (uint32_t)std::min<size_t>(kMacDigitsPerLaunch, count - j0);
      for (uint32_t j = 0; j < m.count; ++j) {
        m.digits[j] = digits[j0 + j];
        m.key_b[j] = key_b[j0 + j];
        m.key_a[j] = key_a[j0 + j];
      }
      m.primes = ctx->dprimes();
      m.first_prime = (uint32_t)first;
      m.n = ctx->n;
      m.towers = (uint32_t)towers;
      m.first_digit = (first_digit != 0 && j0 == 0) ? 1 : 0;
      hip_ok(launch_mac_digits(m, ctx->pick(stream)), "hybrid mac launch");
    }
    return NTT_HIP_OK;
  });
}

int ntt_hip_hybrid_rotate(ntt_hip_context *ctx, uint64_t *out_b, uint64_t *out_a, const uint64_t *ext_b,
                          const uint64_t *ext_a, const uint64_t *c0, size_t first, size_t towers, size_t active_q,
                          const size_t *p, size_t k, uint32_t automorphism, void *stream) {
  return guarded([&] {
    require(ctx && out_b && out_a && ext_b && ext_a && c0 && p && k > 0, "null argument");
    ctx->check_range(first, towers);
    require(active_q >= 1 && active_q <= towers, "invalid active Q count");
    require((automorphism & 1) && automorphism < 2 * ctx->n, "invalid negacyclic automorphism index");
    std::vector<size_t> pv(p, p + k);
    // P towers are context towers outside the rotated Q range (the same rule as ntt_hip_correct)
    for (size_t j : pv) require(j < ctx->q.size() && (j < first || j >= first + active_q), "invalid P tower");
    RotateLaunch r{};
    r.out_b = out_b;
    r.out_a = out_a;
    r.ext_b = ext_b;
    r.ext_a = ext_a;
    r.c0 = c0;
    r.primes = ctx->dprimes();
    r.scale = scale_table(*ctx, first, active_q, pv, 1);
    r.first_prime = (uint32_t)first;
    r.n = ctx->n;
    r.log_n = ctx->log_n;
    r.active_q = (uint32_t)active_q;
    r.automorphism = automorphism;
    DeviceGuard g(ctx->device);
    hip_ok(launch_rotate(r, (uint32_t)towers, 1, ctx->pick(stream)), "rotation launch");
    return NTT_HIP_OK;
  });
}

// ---------------------------------------------------------------- fused key switch
int ntt_hip_keyswitch_create(ntt_hip_keyswitch **out, ntt_hip_context *ctx, size_t L, size_t K, size_t alpha,
                             size_t max_batch) {
  return guarded([&] {
    require(out && ctx, "null argument");
    require(L >= 1 && K >= 1 && L + K == ctx->q.size(), "context basis must be L Q towers then K P towers");
    require(alpha >= 1, "invalid digit size");  // alpha may exceed L: one shortened digit
    require(max_batch >= 1, "invalid batch capacity");
    auto ks = std::make_unique<ntt_hip_keyswitch>();
    ks->ctx = ctx;
    ks->L = (uint32_t)L;
    ks->K = (uint32_t)K;
    ks->alpha = (uint32_t)alpha;
    ks->beta = (uint32_t)((L + alpha - 1) / alpha);
    ks->max_batch = max_batch;
    const size_t T = L + K, n = ctx->n;
    for (uint32_t d = 0; d < ks->beta; ++d) {
      const size_t first = (size_t)d * alpha, count = std::min(alpha, L - first);
      std::vector<size_t> src, dst;
      for (size_t i = 0; i < count; ++i) src.push_back(first + i);
      for (size_t i = 0; i < T; ++i)
        if (i < first || i >= first + count) dst.push_back(i);
      // the 40-prime plan limit applies per target chunk (reference src/resident.cpp:474-484)
      require(count <= 40, "digit wider than 40 towers");
      require(dst.size() <= 40, "complement wider than 40 towers");
      ks->up.push_back(basis_plan(*ctx, src, dst));
    }
    std::vector<size_t> p, q;
    for (size_t i = 0; i < K; ++i) p.push_back(L + i);
    for (size_t i = 0; i < L; ++i) q.push_back(i);
    require(K <= 40 && L <= 40, "ModDown limited to 40 towers per basis");
    ks->down = basis_plan(*ctx, p, q);
    ks->full_up = basis_plan(*ctx, q, p);
    ks->pinv = scale_table(*ctx, 0, L, p, 0);
    ks->pmod = scale_table(*ctx, 0, L, p, 1);
    std::vector<uint32_t> tp(T);
    for (size_t i = 0; i < T; ++i) tp[i] = (uint32_t)i;
    ks->tower_prime = std::make_unique<DeviceBlob>(ctx->device, T * sizeof(uint32_t));
    // complement NTT jobs of one component (digit d, tower t not owned by d),
    // ordered by prime; component c adds c * digit_words
    std::vector<TowerJob> jobs;
    for (size_t t = 0; t < T; ++t)
      for (uint32_t d = 0; d < ks->beta; ++d) {
        const size_t first = (size_t)d * alpha, count = std::min(alpha, L - first);
        if (t >= first && t < first + count) continue;
        TowerJob j{};
        j.offset = ((size_t)d * T + t) * n;
        j.src_offset = j.offset;
        j.prime = (uint32_t)t;
        jobs.push_back(j);
      }
    ks->jobs_per_poly = jobs.size();
    // A launch of the plan takes at most this many workgroup rows (grid y) per
    // component: the complement transforms, both ModDown channels, or
    // conversion chunks of up to 40 targets for both channels. HIP allows
    // 65536 rows per launch.
    const size_t rows_per_component = std::max({ks->jobs_per_poly, 2 * T, size_t{80}});
    require(max_batch <= 65535 / rows_per_component, "batch capacity exceeds the rows of one launch");
    ks->tower_job_begin.assign(T + 1, 0);
    for (const TowerJob &j : jobs) ++ks->tower_job_begin[j.prime + 1];
    for (size_t t = 0; t < T; ++t) ks->tower_job_begin[t + 1] += ks->tower_job_begin[t];
    ks->jobs = std::make_unique<DeviceBlob>(ctx->device, jobs.size() * sizeof(TowerJob));
    // Montgomery single-reduction bound for the MAC: beta * max q < 2^64
    u128 bound = 0;
    for (size_t i = 0; i < T; ++i) bound = std::max<u128>(bound, (u128)ctx->q[i] * ks->beta);
    ks->mac_per_term = bound >= ((u128)1 << 64);
    int sum_log = 0;
    while ((1u << sum_log) < ks->beta) ++sum_log;
    const uint64_t max_q = *std::max_element(ctx->q.begin(), ctx->q.end());
    ks->mac_sum_log = sum_log <= 6 && ((u128)max_q << (sum_log + 1)) <= ((u128)1 << 64) ? sum_log : -1;
    ks->key = std::make_unique<DeviceBlob>(ctx->device, 2 * ks->beta * T * n * sizeof(uint64_t));
    DeviceGuard g(ctx->device);
    HIP_CHECK(hipMemcpy(ks->tower_prime->ptr, tp.data(), T * sizeof(uint32_t), hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(ks->jobs->ptr, jobs.data(), jobs.size() * sizeof(TowerJob), hipMemcpyHostToDevice));
    *out = ks.release();
    return NTT_HIP_OK;
  });
}

int ntt_hip_keyswitch_device_bytes(ntt_hip_keyswitch *ks, size_t *bytes) {
  return guarded([&] {
    require(ks && bytes, "null argument");
    size_t total = 0;
    for (const DeviceBlob *b : {ks->key.get(), ks->jobs.get(), ks->tower_prime.get()})
      if (b) total += b->bytes;
    *bytes = total;
    return NTT_HIP_OK;
  });
}

int ntt_hip_keyswitch_destroy(ntt_hip_keyswitch *ks) {
  return guarded([&] {
    delete ks;
    return NTT_HIP_OK;
  });
}

int ntt_hip_keyswitch_upload_key(ntt_hip_keyswitch *ks, const uint64_t *key_b, const uint64_t *key_a, size_t M,
                                 const uint64_t *q_full) {
  return guarded([&] {
    require(ks && key_b && key_a && q_full, "null argument");
    const size_t L = ks->L, K = ks->K, T = L + K, n = ks->n(), full = M + K;
    require(M >= L, "full context has fewer Q towers than the active level");
    for (size_t i = 0; i < L; ++i) require(q_full[i] == ks->ctx->q[i], "full-context Q prefix mismatch");
    const size_t full_digits = (M + ks->alpha - 1) / ks->alpha;
    require(full_digits >= ks->beta, "key has fewer digits than the active level needs");
    // canonical check and active row selection (i < L: i; P: i + (M - L))
    std::vector<uint64_t> host(2 * (size_t)ks->beta * T * n);
    for (int ch = 0; ch < 2; ++ch) {
      const uint64_t *src = ch == 0 ? key_b : key_a;
      for (size_t d = 0; d < ks->beta; ++d)
        for (size_t t = 0; t < T; ++t) {
          const size_t kt = t < L ? t : t + (M - L);
          const uint64_t q = ks->ctx->q[t];
          const uint64_t *row = src + (d * full + kt) * n;
          uint64_t *dst = host.data() + ((size_t)ch * ks->beta * T + d * T + t) * n;
          for (size_t i = 0; i < n; ++i) {
            require(row[i] < q, "noncanonical key residue");
            dst[i] = row[i];
          }
        }
    }
    DeviceGuard g(ks->ctx->device);
    hipStream_t s = ks->ctx->stream;
    HIP_CHECK(hipMemcpyAsync(ks->key->ptr, host.data(), host.size() * sizeof(uint64_t), hipMemcpyHostToDevice, s));
    // to Montgomery form: rows are (channel, digit, tower): tower prime = t
    std::vector<uint32_t> tp(2 * ks->beta * T);
    for (size_t r = 0; r < tp.size(); ++r) tp[r] = (uint32_t)(r % T);
    DeviceBlob tpd(ks->ctx->device, tp.size() * sizeof(uint32_t));
    HIP_CHECK(hipMemcpyAsync(tpd.ptr, tp.data(), tp.size() * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    hip_ok(launch_to_mont((uint64_t *)ks->key->ptr, ks->ctx->dprimes(), (const uint32_t *)tpd.ptr, (uint32_t)n,
                          (uint32_t)tp.size(), s),
           "key conversion launch");
    HIP_CHECK(hipStreamSynchronize(s));
    ks->key_loaded = true;
    return NTT_HIP_OK;
  });
}

}  // extern "C"

// ---------------------------------------------------------------- key-switch stages
namespace {
struct KsBuffers {
  uint64_t *coeff, *digits, *z, *pc, *corr;
};

// The components' evaluation-form input: contiguous (base + c * L * n) or,
// when `table` is set, at per-component device addresses table[c].
struct KsIn {
  const uint64_t *base = nullptr;
  const uint64_t *const *table = nullptr;
  KsIn(const uint64_t *b) : base(b) {}
  KsIn(const uint64_t *const *t, int) : table(t) {}
};

KsBuffers ks_buffers(const ntt_hip_keyswitch &ks, void *workspace, size_t B) {
  KsBuffers w;
  w.coeff = (uint64_t *)workspace;                   // B * coeff_words
  w.digits = w.coeff + B * ks.coeff_words();         // B * digit_words (beta digits of T towers)
  w.z = w.digits + B * ks.digit_words();             // [2][B] * z_words (b channel, then a)
  w.pc = w.z + 2 * B * ks.z_words();                 // [2][B] * pc_words
  w.corr = w.pc + 2 * B * ks.pc_words();             // [2][B] * corr_words
  return w;
}

// ModUp of B components (evaluation Q input, stride L*n): coefficient form of
// every Q tower, per digit the unreduced CRT into its complement, then the
// forward NTT of every complement tower (one launch, jobs sorted by prime).
TowerMap complement_jobs(const ntt_hip_keyswitch &ks, const KsBuffers &w) {
  TowerMap m{};
  m.base = w.digits;
  m.jobs = (const TowerJob *)ks.jobs->ptr;
  m.tower_stride = ks.n();
  m.poly_stride = ks.digit_words();
  m.towers_per_poly = (uint32_t)ks.jobs_per_poly;
  return m;
}

// INTT of the Q towers and, per digit, the unreduced CRT into its complement
// (coefficient form, in the digit buffers).
void ks_modup_convert(const ntt_hip_keyswitch &ks, KsIn in, size_t B, const KsBuffers &w, hipStream_t s);

void ks_modup(const ntt_hip_keyswitch &ks, KsIn in, size_t B, const KsBuffers &w, hipStream_t s) {
  const ntt_hip_context &c = *ks.ctx;
  ks_modup_convert(ks, in, B, w, s);
  const TowerMap m = complement_jobs(ks, w);
  if (c.ntt_chunk) {
    hip_ok(launch_ntt(m, c.dprimes(), c.log_n, (uint32_t)(B * ks.jobs_per_poly), false, c.ntt_mode(), s, c.ntt_chunk),
           "modup ntt");
    return;
  }
  // jobs are ordered by tower: one launch per run of towers of one headroom class
  const uint32_t T = ks.L + ks.K;
  for (uint32_t t0 = 0; t0 < T;) {
    const int mode = c.ntt_mode_for(t0, 1);
    uint32_t t1 = t0 + 1;
    while (t1 < T && c.ntt_mode_for(t1, 1) == mode) ++t1;
    const uint32_t k0 = ks.tower_job_begin[t0], k1 = ks.tower_job_begin[t1];
    hip_ok(launch_ntt_span(m, c.dprimes(), c.log_n, (uint32_t)B, k0 * (uint32_t)B, (k1 - k0) * (uint32_t)B, false, mode,
                           s),
           "modup ntt");
    t0 = t1;
  }
}

void ks_modup_convert(const ntt_hip_keyswitch &ks, KsIn in, size_t B, const KsBuffers &w, hipStream_t s) {
  const ntt_hip_context &c = *ks.ctx;
  const size_t L = ks.L, T = ks.L + ks.K, n = c.n;
  const PrimeConst *pr = c.dprimes();
  TowerMap im = affine(w.coeff, in.base, n, L, 0, ks.coeff_words(), L * n);
  im.src_table = in.table;
  hip_ok(launch_ntt(im, pr, c.log_n, (uint32_t)(B * L), true, c.ntt_mode(), s, c.ntt_chunk), "intt");
  for (uint32_t d = 0; d < ks.beta; ++d) {
    const auto &plan = *ks.up[d];
    const uint32_t first = d * ks.alpha, count = std::min<uint32_t>(ks.alpha, ks.L - first);
    BasisLaunch a{};
    a.plan = plan.dev;
    a.in = w.coeff + (size_t)first * n;
    a.out = w.digits + (size_t)d * T * n;
    a.in_poly_stride = ks.coeff_words();
    a.out_poly_stride = ks.digit_words();
    a.n = (uint32_t)n;
    a.targets_per_block = c.conv_targets;
    a.plan_host_s = plan.s;
    a.plan_host_t = plan.t;
    a.skip_at = first;
    a.skip_len = count;
    hip_ok(launch_basis_switch(a, plan.per_term, (uint32_t)B, s), "modup conversion");
  }
}

// Key inner product over all digits of `plan`'s workspace with `key`'s
// resident key, both channels, into z ([2][B] * z_words).
void ks_mac(const ntt_hip_keyswitch &key, const ntt_hip_keyswitch &plan, KsIn in, size_t B, const KsBuffers &w,
            uint64_t *z, hipStream_t s, uint32_t first_tower = 0, uint32_t towers = 0) {
  const size_t T = plan.L + plan.K, n = plan.n();
  MacLaunch m{};
  m.own = in.base;
  m.own_table = in.table;
  m.own_poly_stride = plan.L * n;
  m.alpha = plan.alpha;
  m.active_q = plan.L;
  m.digits = w.digits;
  m.key_b = (const uint64_t *)key.key->ptr;
  m.key_a = m.key_b + (size_t)plan.beta * T * n;
  m.out_b = z;
  m.out_a = z + B * plan.z_words();
  m.primes = plan.ctx->dprimes();
  m.tower_prime = (const uint32_t *)plan.tower_prime->ptr;
  m.digit_poly_stride = plan.digit_words();
  m.out_poly_stride = plan.z_words();
  m.n = (uint32_t)n;
  m.towers = (uint32_t)T;
  m.digits_count = plan.beta;
  m.first_tower = first_tower;
  m.launch_towers = towers;
  hip_ok(launch_mac_fused(m, key.mac_per_term, (uint32_t)B, s), "mac");
}

// The fused ModUp row pass + MAC covers N = 2^16 contexts in the wide-lazy
// arithmetic with the 8-residue row pass, and sums that fit its accumulators.
bool fused_mac_applies(const ntt_hip_keyswitch &plan) {
  const ntt_hip_context &c = *plan.ctx;
  const int mode = c.ntt_mode();
  if (plan.fuse_mac == 0 || c.log_n != 16 || (mode & kArithMask) != kLazy || (mode & kRow16Bit)) return false;
  return plan.fuse_mac == 1 ? !plan.mac_per_term : plan.mac_sum_log >= 0;
}

// ModUp with its row pass fused into the key inner product of `key` into z:
// per group of towers, the column pass of their complement transforms, then
// one launch that finishes them in registers and accumulates both channels.
void ks_modup_mac_fused(const ntt_hip_keyswitch &key, const ntt_hip_keyswitch &plan, KsIn in, size_t B,
                        const KsBuffers &w, uint64_t *z, uint32_t G, hipStream_t s) {
  const ntt_hip_context &c = *plan.ctx;
  const uint32_t T = plan.L + plan.K;
  const size_t n = c.n;
  ks_modup_convert(plan, in, B, w, s);
  const TowerMap cols = complement_jobs(plan, w);
  ModUpMacRows mm{};
  mm.digits = w.digits;
  mm.own = in.base;
  mm.own_table = in.table;
  mm.own_poly_stride = plan.L * n;
  mm.key_b = (const uint64_t *)key.key->ptr;
  mm.key_a = mm.key_b + (size_t)plan.beta * T * n;
  mm.out_b = z;
  mm.out_a = z + B * plan.z_words();
  mm.tower_prime = (const uint32_t *)plan.tower_prime->ptr;
  mm.digit_poly_stride = plan.digit_words();
  mm.out_poly_stride = plan.z_words();
  mm.towers = T;
  mm.alpha = plan.alpha;
  mm.active_q = plan.L;
  mm.digits_count = plan.beta;
  mm.polys = (uint32_t)B;
  mm.wide = plan.fuse_mac == 1;
  mm.sum_log = plan.mac_sum_log < 0 ? 0 : (uint32_t)plan.mac_sum_log;
  for (uint32_t g0 = 0; g0 < T; g0 += G) {
    const uint32_t g1 = std::min(T, g0 + G);
    // runs of towers of one headroom class: both passes of a tower take its class
    for (uint32_t t0 = g0; t0 < g1;) {
      const int mode = c.ntt_mode_for(t0, 1);
      uint32_t t1 = t0 + 1;
      while (t1 < g1 && c.ntt_mode_for(t1, 1) == mode) ++t1;
      const uint32_t k0 = plan.tower_job_begin[t0], k1 = plan.tower_job_begin[t1];
      hip_ok(launch_ntt_columns_span(cols, c.dprimes(), (uint32_t)B, k0 * (uint32_t)B, (k1 - k0) * (uint32_t)B, mode,
                                     s),
             "modup column pass");
      mm.first_tower = t0;
      mm.launch_towers = t1 - t0;
      hip_ok(launch_modup_mac_rows(mm, c.dprimes(), mode, s), "modup row pass + mac");
      t0 = t1;
    }
  }
}

// ModUp and the key inner product of `key` into z: the fused row pass where
// it applies, otherwise interleaved by tower groups (mac_group) when the
// plan asks for it.
constexpr size_t kMinGroupBatch = 8;
void ks_modup_mac(const ntt_hip_keyswitch &key, const ntt_hip_keyswitch &plan, KsIn in, size_t B,
                  const KsBuffers &w, uint64_t *z, hipStream_t s) {
  const ntt_hip_context &c = *plan.ctx;
  const uint32_t T = plan.L + plan.K, G = plan.mac_group;
  if (fused_mac_applies(plan)) {
    ks_modup_mac_fused(key, plan, in, B, w, z, G == 0 || G >= T || B < kMinGroupBatch ? T : G, s);
    return;
  }
  if (G == 0 || G >= T || B < kMinGroupBatch) {  // small batches: fewer, larger launches
    ks_modup(plan, in, B, w, s);
    ks_mac(key, plan, in, B, w, z, s);
    return;
  }
  ks_modup_convert(plan, in, B, w, s);
  const TowerMap m = complement_jobs(plan, w);
  for (uint32_t t0 = 0; t0 < T; t0 += G) {
    const uint32_t t1 = std::min(T, t0 + G);
    const uint32_t k0 = plan.tower_job_begin[t0], k1 = plan.tower_job_begin[t1];
    hip_ok(launch_ntt_span(m, c.dprimes(), c.log_n, (uint32_t)B, k0 * (uint32_t)B, (k1 - k0) * (uint32_t)B, false,
                           c.ntt_mode_for(t0, t1 - t0), s),
           "modup ntt group");
    ks_mac(key, plan, in, B, w, z, s, t0, t1 - t0);
  }
}

// ModDown of both channels of z ([2][B] * z_words, evaluation QP): INTT(P) ->
// CRT to Q -> NTT with (Z_Q - NTT(corr)) P^-1 fused into its last pass;
// channel b lands in out_b, channel a in out_a (B * L * n each).
// accumulate: bit 0 adds the B channel into out_b, bit 1 the A channel into
// out_a (mod q) instead of writing them.
// out_table (when set): per-component output addresses, the B channel of
// component c at out_table[c], the A channel at out_table[B + c].
void ks_moddown(const ntt_hip_keyswitch &ks, const uint64_t *z, size_t B, uint64_t *out_b, uint64_t *out_a,
                const KsBuffers &w, hipStream_t s, uint32_t accumulate = 0, uint64_t *const *out_table = nullptr) {
  const ntt_hip_context &c = *ks.ctx;
  const size_t L = ks.L, K = ks.K, n = c.n;
  const PrimeConst *pr = c.dprimes();
  hip_ok(launch_ntt(affine(w.pc, z + L * n, n, K, L, ks.pc_words(), ks.z_words()), pr, c.log_n,
                    (uint32_t)(2 * B * K), true, c.ntt_mode(), s, c.ntt_chunk),
         "moddown intt");
  BasisLaunch a{};
  a.plan = ks.down->dev;
  a.in = w.pc;
  a.out = w.corr;
  a.in_poly_stride = ks.pc_words();
  a.out_poly_stride = ks.corr_words();
  a.n = (uint32_t)n;
  a.targets_per_block = c.conv_targets;
  a.plan_host_s = ks.down->s;
  a.plan_host_t = ks.down->t;
  a.skip_at = 0xffffffffu;
  a.skip_len = 0;
  hip_ok(launch_basis_switch(a, ks.down->per_term, (uint32_t)(2 * B), s), "moddown conversion");
  TowerMap fin = with_correction(affine(w.corr, nullptr, n, L, 0, ks.corr_words(), 0), z, ks.z_words(), out_b, out_a,
                                 (uint32_t)B, L * n, ks.pinv);
  fin.epi.accumulate = accumulate;
  fin.epi.out_table = out_table;
  if (c.ntt_chunk) {
    hip_ok(launch_ntt(fin, pr, c.log_n, (uint32_t)(2 * B * L), false, c.ntt_mode(), s, c.ntt_chunk),
           "moddown ntt + correction");
    return;
  }
  // one launch per run of towers of one headroom class (tower-major index:
  // towers [t0, t1) of all 2B polynomials are indices [t0 2B, t1 2B))
  const uint32_t polys = (uint32_t)(2 * B);
  for (size_t t0 = 0; t0 < L;) {
    const int mode = c.ntt_mode_for(t0, 1);
    size_t t1 = t0 + 1;
    while (t1 < L && c.ntt_mode_for(t1, 1) == mode) ++t1;
    hip_ok(launch_ntt_span(fin, pr, c.log_n, polys, (uint32_t)(t0 * polys), (uint32_t)((t1 - t0) * polys), false, mode,
                           s),
           "moddown ntt + correction");
    t0 = t1;
  }
}
// Lanes of a batch: up to ks.lanes sub-batches of at least kMinLaneBatch.
uint32_t lane_count(const ntt_hip_keyswitch &ks, size_t batch) {
  return (uint32_t)std::max<size_t>(1, std::min<size_t>(ks.lanes, batch / ntt_hip_keyswitch::kMinLaneBatch));
}

// Run fn(first, count, lane, stream) per lane: lane j takes components
// [first_j, first_j + count_j); lane 0 runs on the caller's stream, the others
// on side streams forked from it and joined back, so the caller's stream (or
// its capture) orders the whole batch.
template <class F>
void run_lanes(ntt_hip_keyswitch &ks, size_t batch, hipStream_t s, F &&fn) {
  const uint32_t lanes = lane_count(ks, batch);
  if (lanes <= 1) {
    fn(0, batch, 0u, s);
    return;
  }
  ks.ensure_lanes(lanes);
  HIP_CHECK(hipEventRecord(ks.fork, s));
  size_t first = 0;
  for (uint32_t j = 0; j < lanes; ++j) {
    const size_t count = batch / lanes + (j < batch % lanes ? 1 : 0);
    const hipStream_t st = j == 0 ? s : ks.side[j - 1];
    if (j) HIP_CHECK(hipStreamWaitEvent(st, ks.fork, 0));
    fn(first, count, j, st);
    first += count;
  }
  for (uint32_t j = 1; j < lanes; ++j) {
    HIP_CHECK(hipEventRecord(ks.joins[j - 1], ks.side[j - 1]));
    HIP_CHECK(hipStreamWaitEvent(s, ks.joins[j - 1], 0));
  }
}

// Pointer tables of ntt_hip_keyswitch_run_gather (3 per component), 256-byte padded.
size_t gather_table_bytes(size_t batch) { return (3 * batch * sizeof(void *) + 255) / 256 * 256; }
}  // namespace

extern "C" {

int ntt_hip_keyswitch_workspace_bytes(ntt_hip_keyswitch *ks, size_t batch, size_t *bytes) {
  return guarded([&] {
    require(ks && bytes, "null argument");
    require(batch >= 1 && batch <= ks->max_batch, "batch exceeds the plan capacity");
    *bytes = batch * ks->per_component() * sizeof(uint64_t);
    return NTT_HIP_OK;
  });
}

int ntt_hip_keyswitch_run(ntt_hip_keyswitch *ks, const uint64_t *in, size_t batch, uint64_t *out_b,
                          uint64_t *out_a, void *workspace, void *stream) {
  return guarded([&] {
    require(ks && in && out_b && out_a && workspace, "null argument");
    require(ks->key_loaded, "evaluation key not uploaded");
    require(batch >= 1 && batch <= ks->max_batch, "batch exceeds the plan capacity");
    const hipStream_t s = ks->ctx->pick(stream);
    DeviceGuard g(ks->ctx->device);
    run_lanes(*ks, batch, s, [&](size_t first, size_t count, uint32_t, hipStream_t st) {
      const size_t off = first * ks->L * ks->n();
      const KsBuffers w = ks_buffers(*ks, (uint64_t *)workspace + first * ks->per_component(), count);
      ks_modup_mac(*ks, *ks, in + off, count, w, w.z, st);
      ks_moddown(*ks, w.z, count, out_b + off, out_a + off, w, st);
    });
    return NTT_HIP_OK;
  });
}

int ntt_hip_keyswitch_gather_workspace_bytes(ntt_hip_keyswitch *ks, size_t batch, size_t *bytes) {
  return guarded([&] {
    require(ks && bytes, "null argument");
    require(batch >= 1 && batch <= ks->max_batch, "batch exceeds the plan capacity");
    *bytes = batch * ks->per_component() * sizeof(uint64_t) + gather_table_bytes(batch);
    return NTT_HIP_OK;
  });
}

int ntt_hip_keyswitch_run_gather(ntt_hip_keyswitch *ks, const uint64_t *const *in, size_t batch,
                                 uint64_t *const *out_b, uint64_t *const *out_a, void *workspace, void *stream) {
  return guarded([&] {
    require(ks && in && out_b && out_a && workspace, "null argument");
    require(ks->key_loaded, "evaluation key not uploaded");
    require(batch >= 1 && batch <= ks->max_batch, "batch exceeds the plan capacity");
    for (size_t c = 0; c < batch; ++c) require(in[c] && out_b[c] && out_a[c], "null component pointer");
    const hipStream_t s = ks->ctx->pick(stream);
    DeviceGuard g(ks->ctx->device);
    // the tables behind the workspace proper: the inputs, then per lane its
    // components' B outputs followed by their A outputs (the order of the
    // ModDown epilogue's polynomials)
    const uint32_t lanes = lane_count(*ks, batch);
    std::vector<const void *> table(3 * batch);
    for (size_t c = 0; c < batch; ++c) table[c] = in[c];
    size_t first = 0;
    for (uint32_t j = 0; j < lanes; ++j) {
      const size_t count = batch / lanes + (j < batch % lanes ? 1 : 0);
      for (size_t c = 0; c < count; ++c) {
        table[batch + 2 * first + c] = out_b[first + c];
        table[batch + 2 * first + count + c] = out_a[first + c];
      }
      first += count;
    }
    char *dev_table = (char *)workspace + batch * ks->per_component() * sizeof(uint64_t);
    HIP_CHECK(hipMemcpyAsync(dev_table, table.data(), table.size() * sizeof(void *), hipMemcpyHostToDevice, s));
    const uint64_t *const *in_tab = (const uint64_t *const *)dev_table;
    uint64_t *const *out_tab = (uint64_t *const *)(dev_table + batch * sizeof(void *));
    run_lanes(*ks, batch, s, [&](size_t first_c, size_t count, uint32_t, hipStream_t st) {
      const KsBuffers w = ks_buffers(*ks, (uint64_t *)workspace + first_c * ks->per_component(), count);
      ks_modup_mac(*ks, *ks, KsIn(in_tab + first_c, 0), count, w, w.z, st);
      ks_moddown(*ks, w.z, count, nullptr, nullptr, w, st, 0, out_tab + 2 * first_c);
    });
    return NTT_HIP_OK;
  });
}

}  // extern "C"

// ---------------------------------------------------------------- allocation-free key switch
// The workspace of ntt_hip_keyswitch_run_static lives in this file-scope
// device array: every GPU the library runs on holds its own copy until the
// process ends. Host code never uses the array's name as an address (on the
// host it is only a stand-in for the symbol): the device address comes from
// hipGetSymbolAddress, once per device, with that device current.
namespace {
constexpr size_t kStaticScratchBytes = size_t{136} << 20;  // one component at N = 2^16, L = 24, K = 6, dnum = 4
constexpr int kMaxStaticDevices = 64;
__device__ __align__(256) uint64_t scratch[kStaticScratchBytes / sizeof(uint64_t)];
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxStaticDevices] = {};

// Device address of `scratch` on `device` (current): looked up on the first
// call per device and kept.
void *static_scratch(int device) {
  require(device >= 0 && device < kMaxStaticDevices, "device ordinal beyond the static scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void *ptr = nullptr;
    HIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    require(ptr != nullptr, "static scratch has no device address");
    g_scratch_ptr[device] = ptr;
  }
  return g_scratch_ptr[device];
}
}  // namespace

extern "C" {

int ntt_hip_keyswitch_run_static(ntt_hip_keyswitch *ks, const uint64_t *in, size_t batch, uint64_t *out_b,
                                 uint64_t *out_a, void *stream) {
  return guarded([&] {
    require(ks && in && out_b && out_a, "null argument");
    require(ks->key_loaded, "evaluation key not uploaded");
    require(batch >= 1 && batch <= ks->max_batch, "batch exceeds the plan capacity");
    const size_t need = batch * ks->per_component() * sizeof(uint64_t);
    if (need > kStaticScratchBytes)
      throw ContractError("the batch needs " + std::to_string(need) + " bytes of key-switch workspace; the static "
                          "scratch holds " + std::to_string(kStaticScratchBytes) +
                          " (use ntt_hip_keyswitch_run with a caller workspace)");
    const hipStream_t s = ks->ctx->pick(stream);
    DeviceGuard g(ks->ctx->device);
    void *workspace = static_scratch(ks->ctx->device);
    // one lane on the caller's stream: no side streams or events are created
    const KsBuffers w = ks_buffers(*ks, workspace, batch);
    ks_modup_mac(*ks, *ks, in, batch, w, w.z, s);
    ks_moddown(*ks, w.z, batch, out_b, out_a, w, s);
    return NTT_HIP_OK;
  });
}

int ntt_hip_keyswitch_in_place(ntt_hip_keyswitch *ks, uint64_t *c0, uint64_t *c1, const uint64_t *c2, size_t batch,
                               void *workspace, void *stream) {
  return guarded([&] {
    require(ks && c0 && c1 && workspace, "null argument");
    require(c0 != c1 && (!c2 || (c2 != c0 && c2 != c1)), "ciphertext components must not alias");
    require(ks->key_loaded, "evaluation key not uploaded");
    require(batch >= 1 && batch <= ks->max_batch, "batch exceeds the plan capacity");
    const hipStream_t s = ks->ctx->pick(stream);
    const KsBuffers w = ks_buffers(*ks, workspace, batch);
    const uint64_t *in = c2 ? c2 : c1;
    DeviceGuard g(ks->ctx->device);
    ks_modup_mac(*ks, *ks, in, batch, w, w.z, s);
    // c1 is read only by the ModUp and the MAC above, so the two-component
    // case may overwrite it in the ModDown's last pass (stream order)
    ks_moddown(*ks, w.z, batch, c0, c1, w, s, c2 ? 3u : 1u);
    return NTT_HIP_OK;
  });
}

int ntt_hip_relinearize(ntt_hip_keyswitch *const *keys, size_t extra, uint64_t *c0, uint64_t *c1,
                        const uint64_t *const *higher, size_t batch, void *workspace, void *stream) {
  return guarded([&] {
    require(c0 && c1 && workspace && (extra == 0 || (keys && higher)), "null argument");
    require(c0 != c1, "ciphertext components must not alias");
    if (extra == 0) return NTT_HIP_OK;
    require(keys[0] != nullptr, "null relinearization key");
    const ntt_hip_keyswitch &plan = *keys[0];
    require(batch >= 1 && batch <= plan.max_batch, "batch exceeds the plan capacity");
    for (size_t j = 0; j < extra; ++j) {
      const ntt_hip_keyswitch *k = keys[j];
      require(k && k->ctx == plan.ctx && k->L == plan.L && k->K == plan.K && k->alpha == plan.alpha,
              "relinearization keys must share the context and the (L, K, alpha) plan");
      require(k->key_loaded, "relinearization key not uploaded");
      require(higher[j], "null higher component");
      require(higher[j] != c0 && higher[j] != c1, "ciphertext components must not alias");
    }
    const hipStream_t s = plan.ctx->pick(stream);
    const KsBuffers w = ks_buffers(plan, workspace, batch);
    DeviceGuard g(plan.ctx->device);
    for (size_t j = 0; j < extra; ++j) {
      ks_modup_mac(*keys[j], plan, higher[j], batch, w, w.z, s);
      ks_moddown(plan, w.z, batch, c0, c1, w, s, 3u);
    }
    return NTT_HIP_OK;
  });
}

int ntt_hip_rotate_workspace_bytes(ntt_hip_keyswitch *ks, size_t batch, size_t *bytes) {
  return guarded([&] {
    require(ks && bytes, "null argument");
    require(batch >= 1 && batch <= ks->max_batch, "batch exceeds the plan capacity");
    *bytes = batch * (ks->per_component() + 2 * ks->z_words()) * sizeof(uint64_t);
    return NTT_HIP_OK;
  });
}

int ntt_hip_keyswitch_rotate(ntt_hip_keyswitch *const *keys, size_t rotations, const uint32_t *automorphisms,
                             const uint64_t *c0, const uint64_t *c1, size_t batch, uint64_t *const *out_b,
                             uint64_t *const *out_a, void *workspace, void *stream) {
  return guarded([&] {
    require(keys && automorphisms && c0 && c1 && out_b && out_a && workspace && rotations >= 1, "null argument");
    require(keys[0] != nullptr, "null rotation key");
    require(c0 != c1, "ciphertext components must not alias");
    ntt_hip_keyswitch &plan = *keys[0];
    require(batch >= 1 && batch <= plan.max_batch, "batch exceeds the plan capacity");
    for (size_t r = 0; r < rotations; ++r) {
      const ntt_hip_keyswitch *k = keys[r];
      require(k && k->ctx == plan.ctx && k->L == plan.L && k->K == plan.K && k->alpha == plan.alpha,
              "rotation keys must share the context and the (L, K, alpha) plan");
      require(k->key_loaded, "rotation key not uploaded");
      require(out_b[r] && out_a[r], "null rotation output");
      // every rotation re-reads c0 and c1: an output over either would corrupt the later ones
      require(out_b[r] != out_a[r] && out_b[r] != c0 && out_b[r] != c1 && out_a[r] != c0 && out_a[r] != c1,
              "rotation outputs must not alias each other or the ciphertext");
      require((automorphisms[r] & 1) && automorphisms[r] < 2 * plan.ctx->n, "invalid negacyclic automorphism index");
    }
    ntt_hip_context &c = *plan.ctx;
    const hipStream_t s = c.pick(stream);
    const KsBuffers w = ks_buffers(plan, workspace, batch);
    uint64_t *rotated = (uint64_t *)workspace + batch * plan.per_component();  // [2][B] * z_words
    const size_t L = plan.L, T = plan.L + plan.K, n = c.n;
    DeviceGuard g(c.device);
    ks_modup(plan, c1, batch, w, s);  // digits once for every rotation
    for (size_t r = 0; r < rotations; ++r) {
      ks_mac(*keys[r], plan, c1, batch, w, w.z, s);
      RotateLaunch rl{};
      rl.out_b = rotated;
      rl.out_a = rotated + batch * plan.z_words();
      rl.ext_b = w.z;
      rl.ext_a = w.z + batch * plan.z_words();
      rl.c0 = c0;
      rl.primes = c.dprimes();
      rl.scale = plan.pmod;
      rl.first_prime = 0;
      rl.n = (uint32_t)n;
      rl.log_n = c.log_n;
      rl.active_q = (uint32_t)L;
      rl.automorphism = automorphisms[r];
      rl.ext_poly_stride = rl.out_poly_stride = plan.z_words();
      rl.c0_poly_stride = L * n;
      hip_ok(launch_rotate(rl, (uint32_t)T, (uint32_t)batch, s), "rotation");
      ks_moddown(plan, rotated, batch, out_b[r], out_a[r], w, s);
    }
    return NTT_HIP_OK;
  });
}

}  // extern "C"

// ---------------------------------------------------------------- graphs
namespace {
struct GraphHandle {
  int device;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};
}  // namespace

extern "C" {

int ntt_hip_keyswitch_capture(ntt_hip_keyswitch *ks, const uint64_t *in, size_t batch, uint64_t *out_b,
                              uint64_t *out_a, void *workspace, void *stream, void **graph) {
  return guarded([&] {
    require(ks && graph && stream, "capture needs a key switch, an explicit stream and a graph out");
    DeviceGuard g(ks->ctx->device);
    auto h = std::make_unique<GraphHandle>();
    h->device = ks->ctx->device;
    const hipStream_t s = (hipStream_t)stream;
    HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    const int status = ntt_hip_keyswitch_run(ks, in, batch, out_b, out_a, workspace, stream);
    hipGraph_t captured = nullptr;
    const hipError_t end = hipStreamEndCapture(s, &captured);
    if (status != NTT_HIP_OK) {
      if (captured) (void)hipGraphDestroy(captured);
      const std::string why = g_last_error;
      throw RuntimeError("capture failed: " + why);
    }
    HIP_CHECK(end);
    h->graph = captured;
    HIP_CHECK(hipGraphInstantiate(&h->exec, h->graph, nullptr, nullptr, 0));
    *graph = h.release();
    return NTT_HIP_OK;
  });
}

int ntt_hip_graph_launch(void *graph, void *stream) {
  return guarded([&] {
    require(graph != nullptr, "null graph");
    auto *h = (GraphHandle *)graph;
    DeviceGuard g(h->device);
    HIP_CHECK(hipGraphLaunch(h->exec, (hipStream_t)stream));
    return NTT_HIP_OK;
  });
}

int ntt_hip_graph_destroy(void *graph) {
  return guarded([&] {
    auto *h = (GraphHandle *)graph;
    if (!h) return NTT_HIP_OK;
    DeviceGuard g(h->device);
    if (h->exec) (void)hipGraphExecDestroy(h->exec);
    if (h->graph) (void)hipGraphDestroy(h->graph);
    delete h;
    return NTT_HIP_OK;
  });
}

}  // extern "C"

// ---------------------------------------------------------------- batched ModUp / ModDown
namespace {

void launch_basis(const BasisPlanDev &plan, uint32_t targets, const uint64_t *in, uint64_t *out, uint64_t in_stride,
                  uint64_t out_stride, uint32_t n, uint32_t polys, hipStream_t s) {
  BasisLaunch a{};
  a.plan = plan.dev;
  a.in = in;
  a.out = out;
  a.in_poly_stride = in_stride;
  a.out_poly_stride = out_stride;
  a.n = n;
  a.targets_per_block = targets;
  a.plan_host_s = plan.s;
  a.plan_host_t = plan.t;
  a.skip_at = 0xffffffffu;
  a.skip_len = 0;
  hip_ok(launch_basis_switch(a, plan.per_term, polys, s), "conversion");
}

}  // namespace

extern "C" {

int ntt_hip_modup_workspace_bytes(ntt_hip_keyswitch *ks, size_t batch, size_t *bytes) {
  return guarded([&] {
    require(ks && bytes && batch >= 1, "invalid argument");
    *bytes = batch * (size_t)ks->L * ks->n() * sizeof(uint64_t);
    return NTT_HIP_OK;
  });
}

int ntt_hip_modup_run(ntt_hip_keyswitch *ks, const uint64_t *in, int in_eval, size_t batch, uint64_t *out,
                      void *workspace, void *stream) {
  return guarded([&] {
    require(ks && in && out && workspace && batch >= 1, "null argument");
    ntt_hip_context &c = *ks->ctx;
    const size_t L = ks->L, K = ks->K, T = L + K, n = c.n;
    require(batch * T <= 65535, "batch too large for one launch");
    const hipStream_t s = c.pick(stream);
    const PrimeConst *pr = c.dprimes();
    uint64_t *coeff = (uint64_t *)workspace;  // batch * L * n
    DeviceGuard g(c.device);
    if (in_eval) {
      // Q part of the output is the input itself; coefficients for the lift
      hip_ok(launch_ntt(affine(coeff, in, n, L, 0, L * n, L * n), pr, c.log_n, (uint32_t)(batch * L), true, c.ntt_mode(), s, c.ntt_chunk),
             "modup intt");
      for (size_t b = 0; b < batch; ++b)
        HIP_CHECK(hipMemcpyAsync(out + b * T * n, in + b * L * n, L * n * sizeof(uint64_t), hipMemcpyDeviceToDevice, s));
      launch_basis(*ks->full_up, c.conv_targets, coeff, out + L * n, L * n, T * n, (uint32_t)n, (uint32_t)batch, s);
    } else {
      hip_ok(launch_ntt(affine(out, in, n, L, 0, T * n, L * n), pr, c.log_n, (uint32_t)(batch * L), false, c.ntt_mode(), s, c.ntt_chunk),
             "modup ntt q");
      launch_basis(*ks->full_up, c.conv_targets, in, out + L * n, L * n, T * n, (uint32_t)n, (uint32_t)batch, s);
    }
    hip_ok(launch_ntt(affine(out + L * n, nullptr, n, K, L, T * n, 0), pr, c.log_n, (uint32_t)(batch * K), false,
                      c.ntt_mode(), s, c.ntt_chunk),
           "modup ntt p");
    return NTT_HIP_OK;
  });
}

int ntt_hip_moddown_workspace_bytes(ntt_hip_keyswitch *ks, size_t batch, size_t *bytes) {
  return guarded([&] {
    require(ks && bytes && batch >= 1, "invalid argument");
    *bytes = batch * (size_t)(ks->K + ks->L) * ks->n() * sizeof(uint64_t);
    return NTT_HIP_OK;
  });
}

int ntt_hip_moddown_run(ntt_hip_keyswitch *ks, const uint64_t *in, size_t batch, uint64_t *out, void *workspace,
                        void *stream) {
  return guarded([&] {
    require(ks && in && out && workspace && batch >= 1, "null argument");
    ntt_hip_context &c = *ks->ctx;
    const size_t L = ks->L, K = ks->K, T = L + K, n = c.n;
    require(batch * T <= 65535, "batch too large for one launch");
    const hipStream_t s = c.pick(stream);
    const PrimeConst *pr = c.dprimes();
    uint64_t *pc = (uint64_t *)workspace;  // batch * K * n
    uint64_t *corr = pc + batch * K * n;   // batch * L * n
    DeviceGuard g(c.device);
    hip_ok(launch_ntt(affine(pc, in + L * n, n, K, L, K * n, T * n), pr, c.log_n, (uint32_t)(batch * K), true, c.ntt_mode(),
                      s, c.ntt_chunk),
           "moddown intt");
    launch_basis(*ks->down, c.conv_targets, pc, corr, K * n, L * n, (uint32_t)n, (uint32_t)batch, s);
    hip_ok(launch_ntt(with_correction(affine(corr, nullptr, n, L, 0, L * n, 0), in, T * n, out, out, (uint32_t)batch,
                                      L * n, ks->pinv),
                      pr, c.log_n, (uint32_t)(batch * L), false, c.ntt_mode(), s, c.ntt_chunk),
           "moddown ntt + correction");
    return NTT_HIP_OK;
  });
}

}  // extern "C"

extern "C" int ntt_hip_copy_words(int device, uint64_t *dst, const uint64_t *src, size_t words, void *stream) {
  return guarded([&] {
    require(dst && src, "null copy pointer");
    DeviceGuard g(device);
    hip_ok(launch_copy_words(dst, src, words, (hipStream_t)stream), "copy kernel");
    return NTT_HIP_OK;
  });
}

// ---------------------------------------------------------------- autotuner
// Times candidate launch geometries on scratch data of this context and keeps
// the fastest: the NTT tower chunk (how many towers one pass-A/pass-B pair
// covers before the next group starts, trading launch count for L2/Infinity
// Cache reuse of the inter-pass hand-off) and the basis-conversion targets per
// workgroup (parallelism vs. recomputing the scaled sources).
namespace {
float time_ms(ntt_hip_context &c, const std::function<void()> &body, int reps) {
  hipEvent_t e0, e1;
  HIP_CHECK(hipEventCreate(&e0));
  HIP_CHECK(hipEventCreate(&e1));
  body();  // warm
  HIP_CHECK(hipEventRecord(e0, c.stream));
  for (int r = 0; r < reps; ++r) body();
  HIP_CHECK(hipEventRecord(e1, c.stream));
  HIP_CHECK(hipEventSynchronize(e1));
  float ms = 0;
  HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return ms / reps;
}
}  // namespace

extern "C" {

int ntt_hip_set_tuning(ntt_hip_context *ctx, uint32_t ntt_chunk, uint32_t conv_targets) {
  return guarded([&] {
    require(ctx != nullptr, "null context");
    ctx->ntt_chunk = ntt_chunk;
    ctx->conv_targets = conv_targets;
    return NTT_HIP_OK;
  });
}

int ntt_hip_get_tuning(ntt_hip_context *ctx, uint32_t *ntt_chunk, uint32_t *conv_targets) {
  return guarded([&] {
    require(ctx && ntt_chunk && conv_targets, "null argument");
    *ntt_chunk = ctx->ntt_chunk;
    *conv_targets = ctx->conv_targets;
    return NTT_HIP_OK;
  });
}

int ntt_hip_set_ntt_shape(ntt_hip_context *ctx, int fwd_headroom, uint32_t row_residues, uint32_t wg_per_cu) {
  return guarded([&] {
    require(ctx != nullptr, "null context");
    require(fwd_headroom >= -1 && fwd_headroom <= kMaxHeadroom, "forward headroom must be -1 (from the moduli) or 0..5");
    require(row_residues == 8 || row_residues == 16, "row-pass residues per thread must be 8 or 16");
    require(wg_per_cu <= 15, "workgroups per CU must be 0 (no cap) .. 15");
    ctx->fwd_headroom = fwd_headroom;
    ctx->row16 = row_residues == 16;
    ctx->wg_cap = wg_per_cu;
    return NTT_HIP_OK;
  });
}

int ntt_hip_get_ntt_shape(ntt_hip_context *ctx, int *fwd_headroom, uint32_t *row_residues, uint32_t *wg_per_cu) {
  return guarded([&] {
    require(ctx && fwd_headroom && row_residues && wg_per_cu, "null argument");
    *fwd_headroom = ctx->fwd_headroom;
    *row_residues = ctx->row16 ? 16 : 8;
    *wg_per_cu = ctx->wg_cap;
    return NTT_HIP_OK;
  });
}

int ntt_hip_keyswitch_set_shape(ntt_hip_keyswitch *ks, uint32_t lanes, uint32_t mac_group, uint32_t fuse_mac) {
  return guarded([&] {
    require(ks != nullptr, "null key switch plan");
    require(lanes >= 1 && lanes <= 8, "lanes must be 1..8");
    require(fuse_mac <= 2, "fuse_mac must be 0 (off), 1 (128-bit sums) or 2 (64-bit sums)");
    ks->lanes = lanes;
    ks->mac_group = mac_group;
    ks->fuse_mac = fuse_mac;
    return NTT_HIP_OK;
  });
}

int ntt_hip_keyswitch_get_shape(ntt_hip_keyswitch *ks, uint32_t *lanes, uint32_t *mac_group, uint32_t *fuse_mac,
                                int *fused) {
  return guarded([&] {
    require(ks && lanes && mac_group && fuse_mac && fused, "null argument");
    *lanes = ks->lanes;
    *mac_group = ks->mac_group;
    *fuse_mac = ks->fuse_mac;
    *fused = fused_mac_applies(*ks) ? 1 : 0;
    return NTT_HIP_OK;
  });
}

// The key switch's own tuner: the same coordinate search with alternating
// pairs, timed on the batched key switch itself (seeded uniform residues: the
// clock under load depends on the data), over the context's NTT shape and
// chunk, the plan's lanes, ModUp/MAC fusion and interleave group, and the
// conversion width.
int ntt_hip_keyswitch_autotune(ntt_hip_keyswitch *ks, size_t batch, char *report, size_t capacity) {
  return guarded([&] {
    require(ks && ks->key_loaded, "autotune needs a plan with its key uploaded");
    require(batch >= 1 && batch <= ks->max_batch, "batch exceeds the plan capacity");
    ntt_hip_context &c = *ks->ctx;
    DeviceGuard g(c.device);
    const size_t comp = (size_t)ks->L * c.n;
    DeviceBlob io(c.device, 3 * batch * comp * sizeof(uint64_t));
    DeviceBlob ws(c.device, batch * ks->per_component() * sizeof(uint64_t));
    uint64_t *in = (uint64_t *)io.ptr, *ob = in + batch * comp, *oa = ob + batch * comp;
    // uniform residues below every modulus (canonical in every tower)
    const uint64_t min_q = *std::min_element(c.q.begin(), c.q.end());
    hip_ok(launch_fill_uniform(in, batch * comp, 0x7251d9, min_q, c.stream), "tune operands");
    auto run_ms = [&] {
      return time_ms(c, [&] {
        const int st = ntt_hip_keyswitch_run(ks, in, batch, ob, oa, ws.ptr, c.stream);
        if (st != NTT_HIP_OK) throw RuntimeError(std::string("tune key switch: ") + g_last_error);
      }, 2);
    };
    for (int w = 0; w < 3; ++w) (void)run_ms();
    std::string rep = "{\"keyswitch\":{\"default_ms\":" + std::to_string(run_ms());
    auto sweep = [&](const char *name, size_t count, const std::function<void(size_t)> &set, size_t incumbent) {
      rep += ",\"" + std::string(name) + "\":{";
      size_t pick = incumbent;
      for (size_t i = 0; i < count; ++i) {
        if (i == pick) continue;
        float t_inc = 0, t_cand = 0;
        for (int r = 0; r < 3; ++r) {
          set(pick);
          t_inc += run_ms();
          set(i);
          t_cand += run_ms();
        }
        rep += (rep.back() == '{' ? "" : ",") + std::string("\"") + std::to_string(i) + "\":[" +
               std::to_string(t_cand / 3) + "," + std::to_string(t_inc / 3) + "]";
        if (t_cand < 0.98f * t_inc) pick = i;
      }
      set(pick);
      rep += "}";
    };
    if (c.log_n == 16 && (c.arith & kArithMask) != kExact) {
      const int heads[2] = {-1, 0};
      sweep("fwd_headroom", ((c.arith >> kArithShift) & 7) > 0 ? 2 : 1, [&](size_t i) { c.fwd_headroom = heads[i]; },
            0);
      sweep("row_residues", 2, [&](size_t i) { c.row16 = i == 1; }, 0);
      const uint32_t caps[4] = {0, 6, 5, 4};
      sweep("wg_per_cu", 4, [&](size_t i) { c.wg_cap = caps[i]; }, 0);
    }
    const uint32_t lanes[4] = {ntt_hip_keyswitch::kDefaultLanes, 1, 3, 4};
    sweep("lanes", 4, [&](size_t i) { ks->lanes = lanes[i]; }, 0);
    const uint32_t fuse[3] = {ntt_hip_keyswitch::kDefaultFuseMac, ntt_hip_keyswitch::kDefaultFuseMac == 0 ? 1u : 0u,
                              ntt_hip_keyswitch::kDefaultFuseMac == 2 ? 1u : 2u};
    sweep("fuse_mac", 3, [&](size_t i) { ks->fuse_mac = fuse[i]; }, 0);
    const uint32_t groups[5] = {ntt_hip_keyswitch::kDefaultMacGroup, ntt_hip_keyswitch::kDefaultMacGroup == 2 ? 0u : 2u,
                                1, 4, 8};
    sweep("mac_group", 5, [&](size_t i) { ks->mac_group = groups[i]; }, 0);
    const uint32_t targets[6] = {0, 2, 4, 8, 12, 24};
    sweep("conv_targets", 6, [&](size_t i) { c.conv_targets = targets[i]; }, 0);
    HIP_CHECK(hipStreamSynchronize(c.stream));
    rep += ",\"tuned_ms\":" + std::to_string(run_ms()) + "}";
    rep += ",\"chosen\":{\"fwd_headroom\":" + std::to_string(c.fwd_headroom) +
           ",\"row_residues\":" + std::to_string(c.row16 ? 16 : 8) + ",\"wg_per_cu\":" + std::to_string(c.wg_cap) +
           ",\"lanes\":" + std::to_string(ks->lanes) + ",\"fuse_mac\":" + std::to_string(ks->fuse_mac) +
           ",\"mac_group\":" + std::to_string(ks->mac_group) +
           ",\"conv_targets\":" + std::to_string(c.conv_targets) + "},\"batch\":" + std::to_string(batch) + "}";
    if (report && capacity) std::snprintf(report, capacity, "%s", rep.c_str());
    return NTT_HIP_OK;
  });
}

// Coordinate search: each knob in turn over its candidates (the others held
// at their best so far), a candidate replacing the incumbent only when it is
// clearly faster. NTT knobs are timed on `towers` forward + inverse transforms
// of context tower 0's prime, the conversion width on a first-half -> rest
// basis switch of the context.
int ntt_hip_autotune(ntt_hip_context *ctx, size_t towers, char *report, size_t capacity) {
  return guarded([&] {
    require(ctx && towers >= 1 && towers <= 65535, "autotune needs a context and 1..65535 towers");
    ntt_hip_context &c = *ctx;
    const size_t n = c.n;
    DeviceGuard g(c.device);
    // conversion sweep shape (first half of the basis -> the rest, <= 6 -> <= 24)
    const size_t cs = std::min<size_t>(6, c.q.size() / 2), ct = std::min<size_t>(24, c.q.size() - cs);
    const size_t cpolys = cs ? std::max<size_t>(1, std::min(towers / ct, towers / cs)) : 0;
    // operand a: the NTT's towers and the conversion's sources; b: its targets
    const size_t a_words = std::max(towers, cpolys * cs) * n, b_words = std::max<size_t>(1, cpolys * ct) * n;
    DeviceBlob scratch(c.device, (a_words + b_words) * sizeof(uint64_t));
    uint64_t *a = (uint64_t *)scratch.ptr, *b = a + a_words;
    const uint64_t min_q = *std::min_element(c.q.begin(), c.q.end());
    hip_ok(launch_fill_uniform(a, a_words + b_words, 0x7251d9, min_q, c.stream), "tune operands");
    const TowerMap m = affine(a, nullptr, n, 1, 0, n, 0);
    // one forward and one inverse transform as a standalone transform runs
    // them (the context's chunking)
    auto ntt_ms = [&] {
      return time_ms(c, [&] {
        hip_ok(launch_ntt(m, c.dprimes(), c.log_n, (uint32_t)towers, false, c.ntt_mode(), c.stream,
                          c.transform_chunk()),
               "tune ntt");
        hip_ok(launch_ntt(m, c.dprimes(), c.log_n, (uint32_t)towers, true, c.ntt_mode(), c.stream,
                          c.transform_chunk()),
               "tune ntt");
      }, 3);
    };
    for (int w = 0; w < 4; ++w) (void)ntt_ms();  // steady clocks before the first comparison
    std::string rep = "{\"ntt\":{";
    const float first = ntt_ms();
    rep += "\"default_ms\":" + std::to_string(first);
    // a knob: each candidate against the incumbent in alternating pairs (the
    // clock drifts less within a pair than across a sweep); the incumbent is
    // kept unless the candidate is at least 2% faster in the sum of 3 pairs
    auto sweep = [&](const char *name, size_t count, const std::function<void(size_t)> &set, size_t incumbent) {
      rep += ",\"" + std::string(name) + "\":{";
      size_t pick = incumbent;
      for (size_t i = 0; i < count; ++i) {
        if (i == pick) continue;
        float t_inc = 0, t_cand = 0;
        for (int r = 0; r < 3; ++r) {
          set(pick);
          t_inc += ntt_ms();
          set(i);
          t_cand += ntt_ms();
        }
        rep += (rep.back() == '{' ? "" : ",") + std::string("\"") + std::to_string(i) + "\":[" +
               std::to_string(t_cand / 3) + "," + std::to_string(t_inc / 3) + "]";
        if (t_cand < 0.98f * t_inc) pick = i;
      }
      set(pick);
      rep += "}";
    };
    const bool big = c.log_n == 16 && (c.arith & kArithMask) != kExact;  // the 2^16 two-pass kernels
    const int hmax = (c.arith >> kArithShift) & 7;
    if (big) {
      // forward headroom: from the moduli (index 0) or none (index 1)
      const int heads[2] = {-1, 0};
      sweep("fwd_headroom", hmax > 0 ? 2 : 1, [&](size_t i) { c.fwd_headroom = heads[i]; }, 0);
      const uint32_t rows[2] = {8, 16};
      sweep("row_residues", 2, [&](size_t i) { c.row16 = rows[i] == 16; }, 0);
      const uint32_t caps[4] = {0, 6, 5, 4};
      sweep("wg_per_cu", 4, [&](size_t i) { c.wg_cap = caps[i]; }, 0);
    }
    // tower chunk: 0 (the context's default grouping) or groups smaller than the run
    std::vector<uint32_t> chunks = {0};
    for (uint32_t ch : {1024u, 512u, 448u, 256u, 128u, 64u})
      if (ch < towers) chunks.push_back(ch);
    sweep("ntt_chunk", chunks.size(), [&](size_t i) { c.ntt_chunk = chunks[i]; }, 0);
    rep += ",\"best_ms\":" + std::to_string(ntt_ms()) + "}";
    // conversion geometry: first half of the basis -> the rest (<= 6 -> <= 24)
    if (cs >= 1 && ct >= 1) {
      std::vector<size_t> src, dst;
      for (size_t i = 0; i < cs; ++i) src.push_back(i);
      for (size_t i = 0; i < ct; ++i) dst.push_back(cs + i);
      auto plan = basis_plan(c, src, dst);
      const uint32_t polys = (uint32_t)cpolys;  // polys * cs <= a_words / n, polys * ct <= b_words / n
      rep += ",\"conversion\":{";
      float cbest = 1e30f;
      uint32_t best_targets = 0;
      for (uint32_t tpb : {0u, 1u, 2u, 4u, 6u, 8u, 12u, 16u, 24u}) {
        if (tpb > ct) continue;
        BasisLaunch bl{};
        bl.plan = plan->dev;
        bl.in = a;
        bl.out = b;
        bl.in_poly_stride = cs * n;
        bl.out_poly_stride = ct * n;
        bl.n = (uint32_t)n;
        bl.targets_per_block = tpb;
        bl.plan_host_s = plan->s;
        bl.plan_host_t = plan->t;
        bl.skip_at = 0xffffffffu;
        const float ms = time_ms(c, [&] { hip_ok(launch_basis_switch(bl, plan->per_term, polys, c.stream), "tune conv"); }, 4);
        rep += (rep.back() == '{' ? "" : ",") + std::string("\"") + std::to_string(tpb) + "\":" + std::to_string(ms);
        // 0 (geometry from the grid size) adapts to every batch size; a fixed
        // width must be clearly faster to replace it
        if (tpb == 0 ? ms < cbest : ms < cbest * 0.95f) {
          cbest = ms;
          best_targets = tpb;
        }
      }
      c.conv_targets = best_targets;
      rep += "}";
    }
    HIP_CHECK(hipStreamSynchronize(c.stream));
    rep += ",\"chosen\":{\"fwd_headroom\":" + std::to_string(c.fwd_headroom) +
           ",\"row_residues\":" + std::to_string(c.row16 ? 16 : 8) + ",\"wg_per_cu\":" + std::to_string(c.wg_cap) +
           ",\"ntt_chunk\":" + std::to_string(c.ntt_chunk) + ",\"conv_targets\":" + std::to_string(c.conv_targets) +
           "},\"towers\":" + std::to_string(towers) + ",\"n\":" + std::to_string(n) + "}";
    if (report && capacity) std::snprintf(report, capacity, "%s", rep.c_str());
    return NTT_HIP_OK;
  });
}

}  // extern "C"
