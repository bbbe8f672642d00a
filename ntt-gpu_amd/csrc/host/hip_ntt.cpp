// Synchronous transform class over the C ABI (include/ntt/hip_ntt.hpp).
//
// Every public call starts and ends with host vectors. Up to 512 MiB of
// residues per operand go through pinned staging owned by the transform:
// one parallel host pass checks the residues are canonical while copying them
// into the staging (no device work starts before the check has passed), the
// transfers then run at pinned-memory speed, and the result is copied back out
// in parallel. Larger calls transfer from the caller's memory directly.
#include "ntt/hip_ntt.hpp"

#include <cstring>
#include <string>

#include "ntt_hip.h"

namespace ntt::hip {
namespace {
void ok(int status, const char *what) {
  if (status == NTT_HIP_OK) return;
  const std::string msg = std::string(what) + ": " + ntt_hip_last_error();
  if (status == NTT_HIP_EINVAL) throw std::invalid_argument(msg);
  throw std::runtime_error("HIP: " + msg);
}
}  // namespace

constexpr size_t kStagingLimit = size_t{512} << 20;  // bytes per operand
constexpr uint64_t kBlock = 8192;                      // residues per parallel task

struct HipNttTransform::Impl {
  int device = 0;
  uint64_t n = 0;
  std::vector<uint64_t> moduli;
  std::vector<NegacyclicNtt> refs;
  ntt_hip_context *ctx = nullptr;
  uint64_t *a = nullptr, *b = nullptr;
  uint64_t *pinned = nullptr;  // two operands' staging, or null above the limit
  size_t count = 0;
  ~Impl() {
    if (pinned) ntt_hip_host_free(pinned);
    if (a) ntt_hip_free(device, a);
    if (b) ntt_hip_free(device, b);
    if (ctx) ntt_hip_context_destroy(ctx);
  }
  void *stream() const { return ntt_hip_context_stream(ctx); }
  uint64_t *staging(int slot) const { return pinned ? pinned + (size_t)slot * count : nullptr; }

  // Copy v into dst while checking every residue is canonical; true if all are.
  bool stage_in(uint64_t *dst, const std::vector<uint64_t> &v) const {
    const long long limbs = (long long)moduli.size(), blocks = (long long)((n + kBlock - 1) / kBlock);
    bool bad = false;
#pragma omp parallel for collapse(2) schedule(static) reduction(|| : bad) if (count >= (size_t{1} << 20))
    for (long long k = 0; k < limbs; ++k)
      for (long long blk = 0; blk < blocks; ++blk) {
        const uint64_t m = moduli[k], lo = (uint64_t)blk * kBlock, hi = std::min<uint64_t>(n, lo + kBlock);
        const uint64_t *s = v.data() + (uint64_t)k * n;
        uint64_t *d = dst + (uint64_t)k * n;
        bool b = false;
        for (uint64_t i = lo; i < hi; ++i) {
          b |= s[i] >= m;
          d[i] = s[i];
        }
        bad = bad || b;
      }
    return !bad;
  }
  void stage_out(std::vector<uint64_t> &v, const uint64_t *src) const {
    const long long parts = (long long)((count + kBlock - 1) / kBlock);
#pragma omp parallel for schedule(static) if (count >= (size_t{1} << 20))
    for (long long p = 0; p < parts; ++p) {
      const size_t lo = (size_t)p * kBlock, hi = std::min<size_t>(count, lo + kBlock);
      std::memcpy(v.data() + lo, src + lo, (hi - lo) * 8);
    }
  }

  // Check one operand (and copy it into staging slot 0 or 1); no device work.
  // Returns the host address the upload reads from.
  const uint64_t *check(const std::vector<uint64_t> &v, int slot) const {
    if (v.size() != count) throw std::length_error("HIP buffer must have limbs*N elements");
    if (uint64_t *st = staging(slot)) {
      if (!stage_in(st, v)) throw std::invalid_argument("HIP requires canonical residues");
      return st;
    }
    detail::validate_values(n, moduli, v);
    return v.data();
  }
  void upload(uint64_t *dst, const uint64_t *host) {
    ok(ntt_hip_memcpy(device, dst, host, count * 8, NTT_HIP_H2D, stream()), "upload");
  }
  void get(std::vector<uint64_t> &v, const uint64_t *src) {
    if (uint64_t *st = staging(0)) {
      ok(ntt_hip_memcpy(device, st, src, count * 8, NTT_HIP_D2H, stream()), "download");
      ok(ntt_hip_stream_synchronize(device, stream()), "synchronize");
      stage_out(v, st);
      return;
    }
    ok(ntt_hip_memcpy(device, v.data(), src, count * 8, NTT_HIP_D2H, stream()), "download");
    ok(ntt_hip_stream_synchronize(device, stream()), "synchronize");
  }
  void transform(uint64_t *data, bool inverse) {
    ok(ntt_hip_transform(ctx, data, nullptr, 0, moduli.size(), inverse, stream()), "transform");
  }
};

HipNttTransform::HipNttTransform(uint64_t n, const std::vector<uint64_t> &moduli,
                                 const std::vector<uint64_t> &roots, int device)
    : impl_(std::make_unique<Impl>()) {
  const auto shape = detail::validate_parameters(n, moduli, roots);
  auto &m = *impl_;
  m.device = device;
  m.n = n;
  m.moduli = moduli;
  m.count = shape.values;
  for (size_t k = 0; k < moduli.size(); ++k) m.refs.emplace_back(n, moduli[k], roots[k]);
  ok(ntt_hip_context_create(&m.ctx, device, n, moduli.data(), roots.data(), moduli.size()), "context");
  void *p = nullptr;
  ok(ntt_hip_malloc(device, &p, m.count * 8), "allocate");
  m.a = static_cast<uint64_t *>(p);
  ok(ntt_hip_malloc(device, &p, m.count * 8), "allocate");
  m.b = static_cast<uint64_t *>(p);
  if (m.count * 8 <= kStagingLimit) {
    ok(ntt_hip_host_alloc(&p, 2 * m.count * 8), "pinned staging");
    m.pinned = static_cast<uint64_t *>(p);
  }
}

HipNttTransform::~HipNttTransform() = default;
uint64_t HipNttTransform::n() const { return impl_->n; }
uint64_t HipNttTransform::limbs() const { return impl_->moduli.size(); }
uint64_t HipNttTransform::q(uint64_t limb) const { return impl_->moduli.at(limb); }
const NegacyclicNtt &HipNttTransform::ref(uint64_t limb) const { return impl_->refs.at(limb); }

void HipNttTransform::forward(std::vector<uint64_t> &values) {
  impl_->upload(impl_->a, impl_->check(values, 0));
  impl_->transform(impl_->a, false);
  impl_->get(values, impl_->a);
}

void HipNttTransform::inverse(std::vector<uint64_t> &values) {
  impl_->upload(impl_->a, impl_->check(values, 0));
  impl_->transform(impl_->a, true);
  impl_->get(values, impl_->a);
}

void HipNttTransform::pointwise_multiply(std::vector<uint64_t> &left, const std::vector<uint64_t> &right) {
  const uint64_t *l = impl_->check(left, 0), *r = impl_->check(right, 1);  // both before any device work
  impl_->upload(impl_->a, l);
  impl_->upload(impl_->b, r);
  ok(ntt_hip_pointwise(impl_->ctx, impl_->a, impl_->b, 0, impl_->moduli.size(), impl_->stream()), "pointwise");
  impl_->get(left, impl_->a);
}

std::vector<uint64_t> HipNttTransform::multiply(std::vector<uint64_t> left, const std::vector<uint64_t> &right) {
  const uint64_t *l = impl_->check(left, 0), *r = impl_->check(right, 1);  // both before any device work
  impl_->upload(impl_->a, l);
  impl_->upload(impl_->b, r);
  impl_->transform(impl_->a, false);
  impl_->transform(impl_->b, false);
  ok(ntt_hip_pointwise(impl_->ctx, impl_->a, impl_->b, 0, impl_->moduli.size(), impl_->stream()), "pointwise");
  impl_->transform(impl_->a, true);
  impl_->get(left, impl_->a);
  return left;
}

}  // namespace ntt::hip
