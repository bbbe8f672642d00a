// Synchronous transform class over the C ABI (include/ntt/hip_ntt.hpp).
#include "ntt/hip_ntt.hpp"

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

struct HipNttTransform::Impl {
  int device = 0;
  uint64_t n = 0;
  std::vector<uint64_t> moduli;
  std::vector<NegacyclicNtt> refs;
  ntt_hip_context *ctx = nullptr;
  uint64_t *a = nullptr, *b = nullptr;
  size_t count = 0;
  ~Impl() {
    if (a) ntt_hip_free(device, a);
    if (b) ntt_hip_free(device, b);
    if (ctx) ntt_hip_context_destroy(ctx);
  }
  void *stream() const { return ntt_hip_context_stream(ctx); }
  void put(uint64_t *dst, const std::vector<uint64_t> &v) {
    ok(ntt_hip_memcpy(device, dst, v.data(), count * 8, NTT_HIP_H2D, stream()), "upload");
  }
  void get(std::vector<uint64_t> &v, const uint64_t *src) {
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
}

HipNttTransform::~HipNttTransform() = default;
uint64_t HipNttTransform::n() const { return impl_->n; }
uint64_t HipNttTransform::limbs() const { return impl_->moduli.size(); }
uint64_t HipNttTransform::q(uint64_t limb) const { return impl_->moduli.at(limb); }
const NegacyclicNtt &HipNttTransform::ref(uint64_t limb) const { return impl_->refs.at(limb); }

void HipNttTransform::forward(std::vector<uint64_t> &values) {
  detail::validate_values(impl_->n, impl_->moduli, values);
  impl_->put(impl_->a, values);
  impl_->transform(impl_->a, false);
  impl_->get(values, impl_->a);
}

void HipNttTransform::inverse(std::vector<uint64_t> &values) {
  detail::validate_values(impl_->n, impl_->moduli, values);
  impl_->put(impl_->a, values);
  impl_->transform(impl_->a, true);
  impl_->get(values, impl_->a);
}

void HipNttTransform::pointwise_multiply(std::vector<uint64_t> &left, const std::vector<uint64_t> &right) {
  detail::validate_values(impl_->n, impl_->moduli, left);
  detail::validate_values(impl_->n, impl_->moduli, right);
  impl_->put(impl_->a, left);
  impl_->put(impl_->b, right);
  ok(ntt_hip_pointwise(impl_->ctx, impl_->a, impl_->b, 0, impl_->moduli.size(), impl_->stream()), "pointwise");
  impl_->get(left, impl_->a);
}

std::vector<uint64_t> HipNttTransform::multiply(std::vector<uint64_t> left, const std::vector<uint64_t> &right) {
  detail::validate_values(impl_->n, impl_->moduli, left);
  detail::validate_values(impl_->n, impl_->moduli, right);
  impl_->put(impl_->a, left);
  impl_->put(impl_->b, right);
  impl_->transform(impl_->a, false);
  impl_->transform(impl_->b, false);
  ok(ntt_hip_pointwise(impl_->ctx, impl_->a, impl_->b, 0, impl_->moduli.size(), impl_->stream()), "pointwise");
  impl_->transform(impl_->a, true);
  impl_->get(left, impl_->a);
  return left;
}

}  // namespace ntt::hip
