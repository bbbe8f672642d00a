// Basis conversion, key inner product, ModDown correction, automorphism and
// pointwise kernels for gfx950.
//
// Basis conversion follows the approximate (unreduced) CRT of reference
// include/ntt/basis_conversion.hpp:13-30: y_i = [x_i * Bhat_i^-1]_{b_i},
// out_m = [sum_i y_i * Bhat_i]_m with NO reduction modulo B. Instead of one
// mul_mod + add_mod per term (reference), the sum is accumulated exactly in
// 128 bits and reduced once per output with Montgomery, the hat residues
// being stored as (Bhat_i mod m) * 2^64 mod m so the result is the plain
// residue. A per-term reduction variant covers sums that could exceed
// m * 2^64 (very wide or very many source primes).
//
// Up to 16 sources (every ModUp digit and ModDown of the key-switch shapes)
// take a register kernel specialised on the source count: each scaled source
// y_i is split into three 21-bit limbs y_ij (kept in VGPRs), and the plan
// holds, per (target, source, limb), h_ij = Bhat_i * 2^64 * 2^(21 j) mod m as
// two 32-bit halves (read through scalar loads). Then
//   sum_i y_i Bhat_i 2^64 = sum_ij y_ij h_ij = C_lo + 2^32 C_hi (mod m),
// C_lo = sum y_ij lo32(h_ij), C_hi = sum y_ij hi32(h_ij): 6s v_mad_u64_u32
// into two 64-bit columns that cannot overflow (each below 48 * 2^53), one
// shifted add and one Montgomery reduction per output, exact for any
// source primes (the sum is below 3 s 2^21 m < m 2^64).
#include <algorithm>
#include <utility>

#include "conv_device.hpp"

namespace nttamd {

// ---------------------------------------------------------------- basis switch
// blockIdx.x: 256 coefficients; blockIdx.y: (poly, target chunk).
// Dynamic LDS: s * 256 words for the scaled source residues y_i.
template <bool PerTerm>
__global__ __launch_bounds__(256) void basis_switch_kernel(BasisLaunch a) {
  extern __shared__ uint64_t ys[];
  const BasisConst &p = *a.plan;
  const uint32_t s = p.s, t = p.t;
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  const uint32_t chunks = (t + a.targets_per_block - 1) / a.targets_per_block;
  const uint32_t poly = blockIdx.y / chunks, chunk = blockIdx.y - poly * chunks;
  const uint64_t *in = a.in + (uint64_t)poly * a.in_poly_stride;
  uint64_t *out = a.out + (uint64_t)poly * a.out_poly_stride;
  if (i >= a.n) return;
  for (uint32_t k = 0; k < s; ++k) {
    const uint64_t x = in[(uint64_t)k * a.n + i];
    ys[k * 256 + threadIdx.x] = shoup(x, p.src[k].inv, p.src[k].inv_shoup, p.src[k].q);
  }
  const uint32_t m0 = chunk * a.targets_per_block;
  const uint32_t m1 = min(t, m0 + a.targets_per_block);
  for (uint32_t m = m0; m < m1; ++m) {
    const uint64_t q = p.dst[m].q, qn = p.dst[m].qinv_neg;
    const uint64_t *hat = p.hat_mont + m;  // hat_mont[k * t + m]
    uint64_t r;
    if constexpr (!PerTerm) {
      U128 acc{0, 0};
      for (uint32_t k = 0; k < s; ++k) add_wide(acc, mul_wide(ys[k * 256 + threadIdx.x], hat[(uint64_t)k * t]));
      r = mont_reduce(acc, q, qn);
    } else {
      r = 0;
      for (uint32_t k = 0; k < s; ++k)
        r = add_mod(r, mont_reduce(mul_wide(ys[k * 256 + threadIdx.x], hat[(uint64_t)k * t]), q, qn), q);
    }
    const uint32_t ot = m < a.skip_at ? m : m + a.skip_len;
    out[(uint64_t)ot * a.n + i] = r;
  }
}

// Scalar-load views of the plan (constant address space: uniform reads go
// through the scalar cache).
typedef const __attribute__((address_space(4))) SourceConst *CSrc;
typedef const __attribute__((address_space(4))) TargetConst *CDst;
typedef const __attribute__((address_space(4))) uint32_t *CLimbs;

constexpr uint32_t kRegSources = 16;

// Same grid as basis_switch_kernel; S = s source towers (1..kRegSources).
template <int S>
__global__ __launch_bounds__(256) void basis_switch_regs(BasisLaunch a) {
  const BasisConst &p = *a.plan;
  const CSrc src = (CSrc)(const void *)p.src;
  const CDst dst = (CDst)(const void *)p.dst;
  const CLimbs limbs = (CLimbs)(const void *)p.hat_limbs;
  const uint32_t t = a.plan_host_t;
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  const uint32_t chunks = (t + a.targets_per_block - 1) / a.targets_per_block;
  const uint32_t poly = blockIdx.y / chunks, chunk = blockIdx.y - poly * chunks;
  const uint64_t *in = a.in + (uint64_t)poly * a.in_poly_stride;
  uint64_t *out = a.out + (uint64_t)poly * a.out_poly_stride;
  if (i >= a.n) return;
  constexpr uint32_t kMask21 = (1u << 21) - 1;
  uint32_t y[3 * S];
#pragma unroll
  for (int k = 0; k < S; ++k) {
    const uint64_t v = shoup(in[(uint64_t)k * a.n + i], src[k].inv, src[k].inv_shoup, src[k].q);
    y[3 * k] = (uint32_t)v & kMask21;
    y[3 * k + 1] = (uint32_t)(v >> 21) & kMask21;
    y[3 * k + 2] = (uint32_t)(v >> 42);
  }
  const uint32_t m0 = chunk * a.targets_per_block;
  const uint32_t m1 = min(t, m0 + a.targets_per_block);
  for (uint32_t m = m0; m < m1; ++m) {
    const CLimbs h = limbs + (uint64_t)m * (6 * S);
    uint64_t lo = 0, hi = 0;
#pragma unroll
    for (int j = 0; j < 3 * S; ++j) {
      lo += (uint64_t)y[j] * h[2 * j];
      hi += (uint64_t)y[j] * h[2 * j + 1];
    }
    // lo + 2^32 hi: the shifted column's low word is zero, so one add with
    // carry on the middle words (hi >> 32 < 2^28 takes the carry)
    uint32_t carry;
    const uint32_t mid = __builtin_addc((uint32_t)(lo >> 32), (uint32_t)hi, 0u, &carry);
    const U128 sum{((uint64_t)mid << 32) | (uint32_t)lo, (hi >> 32) + carry};
    const uint64_t r = mont_reduce(sum, dst[m].q, dst[m].qinv_neg);
    const uint32_t ot = m < a.skip_at ? m : m + a.skip_len;
    out[(uint64_t)ot * a.n + i] = r;
  }
}

typedef void (*BasisKernel)(BasisLaunch);
template <int... S>
static BasisKernel regs_kernel(uint32_t s, std::integer_sequence<int, S...>) {
  static constexpr BasisKernel table[] = {basis_switch_regs<S + 1>...};
  return table[s - 1];
}

// Targets per workgroup when the caller leaves it to the launcher: as many as
// keep at least ~2048 workgroups in flight (the y_i are recomputed per chunk).
static uint32_t auto_targets(uint32_t t, uint32_t n, uint32_t polys) {
  const uint32_t blocks = ((n + 255) / 256) * polys;
  uint32_t chunks = (2048 + blocks - 1) / blocks;
  chunks = chunks < 1 ? 1 : (chunks > t ? t : chunks);
  return (t + chunks - 1) / chunks;
}

hipError_t launch_basis_switch(const BasisLaunch &in, bool per_term, uint32_t polys, hipStream_t stream) {
  BasisLaunch a = in;
  if (a.targets_per_block == 0) a.targets_per_block = auto_targets(a.plan_host_t, a.n, polys);
  const uint32_t chunks = (a.plan_host_t + a.targets_per_block - 1) / a.targets_per_block;
  dim3 grid((a.n + 255) / 256, chunks * polys);
  if (a.plan_host_s >= 1 && a.plan_host_s <= kRegSources) {
    hipLaunchKernelGGL(regs_kernel(a.plan_host_s, std::make_integer_sequence<int, kRegSources>()), grid, dim3(256),
                       0, stream, a);
    return hipGetLastError();
  }
  const size_t lds = (size_t)a.plan_host_s * 256 * sizeof(uint64_t);
  if (lds > 64 * 1024) {
    // up to 40 sources: 80 KiB of the 160 KiB gfx950 LDS; a refused opt-in is
    // reported here rather than as a failed launch
    const void *kernel = per_term ? (const void *)basis_switch_kernel<true> : (const void *)basis_switch_kernel<false>;
    const hipError_t e = hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  if (per_term) hipLaunchKernelGGL(basis_switch_kernel<true>, grid, dim3(256), lds, stream, a);
  else hipLaunchKernelGGL(basis_switch_kernel<false>, grid, dim3(256), lds, stream, a);
  return hipGetLastError();
}

// ---------------------------------------------------------------- key inner product
// Reference per-digit semantics (resident_adapter.inc:31-43, resident_conversion.hpp:50-54):
// acc = first ? d*k : acc + d*k (mod q), canonical key rows.
__global__ __launch_bounds__(256) void mac_digit_kernel(uint64_t *b, uint64_t *a, const uint64_t *digit,
                                                        const uint64_t *kb, const uint64_t *ka,
                                                        const PrimeConst *primes, uint32_t first_prime,
                                                        uint32_t n, uint32_t towers, int first_digit) {
  const uint64_t idx = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (uint64_t)n * towers) return;
  const PrimeConst &pc = primes[first_prime + idx / n];
  const uint64_t d = digit[idx];
  // mont(mont(d*k) * 2^128) = d*k mod q
  const uint64_t pb = mont_mul(mont_mul(d, kb[idx], pc.q, pc.qinv_neg), pc.r2, pc.q, pc.qinv_neg);
  const uint64_t pa = mont_mul(mont_mul(d, ka[idx], pc.q, pc.qinv_neg), pc.r2, pc.q, pc.qinv_neg);
  b[idx] = first_digit ? pb : add_mod(b[idx], pb, pc.q);
  a[idx] = first_digit ? pa : add_mod(a[idx], pa, pc.q);
}

hipError_t launch_mac_digit(uint64_t *b, uint64_t *a, const uint64_t *digit, const uint64_t *kb,
                            const uint64_t *ka, const PrimeConst *primes, uint32_t first_prime, uint32_t n,
                            uint32_t towers, bool first_digit, hipStream_t stream) {
  const uint64_t total = (uint64_t)n * towers;
  hipLaunchKernelGGL(mac_digit_kernel, dim3((uint32_t)((total + 255) / 256)), dim3(256), 0, stream, b, a, digit, kb,
                     ka, primes, first_prime, n, towers, (int)first_digit);
  return hipGetLastError();
}

// The same accumulation over m.count digits in one pass (each product a
// Barrett reduction, the sum kept canonical), so b and a are read and written
// once instead of once per digit; results equal the per-digit sequence.
__global__ __launch_bounds__(256) void mac_digits_kernel(MacDigitsLaunch m) {
  const uint64_t idx = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (uint64_t)m.n * m.towers) return;
  const PrimeConst &pc = m.primes[m.first_prime + idx / m.n];
  uint64_t rb = m.first_digit ? 0 : m.b[idx], ra = m.first_digit ? 0 : m.a[idx];
  for (uint32_t j = 0; j < m.count; ++j) {
    const uint64_t d = m.digits[j][idx];
    rb = add_mod(rb, barrett_mul(d, m.key_b[j][idx], pc.q, pc.mu_lo, pc.mu_hi), pc.q);
    ra = add_mod(ra, barrett_mul(d, m.key_a[j][idx], pc.q, pc.mu_lo, pc.mu_hi), pc.q);
  }
  m.b[idx] = rb;
  m.a[idx] = ra;
}

hipError_t launch_mac_digits(const MacDigitsLaunch &m, hipStream_t stream) {
  if (m.count == 0 || m.count > kMacDigitsPerLaunch) return hipErrorInvalidValue;
  const uint64_t total = (uint64_t)m.n * m.towers;
  hipLaunchKernelGGL(mac_digits_kernel, dim3((uint32_t)((total + 255) / 256)), dim3(256), 0, stream, m);
  return hipGetLastError();
}

// Fused inner product over all digits with resident Montgomery-form keys:
// Z[t][i] = sum_d D_d[t][i] * K_d[t][i] for both channels, one reduction per output.
// digits: poly-major (poly, digit, tower, i); keys: (digit, tower, i) in Montgomery form.
// blockIdx.y = tower * polys + poly: the polynomials of one tower run back to
// back, so the tower's key rows (digits x 2 channels) are re-read from L2
// rather than from HBM once per polynomial.
template <bool PerTerm>
__global__ __launch_bounds__(256) void mac_fused_kernel(MacLaunch m, uint32_t polys) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  const uint32_t tower = m.first_tower + blockIdx.y / polys, poly = blockIdx.y % polys;
  if (i >= m.n) return;
  const PrimeConst &pc = m.primes[m.tower_prime[tower]];
  const uint64_t *dg = m.digits + (uint64_t)poly * m.digit_poly_stride + (uint64_t)tower * m.n + i;
  const uint64_t *kb = m.key_b + (uint64_t)tower * m.n + i;
  const uint64_t *ka = m.key_a + (uint64_t)tower * m.n + i;
  const uint64_t step = (uint64_t)m.towers * m.n;
  const uint32_t own_digit = tower < m.active_q ? tower / m.alpha : 0xffffffffu;
  const uint64_t *own_base = m.own_table ? m.own_table[poly] : m.own + (uint64_t)poly * m.own_poly_stride;
  const uint64_t own = tower < m.active_q ? own_base[(uint64_t)tower * m.n + i] : 0;
  uint64_t rb, ra;
  if constexpr (!PerTerm) {
    U128 sb{0, 0}, sa{0, 0};
    for (uint32_t d = 0; d < m.digits_count; ++d) {
      const uint64_t x = d == own_digit ? own : dg[d * step];
      add_wide(sb, mul_wide63(x, kb[d * step]));
      add_wide(sa, mul_wide63(x, ka[d * step]));
    }
    rb = mont_reduce(sb, pc.q, pc.qinv_neg);
    ra = mont_reduce(sa, pc.q, pc.qinv_neg);
  } else {
    rb = ra = 0;
    for (uint32_t d = 0; d < m.digits_count; ++d) {
      const uint64_t x = d == own_digit ? own : dg[d * step];
      rb = add_mod(rb, mont_mul(x, kb[d * step], pc.q, pc.qinv_neg), pc.q);
      ra = add_mod(ra, mont_mul(x, ka[d * step], pc.q, pc.qinv_neg), pc.q);
    }
  }
  const uint64_t o = (uint64_t)poly * m.out_poly_stride + (uint64_t)tower * m.n + i;
  m.out_b[o] = rb;
  m.out_a[o] = ra;
}

// mac_fused_kernel<false> for a digit count D fixed at compile time: the
// digit loop unrolls, so every digit and key load is issued before the first
// product waits on one (with a runtime count each iteration waited for its
// own loads, one memory round trip per digit).
template <int D>
__global__ __launch_bounds__(256) void mac_fused_fixed(MacLaunch m, uint32_t polys) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  const uint32_t tower = m.first_tower + blockIdx.y / polys, poly = blockIdx.y % polys;
  if (i >= m.n) return;
  // the modulus through scalar loads, issued with the others
  typedef const __attribute__((address_space(4))) PrimeConst *CPrime;
  typedef const __attribute__((address_space(4))) uint32_t *CIndex;
  const uint32_t prime = ((CIndex)(const void *)m.tower_prime)[tower];
  const uint64_t q = ((CPrime)(const void *)m.primes)[prime].q;
  const uint64_t qinv_neg = ((CPrime)(const void *)m.primes)[prime].qinv_neg;
  const uint64_t *dg = m.digits + (uint64_t)poly * m.digit_poly_stride + (uint64_t)tower * m.n + i;
  const uint64_t *kb = m.key_b + (uint64_t)tower * m.n + i;
  const uint64_t *ka = m.key_a + (uint64_t)tower * m.n + i;
  const uint64_t step = (uint64_t)m.towers * m.n;
  const uint32_t own_digit = tower < m.active_q ? tower / m.alpha : 0xffffffffu;
  // digit d's own towers come from the input: the address is selected, not
  // the value, so no load waits behind a branch (P towers, tower >= active_q,
  // have no own digit)
  const uint64_t *ownp = (m.own_table ? m.own_table[poly] : m.own + (uint64_t)poly * m.own_poly_stride) +
                        (uint64_t)tower * m.n + i;
  uint64_t x[D], b[D], a[D];
#pragma unroll
  for (int d = 0; d < D; ++d) {
    x[d] = *((uint32_t)d == own_digit ? ownp : dg + d * step);
    b[d] = kb[d * step];
    a[d] = ka[d * step];
  }
  U128 sb{0, 0}, sa{0, 0};
#pragma unroll
  for (int d = 0; d < D; ++d) {
    add_wide(sb, mul_wide63(x[d], b[d]));
    add_wide(sa, mul_wide63(x[d], a[d]));
  }
  const uint64_t o = (uint64_t)poly * m.out_poly_stride + (uint64_t)tower * m.n + i;
  m.out_b[o] = mont_reduce(sb, q, qinv_neg);
  m.out_a[o] = mont_reduce(sa, q, qinv_neg);
}

constexpr uint32_t kMacFixedDigits = 8;
typedef void (*MacKernel)(MacLaunch, uint32_t);
template <int... D>
static MacKernel mac_kernel(uint32_t digits, std::integer_sequence<int, D...>) {
  static constexpr MacKernel table[] = {mac_fused_fixed<D + 1>...};
  return table[digits - 1];
}

hipError_t launch_mac_fused(const MacLaunch &m, bool per_term, uint32_t polys, hipStream_t stream) {
  dim3 grid((m.n + 255) / 256, (m.launch_towers ? m.launch_towers : m.towers) * polys);
  if (per_term) {
    hipLaunchKernelGGL(mac_fused_kernel<true>, grid, dim3(256), 0, stream, m, polys);
  } else if (m.digits_count >= 1 && m.digits_count <= kMacFixedDigits) {
    hipLaunchKernelGGL(mac_kernel(m.digits_count, std::make_integer_sequence<int, kMacFixedDigits>()), grid, dim3(256),
                       0, stream, m, polys);
  } else {
    hipLaunchKernelGGL(mac_fused_kernel<false>, grid, dim3(256), 0, stream, m, polys);
  }
  return hipGetLastError();
}

// Convert canonical rows to Montgomery form (x * 2^64 mod q) in place.
__global__ __launch_bounds__(256) void to_mont_kernel(uint64_t *data, const PrimeConst *primes,
                                                      const uint32_t *tower_prime, uint32_t n, uint32_t towers) {
  const uint64_t idx = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (uint64_t)n * towers) return;
  const PrimeConst &pc = primes[tower_prime[idx / n]];
  data[idx] = mont_mul(data[idx], pc.r2, pc.q, pc.qinv_neg);
}

hipError_t launch_to_mont(uint64_t *data, const PrimeConst *primes, const uint32_t *tower_prime, uint32_t n,
                          uint32_t towers, hipStream_t stream) {
  const uint64_t total = (uint64_t)n * towers;
  hipLaunchKernelGGL(to_mont_kernel, dim3((uint32_t)((total + 255) / 256)), dim3(256), 0, stream, data, primes,
                     tower_prime, n, towers);
  return hipGetLastError();
}

// ---------------------------------------------------------------- ModDown correction
// out = (left - right) * P^-1 mod q (reference resident_conversion.hpp:45-49,
// resident.cpp:404-430). scale[t] = {P^-1 mod q_t, Shoup}.
__global__ __launch_bounds__(256) void correct_kernel(CorrectLaunch c) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  const uint32_t poly = blockIdx.y / c.towers, tower = blockIdx.y - poly * c.towers;
  if (i >= c.n) return;
  const uint64_t q = c.primes[c.first_prime + tower].q;
  const uint64_t o = (uint64_t)tower * c.n + i;
  const uint64_t l = c.left[(uint64_t)poly * c.left_poly_stride + o];
  const uint64_t r = c.right[(uint64_t)poly * c.right_poly_stride + o];
  c.out[(uint64_t)poly * c.out_poly_stride + o] = shoup(sub_mod(l, r, q), c.scale[tower].x, c.scale[tower].y, q);
}

hipError_t launch_correct(const CorrectLaunch &c, uint32_t polys, hipStream_t stream) {
  hipLaunchKernelGGL(correct_kernel, dim3((c.n + 255) / 256, c.towers * polys), dim3(256), 0, stream, c);
  return hipGetLastError();
}

// ---------------------------------------------------------------- pointwise
__global__ __launch_bounds__(256) void pointwise_kernel(uint64_t *out, const uint64_t *right, const PrimeConst *primes,
                                                        uint32_t first_prime, uint32_t n, uint32_t towers) {
  const uint64_t idx = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (uint64_t)n * towers) return;
  const PrimeConst &pc = primes[first_prime + idx / n];
  out[idx] = barrett_mul(out[idx], right[idx], pc.q, pc.mu_lo, pc.mu_hi);
}

hipError_t launch_pointwise(uint64_t *out, const uint64_t *right, const PrimeConst *primes, uint32_t first_prime,
                            uint32_t n, uint32_t towers, hipStream_t stream) {
  const uint64_t total = (uint64_t)n * towers;
  hipLaunchKernelGGL(pointwise_kernel, dim3((uint32_t)((total + 255) / 256)), dim3(256), 0, stream, out, right, primes,
                     first_prime, n, towers);
  return hipGetLastError();
}

// ---------------------------------------------------------------- automorphism
// Reference resident_adapter.inc:44-59 / resident_conversion.hpp:15-29:
// out_b[t][k] = ext_b[t][src] + P*c0[t][src] on Q towers, out_a[t][k] = ext_a[t][src].
__device__ __forceinline__ uint32_t brev(uint32_t x, uint32_t lg) { return __brev(x) >> (32 - lg); }

// blockIdx.y = tower * polys + poly.
__global__ __launch_bounds__(256) void rotate_kernel(RotateLaunch r, uint32_t polys) {
  const uint32_t k = blockIdx.x * 256 + threadIdx.x;
  const uint32_t tower = blockIdx.y / polys, poly = blockIdx.y - tower * polys;
  if (k >= r.n) return;
  const uint32_t odd = 2 * brev(k, r.log_n) + 1;
  const uint32_t src_exp = (uint32_t)(((uint64_t)odd * r.automorphism) & (2ull * r.n - 1));
  const uint32_t src = brev(src_exp >> 1, r.log_n);
  const uint64_t at = (uint64_t)tower * r.n;
  const uint64_t *eb = r.ext_b + poly * r.ext_poly_stride, *ea = r.ext_a + poly * r.ext_poly_stride;
  uint64_t b = eb[at + src];
  if (tower < r.active_q) {
    const uint64_t q = r.primes[r.first_prime + tower].q;
    const uint64_t *c0 = r.c0 + poly * r.c0_poly_stride;
    b = add_mod(b, shoup(c0[at + src], r.scale[tower].x, r.scale[tower].y, q), q);
  }
  r.out_b[poly * r.out_poly_stride + at + k] = b;
  r.out_a[poly * r.out_poly_stride + at + k] = ea[at + src];
}

hipError_t launch_rotate(const RotateLaunch &r, uint32_t towers, uint32_t polys, hipStream_t stream) {
  hipLaunchKernelGGL(rotate_kernel, dim3((r.n + 255) / 256, towers * polys), dim3(256), 0, stream, r, polys);
  return hipGetLastError();
}

}  // namespace nttamd

namespace nttamd {
// Plain word copy: the calibration kernel for HBM byte counters (its
// algorithmic traffic is exactly 16 bytes per word).
__global__ __launch_bounds__(256) void copy_words_kernel(uint64_t *__restrict__ dst, const uint64_t *__restrict__ src,
                                                         uint64_t words) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < words; i += (uint64_t)gridDim.x * 256)
    dst[i] = src[i];
}

hipError_t launch_copy_words(uint64_t *dst, const uint64_t *src, uint64_t words, hipStream_t stream) {
  const uint64_t blocks = std::min<uint64_t>((words + 255) / 256, 65536);
  hipLaunchKernelGGL(copy_words_kernel, dim3((uint32_t)blocks), dim3(256), 0, stream, dst, src, words);
  return hipGetLastError();
}

// Seeded uniform words below `bound` (splitmix64 of the index, scaled by a
// 64x64 high product): the autotuners' operands. Timing on zero operands
// measures a different device: the batched key switch runs 6% faster on
// zeros than on uniform residues (the multipliers draw less power).
__global__ __launch_bounds__(256) void fill_uniform_kernel(uint64_t *__restrict__ dst, uint64_t words, uint64_t seed,
                                                           uint64_t bound) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < words; i += (uint64_t)gridDim.x * 256) {
    uint64_t z = seed + (i + 1) * 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    z ^= z >> 31;
    dst[i] = __umul64hi(z, bound);
  }
}

hipError_t launch_fill_uniform(uint64_t *dst, uint64_t words, uint64_t seed, uint64_t bound, hipStream_t stream) {
  const uint64_t blocks = std::min<uint64_t>((words + 255) / 256, 65536);
  hipLaunchKernelGGL(fill_uniform_kernel, dim3((uint32_t)blocks), dim3(256), 0, stream, dst, words, seed, bound);
  return hipGetLastError();
}
}  // namespace nttamd
