// Negacyclic NTT over 64-bit RNS towers for gfx950.
//
// Semantics (reference include/ntt/ntt.hpp:5-35, transform_stage.hpp:19-80):
// forward maps natural-order coefficients to evaluations at psi^(2 rev(k)+1)
// in bit-reversed order (Cooley-Tukey with T[m+g] = psi^bitrev(m+g)); inverse
// is the Gentleman-Sande mirror with psi^-1 and a final n^-1 scale.
//
// Schedule (our own, gfx950-first): N = R x 256 with R = N/256.
//   pass A: the log2(R) far stages, on columns (stride 256). A workgroup
//           owns 16 adjacent columns so every global access is a 128-byte
//           segment.
//   pass B: the 8 near stages, on 256-element rows; a workgroup owns 16 rows.
// In both passes a thread keeps 16 residues in registers and runs 4 radix-2
// stages on them, then trades through LDS for the other 4 ("P1" layout:
// j = u + 16k, "P2" layout: j = 16u + k, u = lane within the sub-transform).
// The row pass of wide-lazy and 32-bit towers at N = 2^16 instead keeps 8
// residues per thread and exchanges within waves (ntt_pass_b8_fwd / _inv).
// Butterflies for q < 2^61 are wide-lazy (forward residues in [0, 8q),
// inverse in [0, 4q), a Shoup quotient short by at most 2: modarith.hpp) with
// twiddles {W, floor(W 2^64/q)} read as one 16-byte load; a canonical variant
// serves 2^61 <= q < 2^63. Every kernel output is canonical except the
// internal hand-off between the two passes of one transform.
//
// Pass A for R < 256 (N = 2^9..2^15) and whole transforms for N <= 2^12 use
// a plain LDS-staged kernel (one radix-2 stage per barrier).
#include <type_traits>

#include "ntt_device.hpp"

namespace nttamd {

constexpr int kRow = 256;     // pass-B sub-transform length
constexpr int kLdsRow = 256 + 16 + 1;  // padded LDS row (one word per 16, plus one)

// Padding (rather than an XOR swizzle, measured slower: its index arithmetic
// costs more VALU than the extra workgroup per CU it allows) keeps both
// exchange patterns free of bank conflicts.
__device__ __forceinline__ int lds_idx(int row, int j) { return row * kLdsRow + j + (j >> 4); }

// Half-size exchange buffers (17.5 KiB, so the LDS no longer caps a CU at
// four workgroups; registers then allow 4-5): each exchange runs in two
// rounds with a barrier after each half. Measured 1.1% more batched key
// switches than one round through a 35 KiB buffer.
constexpr int kHalfRow = 128 + 8 + 1;
__device__ __forceinline__ int lds_half(int row, int j) { return row * kHalfRow + j + (j >> 4); }

// Column pass (rows of the buffer = the workgroup's 16 columns): round h moves
// the residues with j in [128h, 128h + 128). P1 -> P2: every thread writes its
// k in [8h, 8h + 8) (j = u + 16k), waves with u / 8 == h read j = 16u + k.
template <class V>
__device__ __forceinline__ void xchg_cols_p1_p2(V (&v)[16], V *lds, int cl, int u) {
  V w[16];  // v must stay intact until the second round has written it
  const int uh = u >> 3;  // threadIdx.x / 128: waves 0-1 or 2-3
#pragma unroll
  for (int h = 0; h < 2; ++h) {
#pragma unroll
    for (int k = 0; k < 8; ++k) lds[lds_half(cl, u + 16 * k)] = v[8 * h + k];
    __syncthreads();
    if (uh == h) {
#pragma unroll
      for (int k = 0; k < 16; ++k) w[k] = lds[lds_half(cl, 16 * (u & 7) + k)];
    }
    __syncthreads();
  }
#pragma unroll
  for (int k = 0; k < 16; ++k) v[k] = w[k];
}

// P1 -> P2 in one round through a full 16 x 256 buffer of 32-bit words (the
// packed forward column pass): every thread writes its j = u + 16k, then
// reads j = 16 u2 + k for its own u2.
__device__ __forceinline__ void xchg_cols_full(uint32_t (&v)[16], uint32_t *lds, int cl, int u, int u2) {
#pragma unroll
  for (int k = 0; k < 16; ++k) lds[lds_idx(cl, u + 16 * k)] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 16; ++k) v[k] = lds[lds_idx(cl, 16 * u2 + k)];
  __syncthreads();
}

// P2 -> P1: waves with u / 8 == h write j = 16u + k, every thread reads its
// k in [8h, 8h + 8).
template <class V>
__device__ __forceinline__ void xchg_cols_p2_p1(V (&v)[16], V *lds, int cl, int u) {
  V w[16];
  const int uh = u >> 3;  // threadIdx.x / 128: waves 0-1 or 2-3
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    if (uh == h) {
#pragma unroll
      for (int k = 0; k < 16; ++k) lds[lds_half(cl, 16 * (u & 7) + k)] = v[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 8; ++k) w[8 * h + k] = lds[lds_half(cl, u + 16 * k)];
    __syncthreads();
  }
#pragma unroll
  for (int k = 0; k < 16; ++k) v[k] = w[k];
}

// Row pass (16 rows per workgroup): rows 0-7 (waves 0-1) exchange first, then
// rows 8-15 (waves 2-3), through an 8-row buffer. ToP2: j = u + 16k -> 16u + k.
template <bool ToP2, class V>
__device__ __forceinline__ void xchg_rows(V (&v)[16], V *lds, int rl, int u) {
  const int half = rl >> 3, r8 = rl & 7;  // rows 0-7: waves 0-1
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    if (half == h) {
#pragma unroll
      for (int k = 0; k < 16; ++k) lds[lds_idx(r8, ToP2 ? u + 16 * k : 16 * u + k)] = v[k];
    }
    __syncthreads();
    if (half == h) {
#pragma unroll
      for (int k = 0; k < 16; ++k) v[k] = lds[lds_idx(r8, ToP2 ? 16 * u + k : u + 16 * k)];
    }
    __syncthreads();
  }
}

// Twiddle pairs {W, floor(W 2^64 / q)} read through a global-address-space
// pointer (one 16-byte global load, not a flat load).
typedef uint64_t tw2 __attribute__((ext_vector_type(2)));
typedef const __attribute__((address_space(1))) tw2 *TwPtr;

__device__ __forceinline__ TwPtr tw_table(const ulonglong2 *p) { return (TwPtr)(const void *)p; }

// Entry k of a twiddle table (at most 2^16 entries) as a 32-bit byte offset
// from the table's wave-uniform base: the load takes the scalar-base plus
// vector-offset form, with no 64-bit address arithmetic per twiddle (forward
// 59-bit transform 1% faster).
__device__ __forceinline__ tw2 tw_at(TwPtr tw, uint32_t k) {
  typedef const __attribute__((address_space(1))) char *BytePtr;
  return *(TwPtr)((BytePtr)tw + (k << 4));
}

// The first 256 entries of a twiddle table staged in LDS (the column pass's
// lane-dependent twiddles, the same for every column block of a prime).
typedef const __attribute__((address_space(3))) tw2 *TwLds;
__device__ __forceinline__ tw2 tw_at(TwLds tw, uint32_t k) { return tw[k]; }

// Arithmetic modes (the A template argument):
//   kLazy   q < 2^61: 64-bit wide-lazy (forward residues in [0, 8q), inverse in [0, 4q); ct8/gs8);
//   kExact  2^61 <= q < 2^63: canonical throughout;
//   kNarrow q < 2^30: 32-bit registers and LDS words (ct32/gs32, Harvey ranges).
// Global memory holds uint64 residues in every mode.
template <int A>
using Val = std::conditional_t<A == kNarrow, uint32_t, uint64_t>;

template <int A>
__device__ __forceinline__ void ct_bfly(Val<A> &x, Val<A> &y, tw2 w, const Lazy8 &m) {
  if constexpr (A == kLazy) {
    ct8(x, y, w.x, w.y, m);
  } else if constexpr (A == kNarrow) {
    ct32(x, y, (uint32_t)w.x, (uint32_t)(w.y >> 32), narrow_of(m.q));
  } else {
    const uint64_t t = shoup(y, w.x, w.y, m.q);
    y = sub_mod(x, t, m.q);
    x = add_mod(x, t, m.q);
  }
}

template <int A>
__device__ __forceinline__ void gs_bfly(Val<A> &x, Val<A> &y, tw2 w, const Lazy8 &m) {
  if constexpr (A == kLazy) {
    gs8(x, y, w.x, w.y, m);
  } else if constexpr (A == kNarrow) {
    gs32(x, y, (uint32_t)w.x, (uint32_t)(w.y >> 32), narrow_of(m.q));
  } else {
    const uint64_t s = add_mod(x, y, m.q);
    y = shoup(sub_mod(x, y, m.q), w.x, w.y, m.q);
    x = s;
  }
}

// Final inverse stage with n^-1 folded in: x' = (x+y) n^-1, y' = (x-y) W n^-1,
// both canonical (lazy inputs in [0, 4q), narrow inputs in [0, 2q)).
template <int A>
__device__ __forceinline__ void gs_bfly_last(Val<A> &x, Val<A> &y, const PrimeConst &pc, const Lazy8 &m) {
  if constexpr (A == kLazy) {
    const uint64_t s = x + y;
    const uint64_t d = x + m.q4 - y;
    x = reduce_once(shoup_lazy(s, pc.ninv, pc.ninv_shoup, pc.q), pc.q);
    y = reduce_once(shoup_lazy(d, pc.last_w, pc.last_wp, pc.q), pc.q);
  } else if constexpr (A == kNarrow) {
    const Narrow nm = narrow_of(pc.q);
    const uint32_t s = x + y, d = x - y + nm.q2;  // [0, 4q), (0, 4q)
    x = min_sub(shoup32(s, (uint32_t)pc.ninv, (uint32_t)(pc.ninv_shoup >> 32), nm.q), nm.q);
    y = min_sub(shoup32(d, (uint32_t)pc.last_w, (uint32_t)(pc.last_wp >> 32), nm.q), nm.q);
  } else {
    const uint64_t s = add_mod(x, y, pc.q);
    const uint64_t d = sub_mod(x, y, pc.q);
    x = shoup(s, pc.ninv, pc.ninv_shoup, pc.q);
    y = shoup(d, pc.last_w, pc.last_wp, pc.q);
  }
}

// Forward output -> canonical.
template <int A>
__device__ __forceinline__ Val<A> canon(Val<A> x, const Lazy8 &m) {
  if constexpr (A == kLazy) return canon8(x, m);
  else if constexpr (A == kNarrow) return canon4_32(x, narrow_of(m.q));
  else return x;
}

// ModDown correction operand: sub - NTT value, congruent mod q, below 2^64
// (the exact Shoup product that follows reduces it).
template <int A>
__device__ __forceinline__ uint64_t epi_diff(uint64_t sub, Val<A> x, const Lazy8 &m, uint64_t q) {
  if constexpr (A == kLazy) return sub + m.q4 - sub_if_ge_neg(x, m.nq4);
  else if constexpr (A == kNarrow) return sub_mod(sub, canon4_32(x, narrow_of(q)), q);
  else return sub_mod(sub, x, q);
}

// The forward reduction schedule of headroom class H (make_fwd_plan, modarith.hpp).
template <int H>
struct FwdPlan {
  static constexpr FwdPlanData d = make_fwd_plan(H);
};
template <int H>
struct InvPlan {
  static constexpr InvPlanData d = make_inv_plan(H);
};

// Inverse butterfly of stage g of plan H (wide-lazy towers; every other
// arithmetic, and H = 0, keeps gs_bfly's per-stage reduction).
template <int A, int H, int G>
__device__ __forceinline__ void gs_planned(Val<A> &x, Val<A> &y, tw2 w, const Lazy8 &m, const PrimeConst &pc) {
  if constexpr (A == kLazy && H > 0) {
    constexpr int red = InvPlan<H>::d.red[G];
    const uint64_t s = x + y;
    const uint64_t d = x + (pc.q << InvPlan<H>::d.c_log[G]) - y;
    if constexpr (red == 2) x = reduce2_est(s, pc.est_shift, pc.est_mul, m.nq);
    else if constexpr (red == 1) x = sub_if_ge_neg(s, m.nq4);
    else x = s;
    y = shoup4(d, w.x, w.y, m);
  } else {
    gs_bfly<A>(x, y, w, m);
  }
}

// Last inverse stage of plan H, n^-1 folded in, canonical outputs.
template <int A, int H>
__device__ __forceinline__ void gs_last_planned(Val<A> &x, Val<A> &y, const PrimeConst &pc, const Lazy8 &m) {
  if constexpr (A == kLazy && H > 0) {
    const uint64_t s = x + y;
    const uint64_t d = x + (pc.q << InvPlan<H>::d.c_log[15]) - y;
    x = reduce_once(shoup_lazy(s, pc.ninv, pc.ninv_shoup, pc.q), pc.q);
    y = reduce_once(shoup_lazy(d, pc.last_w, pc.last_wp, pc.q), pc.q);
  } else {
    gs_bfly_last<A>(x, y, pc, m);
  }
}

// 2^64 - 2^k q for k = 0..6 (k <= 2 precomputed per prime; larger ones as
// shifts of 2^64 - 4q, opaque to the optimizer, which otherwise folds the add
// back into a subtract with a carry chain).
struct NegMultiples {
  uint64_t c[7];
};
__device__ __forceinline__ NegMultiples neg_multiples(const Lazy8 &m) {
  NegMultiples r;
  r.c[0] = m.nq;
  r.c[1] = m.nq2;
  r.c[2] = m.nq4;
#pragma unroll
  for (int k = 3; k < 7; ++k) r.c[k] = m.nq4 << (k - 2);
  return r;
}

// Forward butterfly of stage g of plan H.
template <int A, int H, int G>
__device__ __forceinline__ void ct_planned(Val<A> &x, Val<A> &y, tw2 w, const Lazy8 &m, const NegMultiples &neg) {
  if constexpr (A == kLazy && H > 0) {
    constexpr int k = FwdPlan<H>::d.red[G];
    uint64_t xr = x;
    if constexpr (k >= 0) xr = sub_if_ge_neg(x, neg.c[k]);
    const uint64_t xn = mad_lo(xr, y, w.x, mulhi_est(y, w.y), m.nq);
    y = 2 * xr + m.q4 - xn;  // xr - T + 4q; the wrap of 2xr + 4q cancels mod 2^64
    x = xn;
  } else {
    ct_bfly<A>(x, y, w, m);
  }
}

// Forward output of plan H -> canonical: conditional subtracts from bounds
// up to 8q, from 16q and more the quotient estimate (reduce2_est, to [0, 2q))
// and one subtract.
template <int A, int H>
__device__ __forceinline__ Val<A> canon_planned(Val<A> x, const Lazy8 &m, const NegMultiples &neg,
                                                const PrimeConst &pc) {
  if constexpr (A == kLazy && H > 0) {
    constexpr int bl = FwdPlan<H>::d.bound_log;
    if constexpr (bl >= 4) {
      return sub_if_ge_neg(reduce2_est(x, pc.est_shift, pc.est_mul, m.nq), neg.c[0]);
    } else {
#pragma unroll
      for (int k = bl - 1; k >= 0; --k) x = sub_if_ge_neg(x, neg.c[k]);
      return x;
    }
  } else {
    return canon<A>(x, m);
  }
}

// Forward output of plan H -> below 8q (the input range of epi_diff).
template <int A, int H>
__device__ __forceinline__ Val<A> below8_planned(Val<A> x, const Lazy8 &m, const NegMultiples &neg,
                                                 const PrimeConst &pc) {
  if constexpr (A == kLazy && H > 0) {
    constexpr int bl = FwdPlan<H>::d.bound_log;
    if constexpr (bl >= 5) {
      return reduce2_est(x, pc.est_shift, pc.est_mul, m.nq);
    } else {
#pragma unroll
      for (int k = bl - 1; k >= 3; --k) x = sub_if_ge_neg(x, neg.c[k]);
    }
  }
  return x;
}

// Four register stages on v[0..15]. H runs 8,4,2,1 (forward) or 1,2,4,8
// (inverse). Twiddle index = ml * row_factor + group, ml = (8/h) * ml_scale,
// group = g_base * (8/h) + k / (2h).
template <bool Inverse, int A, class Tw>
__device__ __forceinline__ void reg_stages(Val<A> (&v)[16], Tw tw, uint32_t ml_scale, uint32_t row_factor,
                                           uint32_t g_base, const Lazy8 &m) {
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int h = Inverse ? (1 << s) : (8 >> s);
    const uint32_t ml = (8u / h) * ml_scale;
    const uint32_t gb = g_base * (8u / h);
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      if (k & h) continue;
      const uint32_t idx = ml * row_factor + gb + (uint32_t)(k / (2 * h));
      const tw2 w = tw_at(tw, idx);
      if constexpr (Inverse) gs_bfly<A>(v[k], v[k + h], w, m);
      else ct_bfly<A>(v[k], v[k + h], w, m);
    }
  }
}

// Four inverse register stages G0..G0+3 of plan H (distances 1, 2, 4, 8; the
// planned form of reg_stages<true>).
template <int A, int H, int G0, int S = 0, class Tw>
__device__ __forceinline__ void reg_stages_inv(Val<A> (&v)[16], Tw tw, uint32_t ml_scale, uint32_t row_factor,
                                               uint32_t g_base, const Lazy8 &m, const PrimeConst &pc) {
  if constexpr (S < 4) {
    constexpr int h = 1 << S;
    const uint32_t ml = (8u / h) * ml_scale;
    const uint32_t gb = g_base * (8u / h);
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      if (k & h) continue;
      const tw2 w = tw_at(tw, ml * row_factor + gb + (uint32_t)(k / (2 * h)));
      gs_planned<A, H, G0 + S>(v[k], v[k + h], w, m, pc);
    }
    reg_stages_inv<A, H, G0, S + 1>(v, tw, ml_scale, row_factor, g_base, m, pc);
  }
}

// Four forward register stages G0..G0+3 of plan H (the forward form of
// reg_stages; distances 8, 4, 2, 1).
template <int A, int H, int G0, int S = 0, class Tw>
__device__ __forceinline__ void reg_stages_fwd(Val<A> (&v)[16], Tw tw, uint32_t ml_scale, uint32_t row_factor,
                                               uint32_t g_base, const Lazy8 &m, const NegMultiples &neg) {
  if constexpr (S < 4) {
    constexpr int h = 8 >> S;
    const uint32_t ml = (8u / h) * ml_scale;
    const uint32_t gb = g_base * (8u / h);
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      if (k & h) continue;
      const tw2 w = tw_at(tw, ml * row_factor + gb + (uint32_t)(k / (2 * h)));
      ct_planned<A, H, G0 + S>(v[k], v[k + h], w, m, neg);
    }
    reg_stages_fwd<A, H, G0, S + 1>(v, tw, ml_scale, row_factor, g_base, m, neg);
  }
}

// Resolve launch index y (tower-major): destination, source (== destination
// when in place or when this is not the first pass) and prime.
__device__ __forceinline__ void tower_of(const TowerMap &m, uint32_t y, bool first_pass, uint64_t *&data,
                                         const uint64_t *&src, uint32_t &prime, uint32_t &k, uint32_t &poly) {
  y += m.y_offset;
  k = m.polys_mul ? (uint32_t)(((uint64_t)y * m.polys_mul) >> 32) : y / m.polys;
  poly = y - k * m.polys;
  uint64_t off, src_off;
  if (m.jobs) {
    const TowerJob j = m.jobs[k];
    off = j.offset;
    src_off = j.src_offset;
    prime = j.prime;
  } else {
    off = src_off = (uint64_t)k * m.tower_stride;
    prime = m.prime_offset + k;
  }
  data = m.base + (uint64_t)poly * m.poly_stride + off;
  if (first_pass && m.src_table) src = m.src_table[poly] + src_off;
  else src = (first_pass && m.src) ? m.src + (uint64_t)poly * m.src_poly_stride + src_off : data;
}

// Epilogue destinations of tower (k, poly): the subtrahend and the output;
// acc says whether the output accumulates (in-place key switch) or is written.
__device__ __forceinline__ uint64_t *epi_dest(const TowerMap &m, uint32_t k, uint32_t poly, const uint64_t *&sub,
                                              bool &acc) {
  sub = m.epi.sub + (uint64_t)poly * m.epi.sub_poly_stride + (uint64_t)k * m.tower_stride;
  const bool second = poly >= m.epi.split;
  acc = (m.epi.accumulate >> (second ? 1 : 0)) & 1;
  if (m.epi.out_table) return m.epi.out_table[poly] + (uint64_t)k * m.tower_stride;
  return m.epi.out[second ? 1 : 0] + (uint64_t)(second ? poly - m.epi.split : poly) * m.epi.out_poly_stride +
         (uint64_t)k * m.tower_stride;
}

// ---------------------------------------------------------------- packed 32-bit hand-off
// Towers below 2^30 at N = 2^16 hand their values from one pass to the next
// as 32-bit words packed into the first half of the tower: element (row r,
// column c) at 32-bit index packed32(r, c), so the 128-byte segment of row
// r < 128 and column block c / 16 holds rows r and r + 128 of those 16
// columns. Every segment is written whole by one column-pass workgroup (its
// own 16 columns, read by it before) and by one row-pass wave (the pair of
// rows r, r + 128, which it reads before writing), so the packing stays in
// place without races, and each pass moves 256 KiB of hand-off per tower
// instead of 512 KiB.
__device__ __forceinline__ uint32_t packed32(uint32_t r, uint32_t c) {
  return ((r & 127u) << 9) | ((c >> 4) << 5) | ((r >> 7) << 4) | (c & 15u);
}

// ---------------------------------------------------------------- pass A, R = 256
// Inverse column-pass stages 12 + S .. 14 (P1 layout, distances 1, 2, 4;
// uniform twiddles through scalar loads).
template <int A, int H, int S>
__device__ __forceinline__ void a256_inv_tail(Val<A> (&v)[16], TwPtr tw, const Lazy8 &m, const PrimeConst &pc) {
  if constexpr (S < 3) {
    constexpr int h = 1 << S;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      if (k & h) continue;
      const tw2 w = tw_at(tw, (8u / h) + (uint32_t)(k / (2 * h)));
      gs_planned<A, H, 12 + S>(v[k], v[k + h], w, m, pc);
    }
    a256_inv_tail<A, H, S + 1>(v, tw, m, pc);
  }
}

// Column block bx (16 columns of 256) of one tower: data in place, src the
// forward input (== data unless the first pass reads elsewhere).
// Pk (32-bit towers only): the hand-off to or from the row pass is packed
// (packed32) instead of one 64-bit word per residue.
template <bool Inverse, int A, int H, bool Pk = false>
__device__ __forceinline__ void a256_body(uint64_t *data, const uint64_t *src, const PrimeConst &pc, uint32_t bx,
                                          Val<A> *lds, tw2 *twl) {
  static_assert(!Pk || A == kNarrow, "the packed hand-off holds 32-bit residues");
  const TwPtr tw = tw_table(Inverse ? pc.inv : pc.fwd);
  const Lazy8 m = lazy8_of(pc.q, pc.nq, pc.nq2, pc.nq4);
  const int t = threadIdx.x, cl = t & 15, u = t >> 4;
  const uint32_t c = bx * 16 + cl;  // column
  const TwLds twp = (TwLds)(const void *)twl;
  // The four lane-dependent stages (twiddles 16..255, 4 distinct per wave)
  // read the table from LDS: one 16-byte load per thread instead of 15 per
  // thread through the vector cache (the uniform stages keep scalar loads).
  const tw2 mine = tw_at(tw, (uint32_t)t);
  Val<A> v[16];
  if constexpr (!Inverse) {
#pragma unroll
    for (int k = 0; k < 16; ++k) v[k] = src[(uint64_t)(u + 16 * k) * kRow + c];
    twl[t] = mine;  // ordered before its readers by the exchange's barriers
    const NegMultiples neg = neg_multiples(m);
    reg_stages_fwd<A, H, 0>(v, tw, 1, 1, 0, m, neg);
    if constexpr (Pk) {
      // second layout with the row groups u2 and u2 + 8 in one wave (u2 = wave
      // + 4 x lane group), so that each packed store writes both halves of its
      // 128-byte segments (rows r and r + 128) at once, whole lines
      const int u2 = (t >> 6) + 4 * ((t >> 4) & 3);
      xchg_cols_full(v, lds, cl, u, u2);
      reg_stages_fwd<A, H, 4>(v, twp, 16, 1, u2, m, neg);
      uint32_t *h = reinterpret_cast<uint32_t *>(data);
#pragma unroll
      for (int k = 0; k < 16; ++k) h[packed32(16 * u2 + k, c)] = (uint32_t)v[k];
    } else {
      xchg_cols_p1_p2(v, lds, cl, u);
      reg_stages_fwd<A, H, 4>(v, twp, 16, 1, u, m, neg);
#pragma unroll
      for (int k = 0; k < 16; ++k) data[(uint64_t)(16 * u + k) * kRow + c] = v[k];
    }
  } else {
    if constexpr (Pk) {
      const uint32_t *h = reinterpret_cast<const uint32_t *>(data);
#pragma unroll
      for (int k = 0; k < 16; ++k) v[k] = (Val<A>)h[packed32(16 * u + k, c)];
    } else {
#pragma unroll
      for (int k = 0; k < 16; ++k) v[k] = data[(uint64_t)(16 * u + k) * kRow + c];
    }
    twl[t] = mine;
    __syncthreads();
    reg_stages_inv<A, H, 8>(v, twp, 16, 1, u, m, pc);
    xchg_cols_p2_p1(v, lds, cl, u);
    // transform stages 12-14 (h = 1, 2, 4; ml = 8, 4, 2), then the scaled
    // last stage (ml = 1)
    a256_inv_tail<A, H, 0>(v, tw, m, pc);
#pragma unroll
    for (int k = 0; k < 8; ++k) gs_last_planned<A, H>(v[k], v[k + 8], pc, m);
#pragma unroll
    for (int k = 0; k < 16; ++k) data[(uint64_t)(u + 16 * k) * kRow + c] = v[k];
  }
}

template <bool Inverse, int A, int H = 0, bool Pk = false>
__global__ __launch_bounds__(256) void ntt_pass_a256(TowerMap map, const PrimeConst *__restrict__ primes) {
  __shared__ Val<A> lds[(Pk && !Inverse) ? 16 * kLdsRow : 16 * kHalfRow];
  __shared__ tw2 twl[256];
  uint64_t *data;
  const uint64_t *src;
  uint32_t prime, tk, tpoly;
  tower_of(map, blockIdx.y, !Inverse, data, src, prime, tk, tpoly);
  const PrimeConst pc = primes[prime];
  a256_body<Inverse, A, H, Pk>(data, src, pc, blockIdx.x, lds, twl);
}

// ---------------------------------------------------------------- pass B (rows of 256)
// R = rows in the tower (N / 256); row_factor for ml is R + r.
// Epi (forward only): the ModDown correction epilogue, applied after the
// exchange back to the coalesced layout (measured faster than loading the
// subtrahend up front at the uncoalesced positions).
template <bool Inverse, int A, bool Epi>
__global__ __launch_bounds__(256) void ntt_pass_b(TowerMap map, const PrimeConst *__restrict__ primes, uint32_t R) {
  __shared__ Val<A> lds[8 * kLdsRow];
  uint64_t *data;
  const uint64_t *src;
  uint32_t prime, tk, tpoly;
  tower_of(map, blockIdx.y, Inverse, data, src, prime, tk, tpoly);
  const PrimeConst pc = primes[prime];
  const TwPtr tw = tw_table(Inverse ? pc.inv : pc.fwd);
  const Lazy8 m = lazy8_of(pc.q, pc.nq, pc.nq2, pc.nq4);
  const int t = threadIdx.x, u = t & 15, rl = t >> 4;
  const uint32_t r = blockIdx.x * 16 + rl;
  uint64_t *row = data + (uint64_t)r * kRow;
  const uint64_t *src_row = src + (uint64_t)r * kRow;
  const uint32_t rf = R + r;
  Val<A> v[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) v[k] = src_row[u + 16 * k];
  if constexpr (Epi) {
    reg_stages<false, A>(v, tw, 1, rf, 0, m);
    xchg_rows<true>(v, lds, rl, u);
    reg_stages<false, A>(v, tw, 16, rf, u, m);
    xchg_rows<false>(v, lds, rl, u);
    const uint64_t *sub;
    bool acc;
    uint64_t *out = epi_dest(map, tk, tpoly, sub, acc) + (uint64_t)r * kRow;
    sub += (uint64_t)r * kRow;
    const ulonglong2 c = map.epi.scale[tk];
    if (acc) {
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        const uint64_t d = epi_diff<A>(sub[u + 16 * k], v[k], m, pc.q);
        out[u + 16 * k] = add_mod(out[u + 16 * k], shoup(d, c.x, c.y, pc.q), pc.q);
      }
      return;
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const uint64_t d = epi_diff<A>(sub[u + 16 * k], v[k], m, pc.q);
      out[u + 16 * k] = shoup(d, c.x, c.y, pc.q);
    }
    return;
  } else if constexpr (!Inverse) {
    reg_stages<false, A>(v, tw, 1, rf, 0, m);
    xchg_rows<true>(v, lds, rl, u);
    reg_stages<false, A>(v, tw, 16, rf, u, m);
#pragma unroll
    for (int k = 0; k < 16; ++k) v[k] = canon<A>(v[k], m);
    xchg_rows<false>(v, lds, rl, u);
  } else {
    xchg_rows<true>(v, lds, rl, u);
    reg_stages<true, A>(v, tw, 16, rf, u, m);
    xchg_rows<false>(v, lds, rl, u);
    reg_stages<true, A>(v, tw, 1, rf, 0, m);
  }
#pragma unroll
  for (int k = 0; k < 16; ++k) row[u + 16 * k] = v[k];
}

// ---------------------------------------------------------------- row pass, 8 residues per thread
// The row pass of wide-lazy (q < 2^61) and 32-bit (q < 2^30) towers at
// R = 256: 32 threads per 256-element row, 8 rows per workgroup and 2 per
// wave, so every exchange stays inside one wave and needs no workgroup
// barrier (LDS operations of one wave complete in issue order; the wavefront
// fences below only keep the compiler from reordering them). With 8 residues
// a 64-bit thread needs 56-67 VGPRs, 7-8 waves per SIMD, against 98 VGPRs and
// 4 waves for 16 residues.
//
// Layouts of row element j (0..255) over lane u (0..31 within the row) and
// register slot k (0..7):
//   W  j = 64 (k >> 1) + 2u + (k & 1)         slot bits 2, 1, 0 <-> j bits 7, 6, 0
//   M  j = (u >> 3) << 6 | k << 3 | (u & 7)   slot bits <-> j bits 5, 4, 3
//   F  j = u << 3 | k                         slot bits <-> j bits 2, 1, 0
// Global loads and stores use W: each lane moves two adjacent residues as one
// 16-byte access (a wave instruction covers 512 bytes of each of its two
// rows), half the memory instructions of one residue per lane. Forward:
// stages of distance 128, 64 in W, 32, 16, 8 in M, 4, 2, 1 in F, back to W
// for the store; the inverse runs the mirror (F, M, W) and stores from W.
// Twiddle of the butterfly whose upper element is j: forward stage s
// (distance 128 >> s) uses T[(rf << s) + (j >> (8 - s))], inverse stage e
// (distance 2^e) T^-1[(rf << (7 - e)) + (j >> (e + 1))], rf = R + row.
__device__ __forceinline__ void wave_sync8() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
__device__ __forceinline__ int pad16(int j) { return j + (j >> 4); }

enum { kLayW = 0, kLayM = 1, kLayF = 2 };

// j = b8_lane<Lay>(u) | b8_slot<Lay>(k): the lane and slot parts are disjoint
// bit fields, so j >> s = (lane part >> s) + (slot part >> s), the second a
// compile-time constant (a twiddle's lane-dependent index is computed once
// per stage and its slot offset lands in the load's immediate).
template <int Lay>
__device__ __forceinline__ int b8_lane(int u) {
  u &= 31;
  if constexpr (Lay == kLayW) return u << 1;
  else if constexpr (Lay == kLayM) return ((u >> 3) << 6) | (u & 7);
  else return u << 3;
}
template <int Lay>
__host__ __device__ constexpr int b8_slot(int k) {
  return Lay == kLayW ? ((k >> 1) << 6) | (k & 1) : Lay == kLayM ? k << 3 : k;
}
template <int Lay>
__device__ __forceinline__ int b8_pos(int u, int k) { return b8_lane<Lay>(u) | b8_slot<Lay>(k); }

// Slot distance of forward stage s (its layout: W for s < 2, M for s < 5, F).
template <int S>
struct B8Fwd {
  static constexpr int lay = S < 2 ? kLayW : S < 5 ? kLayM : kLayF;
  static constexpr int h = S < 2 ? 4 >> S : S < 5 ? 4 >> (S - 2) : 4 >> (S - 5);
};
// Slot distance of inverse stage e (F for e < 3, M for e < 6, W).
template <int E>
struct B8Inv {
  static constexpr int lay = E < 3 ? kLayF : E < 6 ? kLayM : kLayW;
  static constexpr int h = E < 3 ? 1 << E : E < 6 ? 1 << (E - 3) : 1 << (E - 5);
};

// W-layout row accesses: slots 2a and 2a + 1 of a lane are adjacent words,
// which the compiler merges into one 16-byte access (global memory takes
// accesses aligned to 8 bytes on this target).
template <int A>
__device__ __forceinline__ void b8_load_w(Val<A> (&v)[8], const uint64_t *row, int u) {
#pragma unroll
  for (int k = 0; k < 8; ++k) v[k] = (Val<A>)row[b8_pos<kLayW>(u, k)];
}

__device__ __forceinline__ void b8_store_w(uint64_t *row, int u, const uint64_t (&o)[8]) {
#pragma unroll
  for (int k = 0; k < 8; ++k) row[b8_pos<kLayW>(u, k)] = o[k];
}

// Row r of the packed hand-off in layout W (tower base `data`): slots 2a and
// 2a + 1 are adjacent words of one 16-column block.
__device__ __forceinline__ void b8_load_packed(uint32_t (&v)[8], const uint64_t *data, uint32_t r, int u) {
  const uint32_t *h = reinterpret_cast<const uint32_t *>(data);
#pragma unroll
  for (int k = 0; k < 8; ++k) v[k] = h[packed32(r, (uint32_t)b8_pos<kLayW>(u, k))];
}
__device__ __forceinline__ void b8_store_packed(uint64_t *data, uint32_t r, int u, const uint32_t (&v)[8]) {
  uint32_t *h = reinterpret_cast<uint32_t *>(data);
#pragma unroll
  for (int k = 0; k < 8; ++k) h[packed32(r, (uint32_t)b8_pos<kLayW>(u, k))] = v[k];
}

// Row of thread-row rl (0..7) in workgroup bx of a row pass: rows 8 bx .. 8 bx
// + 7, or with the packed hand-off the row pairs (4 bx + w, 4 bx + w + 128) of
// its four waves w (a wave's 64 lanes then cover both halves of the packed
// segments of row 4 bx + w).
template <bool Pk>
__device__ __forceinline__ uint32_t b8_row_of(uint32_t bx, int rl) {
  if constexpr (Pk) return bx * 4 + (uint32_t)(rl >> 1) + ((uint32_t)(rl & 1) << 7);
  else return bx * 8 + (uint32_t)rl;
}

// Move v from layout From to layout To through the row's LDS buffer.
template <int From, int To, class V>
__device__ __forceinline__ void b8_xchg(V (&v)[8], V *buf, int u) {
#pragma unroll
  for (int k = 0; k < 8; ++k) buf[pad16(b8_pos<From>(u, k))] = v[k];
  wave_sync8();
#pragma unroll
  for (int k = 0; k < 8; ++k) v[k] = buf[pad16(b8_pos<To>(u, k))];
  wave_sync8();
}

// A row's twiddles of its first kB8Staged forward stages (the last ones of the
// inverse) are shared by 2 to 32 lanes each: they are staged in 32 LDS
// entries per row, entry 2^s - 1 + i holding T[(rf << s) + i] (stage s,
// 2^s entries), loaded by lanes 0-30 of the row and read back by all. The
// later stages' twiddles are distinct per lane and stay global loads.
// Measured against staging four stages (16 entries, 8 rather than 7
// workgroups per CU by LDS): inverse 59-bit transform 3% faster, the rest
// within 1.5%.
constexpr int kB8Staged = 5;
__device__ __forceinline__ void b8_stage_tw(tw2 *trow, TwPtr tw, uint32_t rf, int u) {
  if (u < (1 << kB8Staged) - 1) {
    const uint32_t s = 31 - __clz(u + 1), i = (uint32_t)(u + 1) - (1u << s);
    trow[u] = tw_at(tw, (rf << s) + i);
  }
}

// Forward row stages S..SEnd-1 (row stage s is transform stage 8 + s).
template <int A, int H, int S, int SEnd>
__device__ __forceinline__ void b8_fwd_stages(Val<A> (&v)[8], TwPtr tw, TwLds trow, uint32_t rf, int u, const Lazy8 &m,
                                              const NegMultiples &neg) {
  if constexpr (S < SEnd) {
    constexpr int h = B8Fwd<S>::h, lay = B8Fwd<S>::lay;
    const uint32_t lane = (uint32_t)(b8_lane<lay>(u) >> (8 - S));
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      if (k & h) continue;
      const uint32_t off = lane + (uint32_t)(b8_slot<lay>(k) >> (8 - S));
      const tw2 w = S < kB8Staged ? tw_at(trow, (1u << S) - 1 + off) : tw_at(tw, (rf << S) + off);
      ct_planned<A, H, 8 + S>(v[k], v[k + h], w, m, neg);
    }
    b8_fwd_stages<A, H, S + 1, SEnd>(v, tw, trow, rf, u, m, neg);
  }
}

// Inverse row stages E..EEnd-1 (distance 2^e).
template <int A, int H, int E, int EEnd>
__device__ __forceinline__ void b8_inv_stages(Val<A> (&v)[8], TwPtr tw, TwLds trow, uint32_t rf, int u, const Lazy8 &m,
                                              const PrimeConst &pc) {
  if constexpr (E < EEnd) {
    constexpr int h = B8Inv<E>::h, lay = B8Inv<E>::lay;
    const uint32_t lane = (uint32_t)(b8_lane<lay>(u) >> (E + 1));
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      if (k & h) continue;
      const uint32_t off = lane + (uint32_t)(b8_slot<lay>(k) >> (E + 1));
      const tw2 w = 7 - E < kB8Staged ? tw_at(trow, (1u << (7 - E)) - 1 + off) : tw_at(tw, (rf << (7 - E)) + off);
      gs_planned<A, H, E>(v[k], v[k + h], w, m, pc);
    }
    b8_inv_stages<A, H, E + 1, EEnd>(v, tw, trow, rf, u, m, pc);
  }
}

// One row (row index rf - R) of a workgroup: loads in W and all 8 forward
// stages; v ends in layout F.
template <int A, int H, bool Pk = false>
__device__ __forceinline__ void b8_fwd_core(Val<A> (&v)[8], const uint64_t *row, TwPtr tw, uint32_t rf, int u,
                                            const Lazy8 &m, const NegMultiples &neg, Val<A> *buf, tw2 *trow,
                                            const uint64_t *tower = nullptr, uint32_t r = 0) {
  b8_stage_tw(trow, tw, rf, u);
  if constexpr (Pk) b8_load_packed(v, tower, r, u);
  else b8_load_w<A>(v, row, u);
  wave_sync8();
  const TwLds tws = (TwLds)(const void *)trow;
  b8_fwd_stages<A, H, 0, 2>(v, tw, tws, rf, u, m, neg);
  b8_xchg<kLayW, kLayM>(v, buf, u);
  b8_fwd_stages<A, H, 2, 5>(v, tw, tws, rf, u, m, neg);
  b8_xchg<kLayM, kLayF>(v, buf, u);
  b8_fwd_stages<A, H, 5, 8>(v, tw, tws, rf, u, m, neg);
}

// Canonical output of b8_fwd_core, back to W for the store.
template <int A, int H>
__device__ __forceinline__ void b8_fwd_store(Val<A> (&v)[8], uint64_t *row, int u, const Lazy8 &m,
                                             const NegMultiples &neg, const PrimeConst &pc, Val<A> *buf) {
#pragma unroll
  for (int k = 0; k < 8; ++k) buf[pad16(b8_pos<kLayF>(u, k))] = canon_planned<A, H>(v[k], m, neg, pc);
  wave_sync8();
  uint64_t o[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) o[k] = buf[pad16(b8_pos<kLayW>(u, k))];
  b8_store_w(row, u, o);
}

template <int A, bool Epi, int H = 0, bool Pk = false>
__global__ __launch_bounds__(256) void ntt_pass_b8_fwd(TowerMap map, const PrimeConst *__restrict__ primes, uint32_t R) {
  static_assert(!Pk || (A == kNarrow && !Epi), "the packed hand-off: 32-bit towers without an epilogue");
  __shared__ Val<A> lds[8 * kLdsRow];
  __shared__ tw2 tws[8 * (1 << kB8Staged)];
  uint64_t *data;
  const uint64_t *src;
  uint32_t prime, tk, tpoly;
  tower_of(map, blockIdx.y, false, data, src, prime, tk, tpoly);
  const PrimeConst pc = primes[prime];
  const TwPtr tw = tw_table(pc.fwd);
  const Lazy8 m = lazy8_of(pc.q, pc.nq, pc.nq2, pc.nq4);
  const int t = threadIdx.x, u = t & 31, rl = t >> 5;
  const uint32_t r = b8_row_of<Pk>(blockIdx.x, rl);
  uint64_t *row = data + (uint64_t)r * kRow;
  const uint32_t rf = R + r;
  Val<A> *buf = lds + rl * kLdsRow;
  Val<A> v[8];
  const NegMultiples neg = neg_multiples(m);
  b8_fwd_core<A, H, Pk>(v, row, tw, rf, u, m, neg, buf, tws + rl * (1 << kB8Staged), data, r);
  if constexpr (Epi) {
    b8_xchg<kLayF, kLayW>(v, buf, u);
    const uint64_t *sub;
    bool acc;
    uint64_t *out = epi_dest(map, tk, tpoly, sub, acc) + (uint64_t)r * kRow;
    sub += (uint64_t)r * kRow;
    const ulonglong2 c = map.epi.scale[tk];
    uint64_t s[8], o[8];
    b8_load_w<kLazy>(s, sub, u);  // 64-bit words whatever the arithmetic
    if (acc) b8_load_w<kLazy>(o, out, u);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const uint64_t d = epi_diff<A>(s[k], below8_planned<A, H>(v[k], m, neg, pc), m, pc.q);
      const uint64_t y = shoup(d, c.x, c.y, pc.q);
      o[k] = acc ? add_mod(o[k], y, pc.q) : y;
    }
    b8_store_w(out, u, o);
  } else {
    b8_fwd_store<A, H>(v, row, u, m, neg, pc, buf);
  }
}

// ModUp's forward row pass fused with the key inner product (ModUpMacRows):
// a workgroup takes 8 rows of tower t of one polynomial and runs, digit by
// digit, the row pass of the digit's complement tower in registers (or loads
// the owning digit's input rows), multiplies by both channels' key rows and
// adds into 128-bit sums. The complement towers' finished transforms are
// never written out and read back, and the sums are written once.
// Wide: 128-bit sums, one Montgomery reduction per output (digits_count * q
// < 2^64); otherwise each product is reduced to [0, 2q) (mont_half) into a
// 64-bit sum below 2^(sum_log + 1) q, made canonical by sum_log + 1
// conditional subtracts (2^(sum_log + 1) q <= 2^64): half the accumulator
// registers for about twice the multiply instructions. Measured and not
// taken (both spill at 4 waves per SIMD): sums of two digits' products per
// reduction (100+ registers spilled) and the quotient estimate for the
// final reduction (4 spilled, 0.6% slower).
template <int H, bool Wide>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void ntt_pass_b8_fwd_mac(
    ModUpMacRows mm, const PrimeConst *__restrict__ primes) {
  __shared__ uint64_t lds[8 * kLdsRow];
  __shared__ tw2 tws[8 * (1 << kB8Staged)];
  const uint32_t k = blockIdx.y / mm.polys, poly = blockIdx.y - k * mm.polys;
  const uint32_t tower = mm.first_tower + k;
  const PrimeConst pc = primes[mm.tower_prime[tower]];
  const TwPtr tw = tw_table(pc.fwd);
  const Lazy8 m = lazy8_of(pc.q, pc.nq, pc.nq2, pc.nq4);
  const NegMultiples neg = neg_multiples(m);
  const int t = threadIdx.x, u = t & 31, rl = t >> 5;
  const uint32_t r = blockIdx.x * 8 + (uint32_t)rl;
  const uint64_t at = ((uint64_t)tower << 16) + (uint64_t)r * kRow;  // row r of tower t (N = 2^16)
  uint64_t *buf = lds + rl * kLdsRow;
  tw2 *trow = tws + rl * (1 << kB8Staged);
  const uint32_t own_digit = tower < mm.active_q ? tower / mm.alpha : 0xffffffffu;
  const uint64_t *own = (mm.own_table ? mm.own_table[poly] : mm.own + (uint64_t)poly * mm.own_poly_stride) + at;
  const uint64_t *dg = mm.digits + (uint64_t)poly * mm.digit_poly_stride + at;
  const uint64_t step = (uint64_t)mm.towers << 16;  // digit stride of the digit buffer and the key
  U128 sb[8], sa[8];
  uint64_t cb[8], ca[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    sb[j] = sa[j] = U128{0, 0};
    cb[j] = ca[j] = 0;
  }
#pragma unroll 1
  for (uint32_t d = 0; d < mm.digits_count; ++d) {
    uint64_t x[8];
    if (d == own_digit) {
      b8_load_w<kLazy>(x, own, u);
    } else {
      b8_fwd_core<kLazy, H>(x, dg + d * step, tw, kRow + r, u, m, neg, buf, trow);
      // 128-bit sums need canonical values (digits_count * q * q < q * 2^64);
      // a half-reduced product is below x K / 2^64 + q < 2q for any 64-bit
      // x and K < q, so the 64-bit sums take the row pass's lazy outputs
#pragma unroll
      for (int j = 0; j < 8; ++j)
        buf[pad16(b8_pos<kLayF>(u, j))] = Wide ? canon_planned<kLazy, H>(x[j], m, neg, pc) : x[j];
      wave_sync8();
#pragma unroll
      for (int j = 0; j < 8; ++j) x[j] = buf[pad16(b8_pos<kLayW>(u, j))];
      wave_sync8();  // the next digit's exchanges rewrite buf
    }
    uint64_t kb[8], ka[8];
    b8_load_w<kLazy>(kb, mm.key_b + d * step + at, u);
    b8_load_w<kLazy>(ka, mm.key_a + d * step + at, u);
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      if constexpr (Wide) {
        add_wide(sb[j], mul_wide63(x[j], kb[j]));
        add_wide(sa[j], mul_wide63(x[j], ka[j]));
      } else {
        cb[j] += mont_half(mul_wide63(x[j], kb[j]), pc.q, pc.qinv_neg);
        ca[j] += mont_half(mul_wide63(x[j], ka[j]), pc.q, pc.qinv_neg);
      }
    }
  }
  uint64_t ob[8], oa[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    if constexpr (Wide) {
      ob[j] = mont_reduce(sb[j], pc.q, pc.qinv_neg);
      oa[j] = mont_reduce(sa[j], pc.q, pc.qinv_neg);
    } else {
      ob[j] = cb[j];
      oa[j] = ca[j];
    }
  }
  if constexpr (!Wide) {
    // nq << k = -(2^k q) mod 2^64: subtract 2^k q, k = sum_log .. 0, where it fits
    for (int k = (int)mm.sum_log; k >= 0; --k) {
      const uint64_t nc = m.nq << k;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        ob[j] = sub_if_ge_neg(ob[j], nc);
        oa[j] = sub_if_ge_neg(oa[j], nc);
      }
    }
  }
  const uint64_t o = (uint64_t)poly * mm.out_poly_stride + at;
  b8_store_w(mm.out_b + o, u, ob);
  b8_store_w(mm.out_a + o, u, oa);
}

// Inverse row pass, 8 residues per thread (first pass of an inverse
// transform): load in W, stages of distance 1, 2, 4 in F, 8, 16, 32 in M,
// 64, 128 in W, store from W. One row per 32 threads (row 8 bx + thread /
// 32): reads src, writes data.
template <int A, int H, bool Pk = false>
__device__ __forceinline__ void b8_inv_body(uint64_t *data, const uint64_t *src, const PrimeConst &pc, uint32_t R,
                                            uint32_t bx, Val<A> *lds, tw2 *tws) {
  static_assert(!Pk || A == kNarrow, "the packed hand-off holds 32-bit residues");
  const TwPtr tw = tw_table(pc.inv);
  const Lazy8 m = lazy8_of(pc.q, pc.nq, pc.nq2, pc.nq4);
  const int t = threadIdx.x, u = t & 31, rl = t >> 5;
  const uint32_t r = b8_row_of<Pk>(bx, rl);
  uint64_t *row = data + (uint64_t)r * kRow;
  const uint64_t *src_row = src + (uint64_t)r * kRow;
  const uint32_t rf = R + r;
  Val<A> *buf = lds + rl * kLdsRow;
  tw2 *trow = tws + rl * (1 << kB8Staged);
  Val<A> v[8];
  b8_stage_tw(trow, tw, rf, u);  // read back after the exchanges' wave syncs
  b8_load_w<A>(v, src_row, u);
  b8_xchg<kLayW, kLayF>(v, buf, u);
  const TwLds twr = (TwLds)(const void *)trow;
  b8_inv_stages<A, H, 0, 3>(v, tw, twr, rf, u, m, pc);
  b8_xchg<kLayF, kLayM>(v, buf, u);
  b8_inv_stages<A, H, 3, 6>(v, tw, twr, rf, u, m, pc);
  b8_xchg<kLayM, kLayW>(v, buf, u);
  b8_inv_stages<A, H, 6, 8>(v, tw, twr, rf, u, m, pc);
  if constexpr (Pk) {
    // the wave has read rows r and r +- 128 (all the packed segments of its
    // pair) before any of its lanes writes them: the stores depend on every load
    b8_store_packed(data, r, u, v);
  } else {
    uint64_t o[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) o[k] = v[k];
    b8_store_w(row, u, o);
  }
}

template <int A, int H = 0, bool Pk = false>
__global__ __launch_bounds__(256) void ntt_pass_b8_inv(TowerMap map, const PrimeConst *__restrict__ primes, uint32_t R) {
  __shared__ Val<A> lds[8 * kLdsRow];
  __shared__ tw2 tws[8 * (1 << kB8Staged)];
  uint64_t *data;
  const uint64_t *src;
  uint32_t prime, tk, tpoly;
  tower_of(map, blockIdx.y, true, data, src, prime, tk, tpoly);
  const PrimeConst pc = primes[prime];
  b8_inv_body<A, H, Pk>(data, src, pc, R, blockIdx.x, lds, tws);
}

// ---------------------------------------------------------------- generic LDS kernel
// Runs `stages` radix-2 stages on `sub` independent sub-transforms of length
// L = 2^log_len held in LDS: element j of sub-transform b is at
// data[b_origin(b) + j * stride]. Used for whole small transforms
// (stride 1, one sub-transform) and for pass A when R < 256 (stride 256,
// 4096/R columns per workgroup). Forward runs ml = 1..L/2, inverse the mirror
// with n^-1 folded into the last stage when `last` is set.
template <bool Inverse, int A>
__device__ __forceinline__ void lds_generic_stages(const uint64_t *src, const PrimeConst &pc, uint32_t log_len,
                                                   uint32_t stride, uint32_t cols_per_block, uint32_t col0,
                                                   bool finish, Val<A> *lds) {
  const TwPtr tw = tw_table(Inverse ? pc.inv : pc.fwd);
  const Lazy8 m = lazy8_of(pc.q, pc.nq, pc.nq2, pc.nq4);
  const uint32_t L = 1u << log_len;
  const uint32_t total = L * cols_per_block;  // <= 4096
  // LDS layout: lds[col * L + j]; global: data[(col0 + col) + j * stride] (stride > 1)
  // or data[j] (stride == 1, cols_per_block == 1).
  for (uint32_t e = threadIdx.x; e < total; e += blockDim.x) {
    const uint32_t col = e % cols_per_block, j = e / cols_per_block;
    lds[col * L + j] = src[(uint64_t)j * stride + col0 + col];
  }
  __syncthreads();
  const uint32_t half = total / 2;
  for (uint32_t s = 0; s < log_len; ++s) {
    const uint32_t lg_d = Inverse ? s : log_len - 1 - s;  // distance 2^lg_d
    const uint32_t ml = L >> (lg_d + 1);
    const bool last = Inverse && finish && s == log_len - 1;
    for (uint32_t p = threadIdx.x; p < half; p += blockDim.x) {
      const uint32_t col = p / (L / 2), pp = p % (L / 2);
      const uint32_t g = pp >> lg_d;
      const uint32_t j = pp + (g << lg_d);
      Val<A> x = lds[col * L + j], y = lds[col * L + j + (1u << lg_d)];
      if constexpr (Inverse) {
        if (last) gs_bfly_last<A>(x, y, pc, m);
        else gs_bfly<A>(x, y, tw_at(tw, ml + g), m);
      } else {
        ct_bfly<A>(x, y, tw_at(tw, ml + g), m);
      }
      lds[col * L + j] = x;
      lds[col * L + j + (1u << lg_d)] = y;
    }
    __syncthreads();
  }
}

template <bool Inverse, int A>
__global__ __launch_bounds__(256) void ntt_lds_generic(TowerMap map, const PrimeConst *__restrict__ primes,
                                                       uint32_t log_len, uint32_t stride, uint32_t cols_per_block,
                                                       uint32_t finish, uint32_t first_pass) {
  __shared__ Val<A> lds[4096];
  uint64_t *data;
  const uint64_t *src;
  uint32_t prime, tk, tpoly;
  tower_of(map, blockIdx.y, first_pass != 0, data, src, prime, tk, tpoly);
  const PrimeConst pc = primes[prime];
  const Lazy8 m = lazy8_of(pc.q, pc.nq, pc.nq2, pc.nq4);
  const uint32_t L = 1u << log_len;
  const uint32_t total = L * cols_per_block;
  const uint32_t col0 = blockIdx.x * cols_per_block;
  lds_generic_stages<Inverse, A>(src, pc, log_len, stride, cols_per_block, col0, finish != 0, lds);
  const bool epilogue = !Inverse && finish && map.epi.sub;
  const uint64_t *sub = nullptr;
  uint64_t *out = data;
  ulonglong2 c{0, 0};
  bool acc = false;
  if (epilogue) {
    out = epi_dest(map, tk, tpoly, sub, acc);
    c = map.epi.scale[tk];
  }
  for (uint32_t e = threadIdx.x; e < total; e += blockDim.x) {
    const uint32_t col = e % cols_per_block, j = e / cols_per_block;
    const uint64_t at = (uint64_t)j * stride + col0 + col;
    Val<A> xv = lds[col * L + j];
    if (!Inverse && finish) xv = canon<A>(xv, m);
    uint64_t x = xv;
    if (epilogue) {
      x = shoup(sub_mod(sub[at], x, pc.q), c.x, c.y, pc.q);
      if (acc) x = add_mod(out[at], x, pc.q);
    }
    out[at] = x;
  }
}

// ---------------------------------------------------------------- mixed ring dimensions
// Work items are 4096-residue blocks (pass A: 16 columns of 256 at N = 2^16,
// 4096/R columns of R below) or half blocks (pass B: 8 rows of 256). The job
// of block b is found by a binary search over the jobs' prefix counts, read
// through the scalar cache (the item index is uniform in a workgroup).
__device__ __forceinline__ uint32_t mixed_job_of(const MixedJob *jobs, uint32_t count, uint32_t block) {
  typedef const __attribute__((address_space(4))) MixedJob *CJob;
  const CJob cj = (CJob)(const void *)jobs;
  uint32_t lo = 0, hi = count;  // jobs[lo].first <= block < (lo + 1 < count ? jobs[lo + 1].first : blocks)
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (cj[mid].first <= block) lo = mid;
    else hi = mid;
  }
  return lo;
}

template <bool Inverse, int A>
__global__ __launch_bounds__(256) void ntt_mixed_pass_a(const MixedJob *jobs, uint32_t count, uint32_t items) {
  __shared__ Val<A> lds[4096];  // the generic column pass's block; the 2^16 one needs 16 * kHalfRow
  __shared__ tw2 twl[256];       // the 2^16 column pass's staged twiddles
  typedef const __attribute__((address_space(4))) MixedJob *CJob;
  const CJob cj = (CJob)(const void *)jobs;
  for (uint32_t it = blockIdx.x; it < items; it += gridDim.x) {
    const uint32_t j = mixed_job_of(jobs, count, it);
    uint64_t *data = cj[j].data;
    const uint64_t *src = Inverse ? data : cj[j].src;  // forward: first pass
    const uint32_t log_n = cj[j].log_n, bx = it - cj[j].first;
    const PrimeConst pc = *cj[j].prime;
    if (log_n == 16) {
      a256_body<Inverse, A, 0>(data, src, pc, bx, lds, twl);
    } else {
      const uint32_t log_r = log_n - 8, cols = 4096u >> log_r, col0 = bx * cols, L = 1u << log_r;
      lds_generic_stages<Inverse, A>(src, pc, log_r, kRow, cols, col0, Inverse, lds);
      for (uint32_t e = threadIdx.x; e < 4096u; e += blockDim.x) {
        const uint32_t col = e % cols, jr = e / cols;
        data[(uint64_t)jr * kRow + col0 + col] = lds[col * L + jr];
      }
    }
    __syncthreads();  // the next item reuses the LDS
  }
}

template <bool Inverse, int A>
__global__ __launch_bounds__(256) void ntt_mixed_pass_b(const MixedJob *jobs, uint32_t count, uint32_t items) {
  __shared__ Val<A> lds[8 * kLdsRow];
  __shared__ tw2 tws[8 * (1 << kB8Staged)];
  typedef const __attribute__((address_space(4))) MixedJob *CJob;
  const CJob cj = (CJob)(const void *)jobs;
  for (uint32_t it = blockIdx.x; it < items; it += gridDim.x) {
    const uint32_t j = mixed_job_of(jobs, count, it >> 1);  // two row blocks per 4096-residue block
    uint64_t *data = cj[j].data;
    const uint32_t R = 1u << (cj[j].log_n - 8), bx = it - 2 * cj[j].first;
    const PrimeConst pc = *cj[j].prime;
    if constexpr (Inverse) {
      b8_inv_body<A, 0>(data, cj[j].src, pc, R, bx, lds, tws);  // inverse: first pass
    } else {
      const TwPtr tw = tw_table(pc.fwd);
      const Lazy8 m = lazy8_of(pc.q, pc.nq, pc.nq2, pc.nq4);
      const NegMultiples neg = neg_multiples(m);
      const int u = threadIdx.x & 31, rl = threadIdx.x >> 5;
      const uint32_t r = bx * 8 + rl;
      uint64_t *row = data + (uint64_t)r * kRow;
      Val<A> *buf = lds + rl * kLdsRow;
      Val<A> v[8];
      b8_fwd_core<A, 0>(v, row, tw, R + r, u, m, neg, buf, tws + rl * (1 << kB8Staged));
      b8_fwd_store<A, 0>(v, row, u, m, neg, pc, buf);
    }
    wave_sync8();  // a wave's exchange rows are its own; order them against its next item
  }
}

template <int A>
static hipError_t launch_mixed_t(const MixedJob *jobs, uint32_t count, uint32_t blocks, bool inverse,
                                 hipStream_t stream) {
  int device = 0, cus = 0;
  hipError_t e = hipGetDevice(&device);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
  if (e != hipSuccess) return e;
  // persistent grid: a few workgroups per CU (both passes hold at least 5 per CU)
  const uint32_t slots = (uint32_t)(cus > 0 ? cus : 1) * 5u;
  auto pass_a = [&]() {
    hipLaunchKernelGGL((ntt_mixed_pass_a<true, A>), dim3(blocks < slots ? blocks : slots), dim3(256), 0, stream, jobs,
                       count, blocks);
  };
  if (!inverse) {
    hipLaunchKernelGGL((ntt_mixed_pass_a<false, A>), dim3(blocks < slots ? blocks : slots), dim3(256), 0, stream,
                       jobs, count, blocks);
    hipLaunchKernelGGL((ntt_mixed_pass_b<false, A>), dim3(2 * blocks < slots ? 2 * blocks : slots), dim3(256), 0,
                       stream, jobs, count, 2 * blocks);
  } else {
    hipLaunchKernelGGL((ntt_mixed_pass_b<true, A>), dim3(2 * blocks < slots ? 2 * blocks : slots), dim3(256), 0,
                       stream, jobs, count, 2 * blocks);
    pass_a();
  }
  return hipGetLastError();
}

hipError_t launch_ntt_mixed(const MixedJob *jobs, uint32_t count, uint32_t blocks, bool inverse, int arith,
                            hipStream_t stream) {
  if (count == 0 || blocks == 0) return hipSuccess;
  switch (arith & kArithMask) {
    case kLazy: return launch_mixed_t<kLazy>(jobs, count, blocks, inverse, stream);
    case kNarrow: return launch_mixed_t<kNarrow>(jobs, count, blocks, inverse, stream);
    default: return hipErrorInvalidValue;
  }
}

// ---------------------------------------------------------------- launcher
// Extra dynamic LDS per workgroup that leaves room for at most `cap`
// workgroups per CU (160 KiB of LDS; 0 = no cap).
static uint32_t lds_pad_for(uint32_t cap) {
  if (cap == 0) return 0;
  constexpr uint32_t kLdsPerCu = 160 * 1024, kStatic = 17536 + 4096;  // largest static LDS of the 2^16 passes
  const uint32_t per_wg = kLdsPerCu / (cap + 1) + 1;
  return per_wg > kStatic ? per_wg - kStatic : 0;
}

template <bool Inverse, int A, int H = 0>
static hipError_t launch_t(const TowerMap &map, const PrimeConst *primes, uint32_t log_n, uint32_t towers,
                           hipStream_t stream, bool row16 = false, uint32_t pad = 0) {
  const uint32_t n = 1u << log_n;
  if (log_n <= 12) {
    hipLaunchKernelGGL((ntt_lds_generic<Inverse, A>), dim3(1, towers), dim3(256), 0, stream, map, primes, log_n,
                       1u, 1u, 1u, 1u);
    return hipGetLastError();
  }
  const uint32_t R = n / kRow, log_r = log_n - 8;
  // 32-bit towers at N = 2^16 with the 8-residue row pass hand their values
  // over packed (packed32), unless the forward's row pass carries the ModDown
  // epilogue; both passes of the launch agree on the form.
  const bool packed = A == kNarrow && R == 256 && !row16 && (Inverse || !map.epi.sub);
  auto pass_a = [&]() {
    if (R == 256) {
      if constexpr (A == kNarrow) {
        if (packed) {
          hipLaunchKernelGGL((ntt_pass_a256<Inverse, A, H, true>), dim3(kRow / 16, towers), dim3(256), pad, stream,
                             map, primes);
          return;
        }
      }
      hipLaunchKernelGGL((ntt_pass_a256<Inverse, A, H>), dim3(kRow / 16, towers), dim3(256), pad,
                         stream, map, primes);
    } else {
      const uint32_t cols = 4096 / R;
      hipLaunchKernelGGL((ntt_lds_generic<Inverse, A>), dim3(kRow / cols, towers), dim3(256), 0, stream, map,
                         primes, log_r, (uint32_t)kRow, cols, (uint32_t)Inverse, (uint32_t)!Inverse);
    }
  };
  const uint32_t pb = R == 256 ? pad : 0;
  auto pass_b = [&]() {
    if constexpr (Inverse && (A == kLazy || A == kNarrow)) {
      if (R == 256 && !row16) {
        if constexpr (A == kNarrow) {
          if (packed) {
            hipLaunchKernelGGL((ntt_pass_b8_inv<A, H, true>), dim3(R / 8, towers), dim3(256), pb, stream, map, primes,
                               R);
            return;
          }
        }
        hipLaunchKernelGGL((ntt_pass_b8_inv<A, H>), dim3(R / 8, towers), dim3(256), pb, stream, map, primes, R);
        return;
      }
    }
    if constexpr (!Inverse && (A == kLazy || A == kNarrow)) {
      if (R == 256 && !row16) {
        if constexpr (A == kNarrow) {
          if (packed) {
            hipLaunchKernelGGL((ntt_pass_b8_fwd<A, false, H, true>), dim3(R / 8, towers), dim3(256), pb, stream, map,
                               primes, R);
            return;
          }
        }
        if (map.epi.sub)
          hipLaunchKernelGGL((ntt_pass_b8_fwd<A, true, H>), dim3(R / 8, towers), dim3(256), pb, stream, map, primes,
                             R);
        else
          hipLaunchKernelGGL((ntt_pass_b8_fwd<A, false, H>), dim3(R / 8, towers), dim3(256), pb, stream, map, primes,
                             R);
        return;
      }
    }
    // the 16-residue row pass (R < 256, canonical arithmetic, or chosen by
    // the tuner) keeps one reduction per stage: its forward pairs with H = 0
    if (!Inverse && map.epi.sub)
      hipLaunchKernelGGL((ntt_pass_b<false, A, true>), dim3(R / 16, towers), dim3(256), pb, stream, map, primes, R);
    else
      hipLaunchKernelGGL((ntt_pass_b<Inverse, A, false>), dim3(R / 16, towers), dim3(256), pb, stream, map, primes,
                         R);
  };
  if (!Inverse) {
    pass_a();
    pass_b();
  } else {
    pass_b();
    pass_a();
  }
  return hipGetLastError();
}

// One launch of `count` launch indices in the arithmetic `arith` (mode and,
// for kLazy, the forward headroom class).
static hipError_t dispatch(const TowerMap &map, const PrimeConst *primes, uint32_t log_n, uint32_t count, bool inverse,
                           int arith, hipStream_t stream) {
  // tower_of's y / polys as a multiply: exact for y * polys <= 2^32 (the
  // reciprocal's excess e = mul polys - 2^32 < polys keeps y e / 2^32 below
  // the distance of y / polys to the next integer)
  TowerMap m = map;
  const uint64_t y_end = (uint64_t)m.y_offset + count;
  m.polys_mul = m.polys && y_end * m.polys <= (uint64_t{1} << 32) ? ((uint64_t{1} << 32) + m.polys - 1) / m.polys : 0;
  const int mode = arith & kArithMask, h = (arith >> kArithShift) & 7;
  const bool row16 = (arith & kRow16Bit) != 0;
  const uint32_t pad = lds_pad_for((uint32_t)(arith >> kWgCapShift) & 15u);
  if (row16 && h != 0) return hipErrorInvalidValue;
  if (mode == kLazy) {
    // both directions take the class H of the launch's widest modulus: the
    // forward and inverse reduction schedules (make_fwd_plan, make_inv_plan)
    if (inverse) {
      switch (h) {
        case 0: return launch_t<true, kLazy, 0>(m, primes, log_n, count, stream, row16, pad);
        case 1: return launch_t<true, kLazy, 1>(m, primes, log_n, count, stream, row16, pad);
        case 2: return launch_t<true, kLazy, 2>(m, primes, log_n, count, stream, row16, pad);
        case 3: return launch_t<true, kLazy, 3>(m, primes, log_n, count, stream, row16, pad);
        case 4: return launch_t<true, kLazy, 4>(m, primes, log_n, count, stream, row16, pad);
        case 5: return launch_t<true, kLazy, 5>(m, primes, log_n, count, stream, row16, pad);
        default: return hipErrorInvalidValue;
      }
    }
    switch (h) {
      case 0: return launch_t<false, kLazy, 0>(m, primes, log_n, count, stream, row16, pad);
      case 1: return launch_t<false, kLazy, 1>(m, primes, log_n, count, stream, row16, pad);
      case 2: return launch_t<false, kLazy, 2>(m, primes, log_n, count, stream, row16, pad);
      case 3: return launch_t<false, kLazy, 3>(m, primes, log_n, count, stream, row16, pad);
      case 4: return launch_t<false, kLazy, 4>(m, primes, log_n, count, stream, row16, pad);
      case 5: return launch_t<false, kLazy, 5>(m, primes, log_n, count, stream, row16, pad);
      default: return hipErrorInvalidValue;
    }
  }
  if (h != 0) return hipErrorInvalidValue;
  switch (mode * 2 + (inverse ? 1 : 0)) {
    case kNarrow * 2: return launch_t<false, kNarrow>(m, primes, log_n, count, stream, row16, pad);
    case kNarrow * 2 + 1: return launch_t<true, kNarrow>(m, primes, log_n, count, stream, row16, pad);
    case kExact * 2: return launch_t<false, kExact>(m, primes, log_n, count, stream, row16, pad);
    case kExact * 2 + 1: return launch_t<true, kExact>(m, primes, log_n, count, stream, row16, pad);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_ntt_span(const TowerMap &map, const PrimeConst *primes, uint32_t log_n, uint32_t polys, uint32_t y0,
                           uint32_t count, bool inverse, int arith, hipStream_t stream) {
  if (count == 0) return hipSuccess;
  if (polys == 0) return hipErrorInvalidValue;
  TowerMap m = map;
  m.polys = polys;
  m.y_offset = map.y_offset + y0;
  return dispatch(m, primes, log_n, count, inverse, arith, stream);
}

hipError_t launch_ntt_columns_span(const TowerMap &map, const PrimeConst *primes, uint32_t polys, uint32_t y0,
                                   uint32_t count, int arith, hipStream_t stream) {
  if (count == 0) return hipSuccess;
  if (polys == 0 || (arith & kArithMask) != kLazy || (arith & kRow16Bit)) return hipErrorInvalidValue;
  TowerMap m = map;
  m.polys = polys;
  m.y_offset = map.y_offset + y0;
  const uint64_t y_end = (uint64_t)m.y_offset + count;
  m.polys_mul = y_end * m.polys <= (uint64_t{1} << 32) ? ((uint64_t{1} << 32) + m.polys - 1) / m.polys : 0;
  const uint32_t pad = lds_pad_for((uint32_t)(arith >> kWgCapShift) & 15u);
  const dim3 grid(kRow / 16, count);
  switch ((arith >> kArithShift) & 7) {
    case 0: hipLaunchKernelGGL((ntt_pass_a256<false, kLazy, 0>), grid, dim3(256), pad, stream, m, primes); break;
    case 1: hipLaunchKernelGGL((ntt_pass_a256<false, kLazy, 1>), grid, dim3(256), pad, stream, m, primes); break;
    case 2: hipLaunchKernelGGL((ntt_pass_a256<false, kLazy, 2>), grid, dim3(256), pad, stream, m, primes); break;
    case 3: hipLaunchKernelGGL((ntt_pass_a256<false, kLazy, 3>), grid, dim3(256), pad, stream, m, primes); break;
    case 4: hipLaunchKernelGGL((ntt_pass_a256<false, kLazy, 4>), grid, dim3(256), pad, stream, m, primes); break;
    case 5: hipLaunchKernelGGL((ntt_pass_a256<false, kLazy, 5>), grid, dim3(256), pad, stream, m, primes); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_modup_mac_rows(const ModUpMacRows &mm, const PrimeConst *primes, int arith, hipStream_t stream) {
  if (mm.launch_towers == 0) return hipSuccess;
  if (mm.polys == 0 || mm.digits_count == 0 || mm.alpha == 0 || (arith & kArithMask) != kLazy ||
      (arith & kRow16Bit) || (uint64_t)mm.launch_towers * mm.polys > 65535 ||
      mm.first_tower + mm.launch_towers > mm.towers ||
      (!mm.wide && (mm.sum_log > 6 || mm.digits_count > (1u << mm.sum_log))))
    return hipErrorInvalidValue;
  const dim3 grid(kRow / 8, mm.launch_towers * mm.polys);
  const int h = (arith >> kArithShift) & 7;
  if (h > kMaxHeadroom) return hipErrorInvalidValue;
  typedef void (*Kernel)(ModUpMacRows, const PrimeConst *);
  static constexpr Kernel wide[] = {ntt_pass_b8_fwd_mac<0, true>, ntt_pass_b8_fwd_mac<1, true>,
                                    ntt_pass_b8_fwd_mac<2, true>, ntt_pass_b8_fwd_mac<3, true>,
                                    ntt_pass_b8_fwd_mac<4, true>, ntt_pass_b8_fwd_mac<5, true>};
  static constexpr Kernel half[] = {ntt_pass_b8_fwd_mac<0, false>, ntt_pass_b8_fwd_mac<1, false>,
                                    ntt_pass_b8_fwd_mac<2, false>, ntt_pass_b8_fwd_mac<3, false>,
                                    ntt_pass_b8_fwd_mac<4, false>, ntt_pass_b8_fwd_mac<5, false>};
  hipLaunchKernelGGL((mm.wide ? wide : half)[h], grid, dim3(256), 0, stream, mm, primes);
  return hipGetLastError();
}

// `chunk` > 0 splits the towers into groups run pass after pass (A then B for
// one group before the next group), so a group's intermediate hand-off stays
// in L2 / Infinity Cache instead of a full HBM round trip.
hipError_t launch_ntt(const TowerMap &map, const PrimeConst *primes, uint32_t log_n, uint32_t towers, bool inverse,
                      int arith, hipStream_t stream, uint32_t chunk) {
  if (towers == 0) return hipSuccess;
  if (map.towers_per_poly == 0 || towers % map.towers_per_poly != 0) return hipErrorInvalidValue;
  const uint32_t step = chunk ? chunk : towers;
  for (uint32_t first = 0; first < towers; first += step) {
    TowerMap m = map;
    m.polys = towers / map.towers_per_poly;
    m.y_offset = map.y_offset + first;
    const uint32_t count = towers - first < step ? towers - first : step;
    const hipError_t e = dispatch(m, primes, log_n, count, inverse, arith, stream);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace nttamd
