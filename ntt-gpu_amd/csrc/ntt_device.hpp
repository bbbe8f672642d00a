// Device-resident per-prime constants and tower addressing shared by all
// kernels of the library.
#pragma once
#include "modarith.hpp"

namespace nttamd {

// One distinct prime of a context (reference tables: resident_adapter.inc:68-81,
// transform_stage.hpp:19-30). fwd[k] = {psi^bitrev(k), Shoup}, inv[k] the same
// for psi^-1; k = 1..n-1 are used.
struct PrimeConst {
  uint64_t q, two_q, qinv_neg;       // modulus, 2q, -q^-1 mod 2^64
  uint64_t ninv, ninv_shoup;         // n^-1 mod q and its Shoup companion
  uint64_t last_w, last_wp;          // psi^-bitrev(1) * n^-1 and Shoup (inverse last stage)
  uint64_t r2;                       // 2^128 mod q (to enter Montgomery form)
  uint64_t nq, nq2, nq4;             // 2^64 - q, 2^64 - 2q, 2^64 - 4q (precomputed: see Lazy8)
  uint64_t mu_lo, mu_hi;             // floor(2^128 / q) (Barrett, variable x variable products)
  uint32_t est_shift, est_mul;       // b - 24 and floor(2^(b+31) / q), b = bit length of q >= 24 (reduce2_est)
  const ulonglong2 *fwd;
  const ulonglong2 *inv;
};

// Explicit tower of polynomial 0: data at base + offset (elements),
// constants primes[prime]; polynomial p adds p * poly_stride.
struct TowerJob {
  uint64_t offset, src_offset;
  uint32_t prime;
  uint32_t pad;
};

// Which towers a launch touches. Launch index y (blockIdx.y + y_offset) is
// taken tower-major, so the polynomials sharing a prime (and its twiddle
// table) run next to each other: k = y / polys, poly = y % polys, then either
// jobs[k] or the affine map t = k:
//   data = base + poly * poly_stride + (jobs[k].offset | t * tower_stride),
//   prime = jobs[k].prime | prime_offset + t.
// `src` (optional) makes the first pass out of place: it reads from
// src + poly * src_poly_stride + (jobs[k].src_offset | t * tower_stride).
// towers_per_poly is the number of jobs or affine towers; launch_ntt sets polys.
struct TowerMap {
  uint64_t *base;
  const uint64_t *src;
  const TowerJob *jobs;
  uint64_t poly_stride, tower_stride, src_poly_stride;
  uint32_t towers_per_poly, prime_offset;
  uint32_t y_offset;  // launch index of blockIdx.y == 0 (chunked launches)
  uint32_t polys;     // towers / towers_per_poly
  // ceil(2^32 / polys) when (y_offset + count) * polys <= 2^32 for the launch
  // (then y / polys = y * polys_mul >> 32 exactly), else 0 (divide); set by
  // the launcher
  uint64_t polys_mul;
  // When set, polynomial p's first-pass source is src_table[p] (+ the tower's
  // offset) instead of src + p * src_poly_stride (components at scattered
  // device addresses, e.g. the buffers of separate requests).
  const uint64_t *const *src_table;
  // Optional epilogue of a forward transform (ModDown correction, affine maps
  // only): the last pass writes (sub - NTT(x)) * scale[k] mod q to
  // out[poly < split ? 0 : 1] + (poly mod split) * out_poly_stride + k * tower_stride
  // instead of writing NTT(x) back; sub is read at
  // sub + poly * sub_poly_stride + k * tower_stride.
  struct Epilogue {
    const uint64_t *sub;
    uint64_t *out[2];
    const ulonglong2 *scale;  // [k] {c, Shoup(c)}
    uint64_t sub_poly_stride, out_poly_stride;
    uint32_t split;
    uint32_t accumulate;  // bit j: out[j] += result (mod q) instead of out[j] = result
    // When set, polynomial p's output is out_table[p] (+ k * tower_stride)
    // instead of out[p < split ? 0 : 1] + ...
    uint64_t *const *out_table;
  } epi;
};

// Arithmetic of a launch, chosen per context from its largest modulus. For
// kLazy the bits above kArithMask carry the headroom class H of the forward
// transform at N = 2^16: every modulus is below 2^(61 - H), so residues may
// grow to 2^(3 + H) q before they must be reduced (0 <= H <= kMaxHeadroom).
enum Arith : int { kExact = 0, kLazy = 1, kNarrow = 2 };
constexpr int kArithMask = 15, kArithShift = 4, kMaxHeadroom = 5;
inline int headroom_for(uint64_t max_q) {
  int bits = 0;
  while (bits < 64 && (max_q >> bits) != 0) ++bits;
  const int h = 61 - bits;
  return h < 0 ? 0 : h > kMaxHeadroom ? kMaxHeadroom : h;
}
inline int arith_for(uint64_t max_q) {
  if (max_q < (uint64_t{1} << 30)) return kNarrow;
  if (max_q < (uint64_t{1} << 61)) return kLazy | (headroom_for(max_q) << kArithShift);
  return kExact;
}

// Tuned launch shape, carried in the same int above the arithmetic bits:
//   kRow16Bit: the N = 2^16 row pass keeps 16 residues per thread (ntt_pass_b)
//              instead of 8 (ntt_pass_b8_*); the forward then uses headroom 0;
//   bits kWgCapShift..+3: at most that many workgroups per CU for the
//              N = 2^16 passes (0 = no cap), enforced with extra dynamic LDS.
constexpr int kRow16Bit = 1 << 8, kWgCapShift = 12;
inline int ntt_mode(int arith, bool row16, uint32_t wg_cap) {
  if (row16) arith = (arith & kArithMask) | kRow16Bit;  // the 16-residue row pass runs headroom 0
  return arith | (int)((wg_cap & 15u) << kWgCapShift);
}

hipError_t launch_ntt(const TowerMap &map, const PrimeConst *primes, uint32_t log_n, uint32_t towers, bool inverse,
                      int arith, hipStream_t stream, uint32_t chunk = 0);

// One tower of a launch that mixes ring dimensions N = 2^13 .. 2^16 (each job
// its own length and prime constants, e.g. towers of contexts of different N).
struct MixedJob {
  uint64_t *data;           // N residues, transformed in place
  const uint64_t *src;      // forward input (== data when in place)
  const PrimeConst *prime;  // device constants of the tower's prime (its context's table)
  uint32_t log_n;           // 13 .. 16
  uint32_t first;           // 4096-residue blocks of the jobs before this one
};
// Forward or inverse transforms of `count` jobs (device table) holding
// `blocks` 4096-residue blocks in all, in one persistent launch per pass:
// a fixed grid whose workgroups stride over (job, block) work items. Wide-lazy
// (kLazy) and 32-bit (kNarrow) towers, one arithmetic per launch.
hipError_t launch_ntt_mixed(const MixedJob *jobs, uint32_t count, uint32_t blocks, bool inverse, int arith,
                            hipStream_t stream);

// Launch indices [y0, y0 + count) of a map over `polys` polynomials (both
// passes): a range of the tower-major index space, e.g. the jobs of a few
// towers for every polynomial.
hipError_t launch_ntt_span(const TowerMap &map, const PrimeConst *primes, uint32_t log_n, uint32_t polys, uint32_t y0,
                           uint32_t count, bool inverse, int arith, hipStream_t stream);

// The column pass alone of forward transforms at N = 2^16 in the wide-lazy
// arithmetic (arith kLazy with its headroom class, 8-residue row pass shape),
// launch indices [y0, y0 + count) of `map` over `polys` polynomials: the
// towers are left in the hand-off form that launch_modup_mac_rows reads.
hipError_t launch_ntt_columns_span(const TowerMap &map, const PrimeConst *primes, uint32_t polys, uint32_t y0,
                                   uint32_t count, int arith, hipStream_t stream);

// ModUp's row pass fused with the key inner product, N = 2^16, wide-lazy
// towers. For towers [first_tower, first_tower + launch_towers) of `polys`
// polynomials: digit d's value of tower t is the row pass (in registers) of
// its complement tower in `digits`, whose column pass has run, or for the
// digit that owns t (t < active_q, d = t / alpha) the input tower itself;
//   out_b[p][t] = sum_d X_d[t] * key_b[d][t],  out_a likewise,
// keys in Montgomery form. wide: sums kept in 128 bits and reduced once (the
// caller guarantees digits_count * q < 2^64); otherwise each product is
// reduced to [0, 2q) into a 64-bit sum (digits_count <= 2^sum_log and
// 2^(sum_log + 1) q <= 2^64). The digit buffers are only read.
struct ModUpMacRows {
  const uint64_t *digits;  // (poly, digit, tower, i)
  const uint64_t *own;     // (poly, tower, i): evaluation-form input
  const uint64_t *const *own_table;  // when set: polynomial p's input towers at own_table[p]
  const uint64_t *key_b, *key_a;     // (digit, tower, i)
  uint64_t *out_b, *out_a;           // (poly, tower, i)
  const uint32_t *tower_prime;       // [towers]
  uint64_t own_poly_stride, digit_poly_stride, out_poly_stride;
  uint32_t towers, alpha, active_q, digits_count;
  uint32_t first_tower, launch_towers, polys;
  uint32_t wide, sum_log;
};
hipError_t launch_modup_mac_rows(const ModUpMacRows &m, const PrimeConst *primes, int arith, hipStream_t stream);

}  // namespace nttamd
