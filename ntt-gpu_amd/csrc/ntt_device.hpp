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
  const ulonglong2 *fwd;
  const ulonglong2 *inv;
};

// Explicit tower: data at base + offset (elements), constants primes[prime].
struct TowerJob {
  uint64_t offset, src_offset;
  uint32_t prime;
  uint32_t pad;
};

// Which towers a launch touches. blockIdx.y = y selects either jobs[y] or the
// affine map: poly = y / towers_per_poly, t = y % towers_per_poly,
// data = base + poly * poly_stride + t * tower_stride, prime = prime_offset + t.
// `src` (optional) makes the first pass out of place: it reads tower y at
// src + poly * src_poly_stride + t * tower_stride (or src + jobs[y].src_offset).
struct TowerMap {
  uint64_t *base;
  const uint64_t *src;
  const TowerJob *jobs;
  uint64_t poly_stride, tower_stride, src_poly_stride;
  uint32_t towers_per_poly, prime_offset;
  uint32_t y_offset;  // tower index of blockIdx.y == 0 (chunked launches)
};

hipError_t launch_ntt(const TowerMap &map, const PrimeConst *primes, uint32_t log_n, uint32_t towers, bool inverse,
                      bool lazy, hipStream_t stream, uint32_t chunk = 0);

}  // namespace nttamd
