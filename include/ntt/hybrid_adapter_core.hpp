#pragma once
// OpenFHE-free parts of the Hybrid adapter (include/ntt/openfhe_hybrid_adapter.hpp):
// the layouts the adapter derives from a CKKS basis, the slot-shift
// normalisation of rotations, the rotation request plan, tower packing with
// validation, and the candidate builders that fix the adapter's commit order.
// They take plain values or callables, so the CPU tests exercise them without
// OpenFHE (reference behaviour: include/ntt/openfhe_hybrid_adapter.hpp:9-361).
#include <algorithm>
#include <cstdint>
#include <span>
#include <utility>
#include <vector>

#include "ntt/backend_contract.hpp"

namespace ntt::openfhe::core {
using proposal::Domain;
using proposal::Order;
using proposal::PolynomialLayout;
using proposal::require;
using proposal::Tower;

// The CKKS basis as the adapter reads it from a crypto context: ring
// dimension, the full Q chain and the special primes P, digit size alpha
// (OpenFHE's numPerPartQ) and the active level (towers of the input).
struct BasisDescription {
  uint64_t ring_dimension = 0;
  std::vector<Tower> q, p;
  size_t digit_size = 0;
  size_t active_q = 0;
};

// Evaluation-form Q prefix at the input's level: the layout of every
// ciphertext component the adapter packs (openfhe_hybrid_adapter.hpp:84-104).
inline PolynomialLayout source_layout(const BasisDescription &d) {
  require(d.ring_dimension >= 2 && d.ring_dimension <= 65536, "Hybrid adapter supports ring dimensions up to 65536");
  require(d.active_q >= 1 && d.active_q <= d.q.size() && d.active_q <= 40,
          "the input level must be 1..40 and within the Q chain");
  require(!d.p.empty() && d.p.size() <= 40, "the Hybrid basis needs 1..40 special primes");
  auto parameters = proposal::Parameters::create(d.ring_dimension, d.q, d.p, d.digit_size);
  return {parameters, d.active_q, proposal::q_prefix(*parameters, d.active_q), Domain::Evaluation,
          Order::BitReversed};
}

// The Session basis of one adapter transaction: active Q then P, coefficient form.
inline PolynomialLayout session_basis(const PolynomialLayout &source) {
  return {source.parameters, source.active_q, proposal::qp_basis(*source.parameters, source.active_q),
          Domain::Coefficient, Order::Natural};
}

// Evaluation keys cover the full chain: all Q towers, then P.
inline PolynomialLayout full_key_layout(const PolynomialLayout &source) {
  const size_t M = source.parameters->q().size();
  return {source.parameters, M, proposal::qp_basis(*source.parameters, M), Domain::Evaluation,
          Order::BitReversed};
}

inline size_t full_key_digits(const PolynomialLayout &source) {
  const size_t M = source.parameters->q().size(), alpha = source.parameters->alpha();
  return (M + alpha - 1) / alpha;
}

// A rotation by `slot_shift` slots of an N-dimensional CKKS ring: the
// generator 5 has order N/2, so shifts reduce mod N/2 (C++ remainder, sign
// kept); 2N - 1 is OpenFHE's reserved conjugation index and passes unchanged
// (openfhe_hybrid_adapter.hpp:183-189).
inline int32_t normalize_slot_shift(int32_t slot_shift, int32_t ring_dimension) {
  require(ring_dimension >= 4 && (ring_dimension & (ring_dimension - 1)) == 0, "invalid ring dimension");
  if (slot_shift == 2 * ring_dimension - 1) return slot_shift;
  return slot_shift % (ring_dimension / 2);
}

// Hoisted rotations: a zero shift is OpenFHE's EvalFastRotation identity (a
// clone of the input, no key); every other shift needs a key.
struct RotationRequest {
  int32_t shift = 0;
  bool identity = true;
};
inline std::vector<RotationRequest> plan_rotations(std::span<const int32_t> shifts) {
  std::vector<RotationRequest> out;
  out.reserve(shifts.size());
  for (const int32_t s : shifts) out.push_back({s, s == 0});
  return out;
}
inline bool all_identity(std::span<const RotationRequest> plan) {
  return std::all_of(plan.begin(), plan.end(), [](const RotationRequest &r) { return r.identity; });
}

// Pack `towers` towers of n residues, tower by tower, from read(t, i) with
// the layout's moduli checked (canonical residues only).
template <class Read>
std::vector<uint64_t> pack_towers(const PolynomialLayout &layout, Read &&read) {
  const size_t n = layout.parameters->n(), towers = layout.towers.size();
  std::vector<uint64_t> out(towers * n);
  for (size_t t = 0; t < towers; ++t) {
    const uint64_t q = layout.parameters->tower(layout.towers[t]).q;
    for (size_t i = 0; i < n; ++i) {
      const uint64_t v = read(t, i);
      require(v < q, "noncanonical residue in an OpenFHE tower");
      out[t * n + i] = v;
    }
  }
  return out;
}

// In-place key switch (openfhe_hybrid_adapter.hpp:221-248): elements are the
// ciphertext components (2 or 3), the last one key-switched into (b, a). The
// candidate keeps the leading components, brings them to evaluation form
// with to_eval, then c0 += b and either c1 += a (three components) or a as
// the new c1 (two components). The caller commits it in one step.
template <class Poly, class ToEval>
std::vector<Poly> key_switch_candidate(std::span<const Poly> elements, const Poly &b, const Poly &a, ToEval &&to_eval) {
  require(elements.size() == 2 || elements.size() == 3, "Hybrid key switch takes two or three components");
  const size_t leading = elements.size() - 1;
  std::vector<Poly> out(elements.begin(), elements.begin() + leading);
  to_eval(out[0]);
  out[0] += b;
  if (leading == 2) {
    to_eval(out[1]);
    out[1] += a;
  } else {
    out.push_back(a);
  }
  return out;
}

// Relinearization (openfhe_hybrid_adapter.hpp:250-286): c0 and c1 in
// evaluation form, then for each higher component j (in order) its
// contribution (b_j, a_j) = contribute(j) added to c0 and c1.
template <class Poly, class ToEval, class Contribute>
std::vector<Poly> relinearize_candidate(std::span<const Poly> elements, ToEval &&to_eval, Contribute &&contribute) {
  require(elements.size() >= 2, "relinearization needs at least two components");
  std::vector<Poly> out(elements.begin(), elements.begin() + 2);
  to_eval(out[0]);
  to_eval(out[1]);
  for (size_t j = 0; j + 2 < elements.size(); ++j) {
    std::pair<Poly, Poly> ba = contribute(j);
    out[0] += ba.first;
    out[1] += ba.second;
  }
  return out;
}

}  // namespace ntt::openfhe::core
