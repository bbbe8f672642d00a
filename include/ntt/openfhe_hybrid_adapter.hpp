#pragma once
// OpenFHE 1.6.1 bridge of the CKKS Hybrid key-switching core (the entry
// points of reference include/ntt/openfhe_hybrid_adapter.hpp:212-361) over the
// resident Session: hybrid_key_switch_core, hybrid_key_switch_in_place,
// hybrid_relinearize_in_place, hybrid_rotate and hybrid_rotate_many.
//
// Header-only; needs openfhe.h (OpenFHE with NATIVE_SIZE=64) only where it
// is included. `make_session` is any callable PolynomialLayout -> Session
// (ntt::resident::make_hip_session binds the gfx950 transport).
//
// Transactions: every input component and key is validated and packed
// before a session is created; results stay private until every download is
// validated and the session closed; a ciphertext is updated by one
// SetElements at the end, so on any exception the caller's ciphertext is
// unchanged. The OpenFHE-free decisions live in ntt/hybrid_adapter_core.hpp.
#include <map>
#include <memory>
#include <span>
#include <vector>

#include "ntt/hybrid_adapter_core.hpp"
#include "ntt/resident.hpp"
#include "openfhe.h"

namespace ntt::openfhe {

struct HybridContributions {
  lbcrypto::DCRTPoly b, a;
};

namespace bridge {
using lbcrypto::DCRTPoly;
using proposal::PolynomialLayout;
using proposal::require;

// (q, root) of every tower of an OpenFHE element-parameter set of ring dimension n.
inline std::vector<proposal::Tower> towers_of(const DCRTPoly::Params &params, size_t n) {
  require(params.GetCyclotomicOrder() == 2 * n, "OpenFHE parameters of another ring dimension");
  std::vector<proposal::Tower> out;
  for (const auto &t : params.GetParams()) {
    require(t && t->GetCyclotomicOrder() == 2 * n, "OpenFHE tower of another ring dimension");
    out.push_back({t->GetModulus().ConvertToInt(), t->GetRootOfUnity().ConvertToInt()});
  }
  return out;
}

// The adapter's view of `input` under `context`: CKKS, Hybrid key switching,
// reduced-noise scaling off, level = towers of the input.
inline PolynomialLayout describe(const DCRTPoly &input, const lbcrypto::CryptoContext<DCRTPoly> &context) {
  require(bool(context), "missing OpenFHE crypto context");
  const auto crypto = std::dynamic_pointer_cast<lbcrypto::CryptoParametersCKKSRNS>(context->GetCryptoParameters());
  require(crypto && crypto->GetKeySwitchTechnique() == lbcrypto::HYBRID && crypto->GetNoiseScale() == 1 &&
              crypto->GetElementParams() && crypto->GetParamsP() && input.GetParams(),
          "the adapter takes CKKS Hybrid parameters and a parameterised input");
  core::BasisDescription d;
  d.ring_dimension = crypto->GetElementParams()->GetRingDimension();
  d.q = towers_of(*crypto->GetElementParams(), d.ring_dimension);
  d.p = towers_of(*crypto->GetParamsP(), d.ring_dimension);
  d.digit_size = crypto->GetNumPerPartQ();
  d.active_q = input.GetParams()->GetParams().size();
  return core::source_layout(d);
}

// Check `poly` against `layout` (tower count, per-tower modulus and root,
// format, extents) and return its residues tower-major; `format` is the
// format the polynomial must be in.
inline std::vector<uint64_t> checked_residues(const DCRTPoly &poly, const PolynomialLayout &layout, Format format) {
  const size_t n = layout.parameters->n();
  require(poly.GetFormat() == format && poly.GetParams() && poly.GetAllElements().size() == layout.towers.size(),
          "OpenFHE polynomial is not complete in the expected format");
  const auto declared = towers_of(*poly.GetParams(), n);
  require(declared.size() == layout.towers.size(), "OpenFHE polynomial parameters of another level");
  for (size_t t = 0; t < layout.towers.size(); ++t) {
    const auto want = layout.parameters->tower(layout.towers[t]);
    const auto &tower = poly.GetElementAtIndex(t);
    require(declared[t] == want && tower.GetParams() && tower.GetParams()->GetCyclotomicOrder() == 2 * n &&
                tower.GetFormat() == format && tower.GetModulus().ConvertToInt() == want.q &&
                tower.GetRootOfUnity().ConvertToInt() == want.psi && !tower.IsEmpty() &&
                tower.GetValues().GetLength() == n && tower.GetValues().GetModulus().ConvertToInt() == want.q,
            "OpenFHE tower does not match the Hybrid basis");
  }
  return core::pack_towers(layout, [&](size_t t, size_t i) {
    return (uint64_t)poly.GetElementAtIndex(t).GetValues()[i].ConvertToInt();
  });
}

// Evaluation-form DCRTPoly over `params` from tower-major residues.
inline DCRTPoly to_dcrt(const std::shared_ptr<DCRTPoly::Params> &params, const PolynomialLayout &layout,
                        std::span<const uint64_t> residues) {
  proposal::validate_host_values(layout, residues);
  const size_t n = layout.parameters->n();
  DCRTPoly out(params, Format::EVALUATION, false);
  for (size_t t = 0; t < layout.towers.size(); ++t) {
    const auto &tp = params->GetParams()[t];
    lbcrypto::NativeVector values(n, tp->GetModulus());
    for (size_t i = 0; i < n; ++i) values[i] = lbcrypto::NativeInteger(residues[t * n + i]);
    lbcrypto::NativePoly tower(tp, Format::EVALUATION);
    tower.SetValues(std::move(values), Format::EVALUATION);
    out.SetElementAtIndex(t, std::move(tower));
  }
  return out;
}

// An evaluation key packed on the host: per full-chain digit, its b and a rows
// over the full QP basis (the Session picks the active rows).
struct PackedKey {
  PolynomialLayout layout;
  std::vector<std::vector<uint64_t>> b, a;

  PackedKey(const lbcrypto::EvalKey<DCRTPoly> &key, const PolynomialLayout &source)
      : layout(core::full_key_layout(source)) {
    require(bool(key), "missing OpenFHE evaluation key");
    const auto &kb = key->GetBVector();
    const auto &ka = key->GetAVector();
    const size_t digits = core::full_key_digits(source);
    require(kb.size() == digits && ka.size() == digits, "OpenFHE Hybrid key has the wrong digit count");
    for (size_t d = 0; d < digits; ++d) {
      b.push_back(checked_residues(kb[d], layout, Format::EVALUATION));
      a.push_back(checked_residues(ka[d], layout, Format::EVALUATION));
    }
  }
  resident::HybridKey upload(resident::Session &session) const {
    std::vector<resident::HybridKeyDigit> rows;
    for (size_t d = 0; d < b.size(); ++d) rows.push_back({b[d], a[d]});
    return session.upload_hybrid_key(layout, rows);
  }
};

// ModDown of both channels, download, validation; the A download is waited
// on last (its completion proves the earlier work complete too).
inline HybridContributions finish(resident::Session &session, const resident::HybridContributions &extended,
                                  const PolynomialLayout &source, const std::shared_ptr<DCRTPoly::Params> &params) {
  const resident::Polynomial b = session.mod_down(extended.b), a = session.mod_down(extended.a);
  proposal::require_layout(b.layout(), source);
  proposal::require_layout(a.layout(), source);
  const resident::Download db = session.download(b), da = session.download(a);
  const std::span<const uint64_t> va = da.wait_values();
  return {to_dcrt(params, source, db.wait_values()), to_dcrt(params, source, va)};
}

// One key switch (optionally followed by sigma_k(B + P c0), sigma_k(A) before
// ModDown) in a session of its own. `input` and the key are only read.
template <class MakeSession>
HybridContributions transact(const DCRTPoly &input, const lbcrypto::EvalKey<DCRTPoly> &key,
                             const PolynomialLayout &source, MakeSession &make_session,
                             std::span<const uint64_t> c0 = {}, uint32_t automorphism = 0) {
  const std::vector<uint64_t> evaluations = checked_residues(input, source, Format::EVALUATION);
  const PackedKey packed(key, source);
  resident::Session session = make_session(core::session_basis(source));
  const resident::HybridKey resident_key = packed.upload(session);
  const std::vector<resident::Polynomial> digits = session.hybrid_digits(session.upload(source, evaluations));
  resident::HybridContributions extended = session.hybrid_accumulate(digits, resident_key);
  if (!c0.empty()) extended = session.hybrid_rotate_extended(extended, session.upload(source, c0), automorphism);
  HybridContributions out = finish(session, extended, source, input.GetParams());
  session.close();
  return out;
}

inline uint32_t automorphism_of(const lbcrypto::CryptoContext<DCRTPoly> &context, int32_t slot_shift,
                                size_t ring_dimension) {
  const int32_t s = core::normalize_slot_shift(slot_shift, (int32_t)ring_dimension);
  return context->FindAutomorphismIndex(static_cast<uint32_t>(s));
}

inline lbcrypto::EvalKey<DCRTPoly> rotation_key(const std::map<uint32_t, lbcrypto::EvalKey<DCRTPoly>> &keys,
                                                uint32_t automorphism,
                                                const lbcrypto::CryptoContext<DCRTPoly> &context) {
  const auto it = keys.find(automorphism);
  require(it != keys.end() && it->second && it->second->GetCryptoContext() == context,
          "no rotation key of this context for the automorphism");
  return it->second;
}

inline lbcrypto::Ciphertext<DCRTPoly> as_ciphertext(lbcrypto::ConstCiphertext<DCRTPoly> &like,
                                                    HybridContributions parts) {
  std::vector<DCRTPoly> elements;
  elements.push_back(std::move(parts.b));
  elements.push_back(std::move(parts.a));
  auto out = like->CloneEmpty();
  out->SetElements(std::move(elements));
  return out;
}

inline void to_evaluation(DCRTPoly &p) { p.SetFormat(Format::EVALUATION); }
}  // namespace bridge

// KeySwitchHybrid core of one evaluation-form polynomial: (B, A) at the
// input's level (openfhe_hybrid_adapter.hpp:212-219).
template <class MakeSession>
HybridContributions hybrid_key_switch_core(const lbcrypto::DCRTPoly &input,
                                           const lbcrypto::EvalKey<lbcrypto::DCRTPoly> &key,
                                           MakeSession make_session) {
  proposal::require(bool(key), "missing OpenFHE evaluation key");
  const auto source = bridge::describe(input, key->GetCryptoContext());
  return bridge::transact(input, key, source, make_session);
}

// Key switch of the last component of a 2- or 3-component ciphertext, added
// into it in place (openfhe_hybrid_adapter.hpp:221-248).
template <class MakeSession>
void hybrid_key_switch_in_place(lbcrypto::Ciphertext<lbcrypto::DCRTPoly> &ciphertext,
                                const lbcrypto::EvalKey<lbcrypto::DCRTPoly> &key, MakeSession make_session) {
  proposal::require(ciphertext && ciphertext->GetCryptoContext() && key &&
                        ciphertext->GetCryptoContext() == key->GetCryptoContext(),
                    "ciphertext and evaluation key must share an OpenFHE context");
  const auto &elements = ciphertext->GetElements();
  proposal::require(elements.size() == 2 || elements.size() == 3, "Hybrid key switch takes two or three components");
  const auto source = bridge::describe(elements.back(), ciphertext->GetCryptoContext());
  for (size_t j = 0; j + 1 < elements.size(); ++j)
    (void)bridge::checked_residues(elements[j], source, elements[j].GetFormat());
  const HybridContributions c = bridge::transact(elements.back(), key, source, make_session);
  std::vector<lbcrypto::DCRTPoly> next =
      core::key_switch_candidate<lbcrypto::DCRTPoly>(elements, c.b, c.a, bridge::to_evaluation);
  ciphertext->SetElements(std::move(next));
}

// Relinearization of components 2.. with the context's EvalMult keys in
// their vector order (openfhe_hybrid_adapter.hpp:250-286).
template <class MakeSession>
void hybrid_relinearize_in_place(lbcrypto::Ciphertext<lbcrypto::DCRTPoly> &ciphertext, MakeSession make_session) {
  proposal::require(ciphertext && ciphertext->GetCryptoContext(), "missing OpenFHE ciphertext or context");
  const auto context = ciphertext->GetCryptoContext();
  const auto &elements = ciphertext->GetElements();
  proposal::require(elements.size() >= 2, "relinearization needs at least two components");
  const auto &keys = context->GetEvalMultKeyVector(ciphertext->GetKeyTag());
  const size_t extra = elements.size() - 2;
  proposal::require(keys.size() >= extra, "not enough OpenFHE relinearization keys");
  for (size_t j = 0; j < extra; ++j)
    proposal::require(keys[j] && keys[j]->GetCryptoContext() == context, "relinearization key of another context");
  const auto source = bridge::describe(elements.back(), context);
  for (const auto &e : elements) (void)bridge::checked_residues(e, source, e.GetFormat());
  std::vector<lbcrypto::DCRTPoly> next = core::relinearize_candidate<lbcrypto::DCRTPoly>(
      elements, bridge::to_evaluation, [&](size_t j) {
        lbcrypto::DCRTPoly component = elements[j + 2];
        bridge::to_evaluation(component);
        HybridContributions c = bridge::transact(component, keys[j], source, make_session);
        return std::make_pair(std::move(c.b), std::move(c.a));
      });
  ciphertext->SetElements(std::move(next));
}

// Rotation by `slot_shift` slots: OpenFHE applies the automorphism to the
// extended QP contributions before ModDown (openfhe_hybrid_adapter.hpp:288-303).
template <class MakeSession>
lbcrypto::Ciphertext<lbcrypto::DCRTPoly> hybrid_rotate(lbcrypto::ConstCiphertext<lbcrypto::DCRTPoly> &ciphertext,
                                                       int32_t slot_shift, MakeSession make_session) {
  proposal::require(ciphertext && ciphertext->GetCryptoContext(), "missing OpenFHE ciphertext or context");
  const auto context = ciphertext->GetCryptoContext();
  const auto &elements = ciphertext->GetElements();
  proposal::require(elements.size() == 2, "rotation takes a relinearized (two-component) ciphertext");
  const auto source = bridge::describe(elements[1], context);
  const uint32_t k = bridge::automorphism_of(context, slot_shift, source.parameters->n());
  const auto keys = context->GetEvalAutomorphismKeyMapPtr(ciphertext->GetKeyTag());
  const auto key = bridge::rotation_key(*keys, k, context);
  const std::vector<uint64_t> c0 = bridge::checked_residues(elements[0], source, Format::EVALUATION);
  return bridge::as_ciphertext(ciphertext, bridge::transact(elements[1], key, source, make_session, c0, k));
}

// Hoisted rotations: the digits of c1 are computed once in one session and
// shared by every shift; a zero shift returns a clone of the input
// (OpenFHE's EvalFastRotation identity), and all-zero shifts open no session
// (openfhe_hybrid_adapter.hpp:305-361).
template <class MakeSession>
std::vector<lbcrypto::Ciphertext<lbcrypto::DCRTPoly>> hybrid_rotate_many(
    lbcrypto::ConstCiphertext<lbcrypto::DCRTPoly> &ciphertext, std::span<const int32_t> slot_shifts,
    MakeSession make_session) {
  proposal::require(ciphertext && ciphertext->GetCryptoContext(), "missing OpenFHE ciphertext or context");
  const auto &elements = ciphertext->GetElements();
  proposal::require(elements.size() == 2, "rotation takes a relinearized (two-component) ciphertext");
  const std::vector<core::RotationRequest> plan = core::plan_rotations(slot_shifts);
  std::vector<lbcrypto::Ciphertext<lbcrypto::DCRTPoly>> out;
  if (core::all_identity(plan)) {
    for (size_t r = 0; r < plan.size(); ++r) out.push_back(ciphertext->Clone());
    return out;
  }
  const auto context = ciphertext->GetCryptoContext();
  const auto source = bridge::describe(elements[1], context);
  const auto keys = context->GetEvalAutomorphismKeyMapPtr(ciphertext->GetKeyTag());
  // every key is found and packed before the session opens
  std::vector<uint32_t> autos(plan.size(), 0);
  std::vector<std::unique_ptr<bridge::PackedKey>> packed(plan.size());
  for (size_t r = 0; r < plan.size(); ++r) {
    if (plan[r].identity) continue;
    autos[r] = bridge::automorphism_of(context, plan[r].shift, source.parameters->n());
    packed[r] = std::make_unique<bridge::PackedKey>(bridge::rotation_key(*keys, autos[r], context), source);
  }
  const std::vector<uint64_t> c0v = bridge::checked_residues(elements[0], source, Format::EVALUATION);
  const std::vector<uint64_t> c1v = bridge::checked_residues(elements[1], source, Format::EVALUATION);
  resident::Session session = make_session(core::session_basis(source));
  const resident::Polynomial c0 = session.upload(source, c0v);
  const std::vector<resident::Polynomial> digits = session.hybrid_digits(session.upload(source, c1v));
  for (size_t r = 0; r < plan.size(); ++r) {
    if (plan[r].identity) {
      out.push_back(ciphertext->Clone());
      continue;
    }
    const resident::HybridKey key = packed[r]->upload(session);
    const resident::HybridContributions ext = session.hybrid_accumulate(digits, key);
    const resident::HybridContributions rot = session.hybrid_rotate_extended(ext, c0, autos[r]);
    out.push_back(bridge::as_ciphertext(ciphertext, bridge::finish(session, rot, source, elements[0].GetParams())));
  }
  session.close();
  return out;
}

}  // namespace ntt::openfhe
