#pragma once
// Test double of the small part of the OpenFHE 1.6.1 API that
// include/ntt/openfhe_hybrid_adapter.hpp calls (test infrastructure only; it
// is not OpenFHE and holds no OpenFHE code). Elements are plain vectors of
// residues; SetFormat converts with the C oracle's negacyclic transform
// (OpenFHE's evaluation layout: bit-reversed order, points psi^(2 rev(k)+1)),
// and FindAutomorphismIndex returns 5^i mod 2N (5^-1 for negative i, 2N - 1
// kept for conjugation), the CKKS rotation convention.
#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "support/oracle_api.h"

enum Format { EVALUATION = 0, COEFFICIENT = 1 };

namespace lbcrypto {

class NativeInteger {
  uint64_t v_ = 0;

 public:
  NativeInteger(uint64_t v = 0) : v_(v) {}
  uint64_t ConvertToInt() const { return v_; }
  bool operator==(const NativeInteger &o) const { return v_ == o.v_; }
};

class NativeVector {
  std::vector<NativeInteger> d_;
  NativeInteger mod_;

 public:
  NativeVector() = default;
  NativeVector(size_t n, const NativeInteger &m) : d_(n), mod_(m) {}
  size_t GetLength() const { return d_.size(); }
  NativeInteger &operator[](size_t i) { return d_.at(i); }
  const NativeInteger &operator[](size_t i) const { return d_.at(i); }
  const NativeInteger &GetModulus() const { return mod_; }
};

class ILNativeParams {
  uint32_t m_;
  NativeInteger q_, root_;

 public:
  ILNativeParams(uint32_t m, uint64_t q, uint64_t root) : m_(m), q_(q), root_(root) {}
  uint32_t GetCyclotomicOrder() const { return m_; }
  uint32_t GetRingDimension() const { return m_ / 2; }
  const NativeInteger &GetModulus() const { return q_; }
  const NativeInteger &GetRootOfUnity() const { return root_; }
};

class NativePoly {
  std::shared_ptr<ILNativeParams> params_;
  Format format_ = EVALUATION;
  NativeVector values_;
  bool empty_ = true;

 public:
  NativePoly() = default;
  NativePoly(const std::shared_ptr<ILNativeParams> &p, Format f) : params_(p), format_(f) {}
  const std::shared_ptr<ILNativeParams> &GetParams() const { return params_; }
  Format GetFormat() const { return format_; }
  const NativeInteger &GetModulus() const { return params_->GetModulus(); }
  const NativeInteger &GetRootOfUnity() const { return params_->GetRootOfUnity(); }
  bool IsEmpty() const { return empty_; }
  const NativeVector &GetValues() const { return values_; }
  void SetValues(NativeVector &&v, Format f) {
    values_ = std::move(v);
    format_ = f;
    empty_ = false;
  }
  // test helper: the residues as words
  std::vector<uint64_t> words() const {
    std::vector<uint64_t> w(values_.GetLength());
    for (size_t i = 0; i < w.size(); ++i) w[i] = values_[i].ConvertToInt();
    return w;
  }
  void set_words(const std::vector<uint64_t> &w, Format f) {
    NativeVector v(w.size(), params_->GetModulus());
    for (size_t i = 0; i < w.size(); ++i) v[i] = NativeInteger(w[i]);
    SetValues(std::move(v), f);
  }
};

class ILDCRTParams {
  uint32_t m_;
  std::vector<std::shared_ptr<ILNativeParams>> towers_;

 public:
  ILDCRTParams(uint32_t m, std::vector<std::shared_ptr<ILNativeParams>> t) : m_(m), towers_(std::move(t)) {}
  uint32_t GetCyclotomicOrder() const { return m_; }
  uint32_t GetRingDimension() const { return m_ / 2; }
  const std::vector<std::shared_ptr<ILNativeParams>> &GetParams() const { return towers_; }
};

class DCRTPoly {
  std::shared_ptr<ILDCRTParams> params_;
  Format format_ = EVALUATION;
  std::vector<NativePoly> towers_;

 public:
  using Params = ILDCRTParams;
  DCRTPoly() = default;
  DCRTPoly(const std::shared_ptr<Params> &p, Format f, bool zero = false) : params_(p), format_(f) {
    for (const auto &t : p->GetParams()) {
      towers_.emplace_back(t, f);
      if (zero) towers_.back().set_words(std::vector<uint64_t>(t->GetRingDimension(), 0), f);
    }
  }
  Format GetFormat() const { return format_; }
  const std::shared_ptr<Params> &GetParams() const { return params_; }
  const std::vector<NativePoly> &GetAllElements() const { return towers_; }
  const NativePoly &GetElementAtIndex(size_t i) const { return towers_.at(i); }
  void SetElementAtIndex(size_t i, NativePoly &&t) { towers_.at(i) = std::move(t); }
  void SetFormat(Format f) {
    if (f == format_) return;
    for (auto &t : towers_) {
      std::vector<uint64_t> w = t.words();
      uint64_t q = t.GetModulus().ConvertToInt(), psi = t.GetRootOfUnity().ConvertToInt();
      if (or_ntt_towers(w.data(), w.size(), 1, &q, &psi, f == COEFFICIENT ? 1 : 0))
        throw std::runtime_error("double: transform rejected");
      t.set_words(w, f);
    }
    format_ = f;
  }
  DCRTPoly &operator+=(const DCRTPoly &o) {
    if (o.format_ != format_ || o.towers_.size() != towers_.size()) throw std::runtime_error("double: add mismatch");
    for (size_t t = 0; t < towers_.size(); ++t) {
      std::vector<uint64_t> a = towers_[t].words();
      const std::vector<uint64_t> b = o.towers_[t].words();
      const uint64_t q = towers_[t].GetModulus().ConvertToInt();
      for (size_t i = 0; i < a.size(); ++i) a[i] = (a[i] + b[i]) % q;
      towers_[t].set_words(a, format_);
    }
    return *this;
  }
};

enum KeySwitchTechnique { BV, HYBRID };

class CryptoParametersBase {
 public:
  virtual ~CryptoParametersBase() = default;
};

class CryptoParametersCKKSRNS : public CryptoParametersBase {
 public:
  std::shared_ptr<ILDCRTParams> q, p;
  uint32_t per_part = 1, noise_scale = 1;
  KeySwitchTechnique technique = HYBRID;
  KeySwitchTechnique GetKeySwitchTechnique() const { return technique; }
  uint32_t GetNoiseScale() const { return noise_scale; }
  const std::shared_ptr<ILDCRTParams> &GetElementParams() const { return q; }
  const std::shared_ptr<ILDCRTParams> &GetParamsP() const { return p; }
  uint32_t GetNumPerPartQ() const { return per_part; }
};

template <class E>
class CryptoContextImpl;
template <class E>
using CryptoContext = std::shared_ptr<CryptoContextImpl<E>>;

template <class E>
class EvalKeyImpl {
 public:
  std::vector<E> b, a;
  CryptoContext<E> context;
  const std::vector<E> &GetBVector() const { return b; }
  const std::vector<E> &GetAVector() const { return a; }
  CryptoContext<E> GetCryptoContext() const { return context; }
};
template <class E>
using EvalKey = std::shared_ptr<EvalKeyImpl<E>>;

template <class E>
class CryptoContextImpl {
 public:
  std::shared_ptr<CryptoParametersBase> params;
  std::map<std::string, std::vector<EvalKey<E>>> mult_keys;
  std::map<std::string, std::shared_ptr<std::map<uint32_t, EvalKey<E>>>> rotation_keys;
  uint32_t cyclotomic_order = 0;
  const std::shared_ptr<CryptoParametersBase> &GetCryptoParameters() const { return params; }
  const std::vector<EvalKey<E>> &GetEvalMultKeyVector(const std::string &tag) const { return mult_keys.at(tag); }
  std::shared_ptr<std::map<uint32_t, EvalKey<E>>> GetEvalAutomorphismKeyMapPtr(const std::string &tag) const {
    return rotation_keys.at(tag);
  }
  uint32_t FindAutomorphismIndex(uint32_t index) const {
    const int32_t i = static_cast<int32_t>(index);
    const uint64_t m = cyclotomic_order;
    if (i == 0) return 1;
    if ((uint64_t)i == m - 1) return (uint32_t)i;
    uint64_t g = 5;
    if (i < 0) {  // 5^-1 mod m
      for (uint64_t x = 1; x < m; x += 2)
        if (x * 5 % m == 1) g = x;
    }
    uint64_t r = 1;
    for (int32_t k = 0; k < (i < 0 ? -i : i); ++k) r = r * g % m;
    return (uint32_t)r;
  }
};

template <class E>
class CiphertextImpl {
 public:
  std::vector<E> elements;
  CryptoContext<E> context;
  std::string tag = "key";
  const std::vector<E> &GetElements() const { return elements; }
  void SetElements(std::vector<E> &&e) { elements = std::move(e); }
  CryptoContext<E> GetCryptoContext() const { return context; }
  const std::string &GetKeyTag() const { return tag; }
  std::shared_ptr<CiphertextImpl<E>> CloneEmpty() const {
    auto c = std::make_shared<CiphertextImpl<E>>();
    c->context = context;
    c->tag = tag;
    return c;
  }
  std::shared_ptr<CiphertextImpl<E>> Clone() const { return std::make_shared<CiphertextImpl<E>>(*this); }
};
template <class E>
using Ciphertext = std::shared_ptr<CiphertextImpl<E>>;
template <class E>
using ConstCiphertext = const std::shared_ptr<const CiphertextImpl<E>>;

}  // namespace lbcrypto
