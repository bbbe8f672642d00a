#pragma once
// Vector files: the text format of the reference's CLI and oracle fixtures
// (reference include/ntt/vector_io.hpp:1-13, read_vector_file :50-93,
// reduce_decimal_mod :97-104). '#' lines are comments; then "N P", P lines
// "q psi", the word COEFFS and N non-negative decimal coefficients of any
// length (they live in [0, prod q), far beyond 64 bits).
#include <cstdint>
#include <fstream>
#include <istream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace ntt {

struct VectorFile {
  uint64_t n = 0;
  std::vector<uint64_t> q;
  std::vector<uint64_t> psi;
  std::vector<std::string> coeffs;
};

inline VectorFile parse_vector_file(std::istream &in) {
  std::vector<std::string> words;
  for (std::string line; std::getline(in, line);) {
    if (!line.empty() && line.front() == '#') continue;
    std::istringstream fields(line);
    for (std::string w; fields >> w;) words.push_back(std::move(w));
  }
  size_t at = 0;
  auto take = [&]() -> const std::string & {
    if (at == words.size()) throw std::runtime_error("vector file: truncated");
    return words[at++];
  };
  auto number = [&](const char *what) {
    const std::string &w = take();
    if (w.empty() || w.find_first_not_of("0123456789") != std::string::npos)
      throw std::runtime_error(std::string("vector file: bad ") + what);
    return std::stoull(w);
  };
  VectorFile f;
  f.n = number("N");
  const uint64_t primes = number("prime count");
  if (f.n < 2 || (f.n & (f.n - 1))) throw std::runtime_error("vector file: N must be a power of two >= 2");
  if (primes == 0) throw std::runtime_error("vector file: need >= 1 prime");
  for (uint64_t k = 0; k < primes; ++k) {
    f.q.push_back(number("modulus"));
    f.psi.push_back(number("root"));
  }
  if (take() != "COEFFS") throw std::runtime_error("vector file: expected COEFFS");
  f.coeffs.reserve(f.n);
  for (uint64_t i = 0; i < f.n; ++i) {
    const std::string &c = take();
    if (c.empty() || c.find_first_not_of("0123456789") != std::string::npos)
      throw std::runtime_error("vector file: bad coefficient");
    f.coeffs.push_back(c);
  }
  return f;
}

inline VectorFile read_vector_file(const std::string &path) {
  std::ifstream in(path);
  if (!in) throw std::runtime_error("cannot open vector file: " + path);
  return parse_vector_file(in);
}

// A decimal string of any length mod q: Horner's rule over the digits with
// 128-bit intermediates, so every q < 2^64 is accepted.
inline uint64_t reduce_decimal_mod(const std::string &decimal, uint64_t q) {
  if (q == 0) throw std::invalid_argument("modulus must be positive");
  unsigned __int128 r = 0;
  for (char ch : decimal) {
    if (ch < '0' || ch > '9') throw std::invalid_argument("not a decimal number");
    r = (r * 10 + static_cast<unsigned>(ch - '0')) % q;
  }
  return static_cast<uint64_t>(r);
}

// Limb-major residues of the file's coefficients: limb k at k * n.
inline std::vector<uint64_t> stage_limbs(const VectorFile &f) {
  std::vector<uint64_t> out(f.q.size() * f.n);
  for (size_t k = 0; k < f.q.size(); ++k)
    for (uint64_t i = 0; i < f.n; ++i) out[k * f.n + i] = reduce_decimal_mod(f.coeffs[i], f.q[k]);
  return out;
}

}  // namespace ntt
