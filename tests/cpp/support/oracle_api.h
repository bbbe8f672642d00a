#pragma once
// C declarations of the CPU oracle (oracle/ntt_oracle.c) used as the checker
// by the C++ tests. Test infrastructure only.
#include <stdint.h>

extern "C" {
int or_key_switch(const uint64_t *in, uint64_t n, const uint64_t *q, const uint64_t *qpsi, uint64_t L, uint64_t M,
                  const uint64_t *p, const uint64_t *ppsi, uint64_t K, uint64_t alpha, const uint64_t *key_b,
                  const uint64_t *key_a, uint64_t *out_b, uint64_t *out_a);
int or_ntt_towers(uint64_t *a, uint64_t n, uint64_t towers, const uint64_t *q, const uint64_t *psi, int inverse);
int or_basis_switch(const uint64_t *in, uint64_t n, const uint64_t *src, uint64_t s, const uint64_t *dst, uint64_t t,
                    uint64_t *out);
int or_mod_up(const uint64_t *in, int in_eval, uint64_t n, const uint64_t *q, const uint64_t *qpsi, uint64_t L,
              const uint64_t *p, const uint64_t *ppsi, uint64_t K, uint64_t *out);
uint64_t or_automorphism_source(uint64_t dest, uint64_t n, uint32_t k);
int or_hybrid_digits(const uint64_t *in, uint64_t n, const uint64_t *q, const uint64_t *qpsi, uint64_t L,
                     const uint64_t *p, const uint64_t *ppsi, uint64_t K, uint64_t alpha, uint64_t *out);
void or_hybrid_rotate(const uint64_t *ext_b, const uint64_t *ext_a, const uint64_t *c0, uint64_t n,
                      const uint64_t *q, uint64_t L, const uint64_t *p, uint64_t K, uint32_t k, uint64_t *out_b,
                      uint64_t *out_a);
int or_mod_down(const uint64_t *in, uint64_t n, const uint64_t *q, const uint64_t *qpsi, uint64_t L,
                const uint64_t *p, const uint64_t *ppsi, uint64_t K, uint64_t *out);
}
