/*
 * ntt_serve.h — C ABI of the continuous-batching server (libntt_host.so).
 * No reference counterpart: the reference serves one synchronous Hybrid
 * graph per call (include/ntt/openfhe_hybrid_adapter.hpp:163-181). Requests
 * of mixed ring dimension, level and operation are batched as they arrive;
 * keys and tables stay resident. Status codes as in ntt_hip.h.
 */
#ifndef NTT_SERVE_H
#define NTT_SERVE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif
#pragma GCC visibility push(default)

#define NTT_SERVE_KEYSWITCH 0
#define NTT_SERVE_FORWARD 1
#define NTT_SERVE_INVERSE 2
#define NTT_SERVE_MODUP 3
#define NTT_SERVE_MODDOWN 4

typedef struct ntt_serve ntt_serve;

const char *ntt_serve_last_error(void);
/* hbm_budget_bytes bounds everything the server holds on the device (resident
 * plans, keys and per-ring pools, request buffers and workspaces); streams =
 * concurrent batches; max_batch = requests per launch. */
int ntt_serve_create(ntt_serve **server, int device, size_t hbm_budget_bytes, size_t streams, size_t max_batch);
int ntt_serve_destroy(ntt_serve *server); /* drains every admitted request */
/* Key-switching parameter set: M Q towers, K P towers, digit size alpha and
 * the full-context key (ceil(M/alpha) digits of (M+K)*n words per channel). */
int ntt_serve_add_params(ntt_serve *server, uint64_t n, const uint64_t *q, const uint64_t *q_psi, size_t M,
                         const uint64_t *p, const uint64_t *p_psi, size_t K, size_t alpha, const uint64_t *key_b,
                         const uint64_t *key_a, uint32_t *params_id);
int ntt_serve_output_words(ntt_serve *server, uint32_t params_id, int op, size_t level, size_t *words);
/* Copies `words` input residues (level*n; ModDown: (level+K)*n). */
int ntt_serve_submit(ntt_serve *server, uint32_t params_id, int op, size_t level, const uint64_t *in, size_t words,
                     uint64_t *ticket);
/* Device-resident request: in/out/out2 are the caller's device buffers
 * (out2: the A contribution of a key switch, NULL otherwise); ciphertexts stay
 * in HBM, nothing crosses the host; inputs must be canonical. Collect with
 * ntt_serve_wait(..., out = NULL, ...). */
int ntt_serve_submit_device(ntt_serve *server, uint32_t params_id, int op, size_t level, const uint64_t *in,
                            uint64_t *out, uint64_t *out2, uint64_t *ticket);
int ntt_serve_ready(ntt_serve *server, uint64_t ticket); /* NTT_HIP_OK ready, NTT_HIP_NOT_READY pending */
/* Blocks until done; out2 receives the A contribution of a key switch; out =
 * NULL for a device-resident request. */
int ntt_serve_wait(ntt_serve *server, uint64_t ticket, uint64_t *out, size_t out_words, uint64_t *out2,
                   size_t out2_words, double *latency_ms, size_t *batch_size);
/* submitted, completed, batches, max concurrent batches, device bytes held, peak */
int ntt_serve_stats(ntt_serve *server, size_t *stats6);
/* the first `count` (<= 9) of: the six above, resident bytes (contexts, keys,
 * per-ring pools), resident plans, plans evicted to make room */
int ntt_serve_stats_ex(ntt_serve *server, size_t *stats, size_t count);

#pragma GCC visibility pop
#ifdef __cplusplus
}
#endif
#endif /* NTT_SERVE_H */
