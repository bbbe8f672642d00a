/*
 * ntt_hip.h — C ABI of the gfx950 CKKS hot-path library (libntt_hip.so).
 *
 * Plain pointers and sizes only. Device buffers are raw device pointers to
 * tower-major uint64 residues: element (tower t, index i) at t * n + i
 * (reference docs/openfhe-representation-contract.md:385-404). Every call
 * returns NTT_HIP_OK or a negative status; ntt_hip_last_error() gives the
 * message of the calling thread's last failure. NTT_HIP_EINVAL is a contract
 * violation detected before any device work (the reference's ContractError,
 * include/ntt/backend_contract.hpp:20-26); NTT_HIP_ERUNTIME is a HIP runtime
 * failure (the reference's std::runtime_error("CUDA: ..."),
 * src/cuda/ntt_kernels.cu:40-46).
 *
 * Streams and events are opaque handles (hipStream_t / hipEvent_t). A NULL
 * stream argument means the context's own ordered stream
 * (src/cuda/resident_adapter.inc:91).
 */
#ifndef NTT_HIP_H
#define NTT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif
#pragma GCC visibility push(default)

#define NTT_HIP_OK 0
#define NTT_HIP_NOT_READY 1
#define NTT_HIP_EINVAL (-1)
#define NTT_HIP_ERUNTIME (-2)

#define NTT_HIP_H2D 1
#define NTT_HIP_D2H 2
#define NTT_HIP_D2D 3

typedef struct ntt_hip_context ntt_hip_context;
typedef struct ntt_hip_keyswitch ntt_hip_keyswitch;

/* ---- status ---- */
const char *ntt_hip_last_error(void);
int ntt_hip_abi_version(void);

/* ---- device, memory, streams, events: the plumbing of the reference CUDA
 * Driver (src/cuda/resident_adapter.inc:131-156 Buffer/Event,
 * :178-197 upload/copy/copy_range, :281-301 download/record/ready/wait/quiesce,
 * src/cuda/ntt_kernels.cu:57-75 DeviceGuard). ---- */
int ntt_hip_device_count(int *count);
int ntt_hip_device_name(int device, char *name, size_t capacity);
/* Architecture name (gcnArchName), device memory, compute units, driver and
 * runtime versions: the device description of the reference CLI's info and
 * JSONL workload lines (src/cuda/ntt_kernels.cu:491-506, :664-680). */
int ntt_hip_device_properties(int device, char *arch, size_t arch_capacity, uint64_t *memory_bytes,
                              int *compute_units, int *driver_version, int *runtime_version);
int ntt_hip_malloc(int device, void **ptr, size_t bytes);
int ntt_hip_free(int device, void *ptr);
int ntt_hip_host_alloc(void **ptr, size_t bytes); /* pinned host staging */
int ntt_hip_host_free(void *ptr);
/* kind: NTT_HIP_H2D / D2H / D2D; asynchronous on `stream` */
int ntt_hip_memcpy(int device, void *dst, const void *src, size_t bytes, int kind, void *stream);
int ntt_hip_stream_create(int device, void **stream);
int ntt_hip_stream_destroy(int device, void *stream);
int ntt_hip_stream_synchronize(int device, void *stream);
int ntt_hip_event_create(int device, void **event, int timing);
int ntt_hip_event_destroy(int device, void *event);
int ntt_hip_event_record(int device, void *event, void *stream);
int ntt_hip_event_query(int device, void *event); /* NTT_HIP_OK ready, NTT_HIP_NOT_READY pending */
int ntt_hip_event_synchronize(int device, void *event);
int ntt_hip_event_elapsed(float *ms, void *start, void *stop);
/* dst[i] = src[i] for `words` words with a plain streaming kernel: the
 * calibration kernel for HBM byte counters (no reference counterpart). */
int ntt_hip_copy_words(int device, uint64_t *dst, const uint64_t *src, size_t words, void *stream);

/* ---- context: one fixed basis of towers (q_t, psi_t) of ring dimension n
 * with its device twiddle tables (reference resident_adapter.inc:60-130
 * Context; CudaNttTransform ctor src/cuda/ntt_kernels.cu:336-386;
 * preflight include/ntt/cuda_validation.hpp:17-37: power-of-two n in
 * [2, 65536], prime q < 2^63 with 2n | q-1, psi^n = -1). ---- */
int ntt_hip_context_create(ntt_hip_context **ctx, int device, uint64_t n, const uint64_t *q, const uint64_t *psi,
                           size_t towers);
int ntt_hip_context_destroy(ntt_hip_context *ctx);
void *ntt_hip_context_stream(ntt_hip_context *ctx);
int ntt_hip_context_device(ntt_hip_context *ctx);
/* bytes of device tables held by the context (reference Driver::constant_bytes,
 * resident_driver.hpp:28 / resident_adapter.inc:171-174) */
size_t ntt_hip_context_table_bytes(ntt_hip_context *ctx);

/* Device bytes a context holds: its twiddle tables and prime records plus the
 * basis-conversion plans and scale tables it has cached (admission budgets). */
int ntt_hip_context_device_bytes(ntt_hip_context *ctx, size_t *bytes);

/* ---- autotuner (no reference counterpart): times candidate launch
 * shapes on `towers` scratch towers (forward + inverse transforms of the
 * context's first prime; a first-half -> rest basis switch) and keeps the
 * fastest in the context, one knob at a time, a candidate replacing the
 * current choice only when clearly faster: for N = 2^16 the forward headroom
 * (reductions skipped while residues stay below 2^64, or one per stage), the
 * row-pass residues per thread (8 or 16) and the workgroups per CU (occupancy
 * cap); for every N the tower chunk (towers per pass-A/pass-B pair; 0 = all)
 * and the basis-conversion targets per workgroup. `report` receives a JSON
 * line with every timing and the choice. Every shape is bit-exact. ---- */
int ntt_hip_autotune(ntt_hip_context *ctx, size_t towers, char *report, size_t capacity);
int ntt_hip_set_tuning(ntt_hip_context *ctx, uint32_t ntt_chunk, uint32_t conv_targets);
int ntt_hip_get_tuning(ntt_hip_context *ctx, uint32_t *ntt_chunk, uint32_t *conv_targets);
/* fwd_headroom: -1 = from the moduli (default) or at most 0..5; row_residues:
 * 8 (default) or 16; wg_per_cu: 0 (no cap, default) or 1..15. */
int ntt_hip_set_ntt_shape(ntt_hip_context *ctx, int fwd_headroom, uint32_t row_residues, uint32_t wg_per_cu);
/* The same search timed on `batch` fused key switches of the plan (its key
 * uploaded), over the context's NTT shape and conversion width and the
 * plan's concurrent lanes, ModUp/MAC fusion and MAC/transform interleave
 * group; allocates its own buffers for the duration of the call. */
int ntt_hip_keyswitch_autotune(ntt_hip_keyswitch *ks, size_t batch, char *report, size_t capacity);
/* The plan's key-switch shape: lanes 1..8 (concurrent sub-batches),
 * mac_group (towers per ModUp/MAC interleave group, 0 = all at once) and
 * fuse_mac: 0 = the ModUp's last NTT pass writes the digits and the key inner
 * product reads them back, 1 = that pass and the inner product run as one
 * kernel with 128-bit sums, 2 = the same with 64-bit sums (defaults: lanes
 * 2, mac_group 0, fuse_mac 2). The fusion applies at N = 2^16 with moduli
 * below 2^61, the 8-residue row pass and sums that fit (128-bit: digits * q
 * < 2^64; 64-bit: 2 q times the digit count rounded up to a power of two <=
 * 2^64); otherwise the unfused path runs. Every shape is bit-exact. `fused`
 * reports whether the fused kernel applies to the plan as set. */
int ntt_hip_keyswitch_set_shape(ntt_hip_keyswitch *ks, uint32_t lanes, uint32_t mac_group, uint32_t fuse_mac);
int ntt_hip_keyswitch_get_shape(ntt_hip_keyswitch *ks, uint32_t *lanes, uint32_t *mac_group, uint32_t *fuse_mac,
                                int *fused);
int ntt_hip_get_ntt_shape(ntt_hip_context *ctx, int *fwd_headroom, uint32_t *row_residues, uint32_t *wg_per_cu);

/* ---- Driver seam operations (include/ntt/resident_driver.hpp:13-57) on
 * contiguous towers [first, first + towers) of the context basis ---- */
/* Driver::transform (resident_adapter.inc:198-203): in place on `data`;
 * `src` != NULL reads the input from src (out of place). Forward: natural
 * coefficients -> bit-reversed evaluations; inverse: the exact inverse. */
int ntt_hip_transform(ntt_hip_context *ctx, uint64_t *data, const uint64_t *src, size_t first, size_t towers,
                      int inverse, void *stream);
/* Batched form: `polys` polynomials of `towers` towers each, poly p at
 * data + p * poly_stride (elements). One launch for all of them. */
int ntt_hip_transform_batch(ntt_hip_context *ctx, uint64_t *data, const uint64_t *src, size_t polys,
                            size_t poly_stride, size_t src_poly_stride, size_t first, size_t towers, int inverse,
                            void *stream);
/* Transforms of arbitrary towers in one launch (no reference counterpart:
 * the reference launches one basis per context): job j transforms the tower
 * at data + jobs[j].offset (reading src + jobs[j].src_offset when src != NULL)
 * with the prime of context tower jobs[j].tower. Polynomials of different
 * tower counts, levels or parameter sets that share the context's primes run
 * together. The host array is checked (tower indices) and staged into
 * `jobs_workspace` (device, count * sizeof(ntt_hip_tower_job) bytes) on
 * `stream`, which must stay untouched until the launch has run. Not for
 * stream capture: the staging copy reads the caller's host array, which is
 * only guaranteed during the call. */
typedef struct ntt_hip_tower_job {
  uint64_t offset, src_offset; /* elements */
  uint32_t tower, reserved;
} ntt_hip_tower_job;
int ntt_hip_transform_jobs(ntt_hip_context *ctx, uint64_t *data, const uint64_t *src, const ntt_hip_tower_job *jobs,
                           size_t count, void *jobs_workspace, int inverse, void *stream);
/* Transforms of towers of different ring dimensions in one launch per pass
 * (no reference counterpart: the reference transforms one basis of one N per
 * call). Job j transforms the N residues at jobs[j].data (its first pass
 * reading jobs[j].src instead when non-NULL) with tower jobs[j].tower of context
 * jobs[j].ctx; the contexts may differ in N (2^13 .. 2^16) and basis but
 * share one device, and all their moduli are below 2^61 (towers below 2^30
 * and wider ones run as two launches). The kernels are persistent: a fixed
 * grid whose workgroups stride over (tower, block) work items. The table is
 * checked on the host and staged into `jobs_workspace` (device,
 * ntt_hip_transform_mixed_workspace_bytes(count) bytes) on `stream` (NULL:
 * the first job's context stream), which must stay untouched until the
 * launch has run. Not for stream capture (the table is staged from a host
 * array that lives only during the call). */
typedef struct ntt_hip_mixed_job {
  ntt_hip_context *ctx;
  uint64_t *data;
  const uint64_t *src;
  uint32_t tower, reserved;
} ntt_hip_mixed_job;
int ntt_hip_transform_mixed_workspace_bytes(size_t count, size_t *bytes);
int ntt_hip_transform_mixed(const ntt_hip_mixed_job *jobs, size_t count, void *jobs_workspace, int inverse,
                            void *stream);
/* Driver::pointwise (resident_adapter.inc:204-213): out[i] = out[i] * right[i] mod q */
int ntt_hip_pointwise(ntt_hip_context *ctx, uint64_t *out, const uint64_t *right, size_t first, size_t towers,
                      void *stream);
/* Driver::switch_basis (resident_adapter.inc:254-266, include/ntt/basis_conversion.hpp:13-114):
 * coefficient-domain unreduced CRT from source towers src[0..s) (indices into
 * the context basis, input tower-major s*n) to target towers dst[0..t)
 * (output t*n). 1..40 each, distinct and disjoint. */
int ntt_hip_switch_basis(ntt_hip_context *ctx, uint64_t *out, const uint64_t *in, const size_t *src, size_t s,
                         const size_t *dst, size_t t, void *stream);
/* Driver::correct (resident_adapter.inc:267-280, src/resident.cpp:404-430):
 * out = (left - right) * P^-1 mod q on towers [first, first+towers); P is the
 * product of context towers p[0..k). */
int ntt_hip_correct(ntt_hip_context *ctx, uint64_t *out, const uint64_t *left, const uint64_t *right, size_t first,
                    size_t towers, const size_t *p, size_t k, void *stream);
/* Driver::hybrid_mac (resident_adapter.inc:214-231): b,a = (first ? 0 : b,a) + digit * key_{b,a} */
int ntt_hip_hybrid_mac(ntt_hip_context *ctx, uint64_t *b, uint64_t *a, const uint64_t *digit, const uint64_t *key_b,
                       const uint64_t *key_a, size_t first, size_t towers, int first_digit, void *stream);
/* Consecutive Driver::hybrid_mac calls on the same outputs (src/resident.cpp:530-570
 * issues one per digit) in one pass: b,a = (first ? 0 : b,a) + sum_j digits[j] *
 * key_{b,a}[j], equal to calling ntt_hip_hybrid_mac for j = 0..count-1 in order
 * (first_digit applying to j = 0). */
int ntt_hip_hybrid_mac_digits(ntt_hip_context *ctx, uint64_t *b, uint64_t *a, const uint64_t *const *digits,
                              const uint64_t *const *key_b, const uint64_t *const *key_a, size_t count, size_t first,
                              size_t towers, int first_digit, void *stream);
/* Driver::hybrid_rotate (resident_adapter.inc:232-253): sigma_k(B + P*c0), sigma_k(A)
 * on evaluation QP towers [first, first+towers) whose first active_q are Q;
 * P = product of context towers p[0..k). */
int ntt_hip_hybrid_rotate(ntt_hip_context *ctx, uint64_t *out_b, uint64_t *out_a, const uint64_t *ext_b,
                          const uint64_t *ext_a, const uint64_t *c0, size_t first, size_t towers, size_t active_q,
                          const size_t *p, size_t k, uint32_t automorphism, void *stream);

/* ---- fused Hybrid key switch (KeySwitchHybrid core:
 * include/ntt/openfhe_hybrid_adapter.hpp:163-181,212-219; graph
 * src/resident.cpp:458-570 + 448-457). The context basis must be Q towers
 * [0, L) followed by P towers [L, L + K); alpha = towers per digit. ---- */
int ntt_hip_keyswitch_create(ntt_hip_keyswitch **ks, ntt_hip_context *ctx, size_t L, size_t K, size_t alpha,
                             size_t max_batch);
int ntt_hip_keyswitch_destroy(ntt_hip_keyswitch *ks);
/* Device bytes the plan owns (resident key, job tables), beside its context's. */
int ntt_hip_keyswitch_device_bytes(ntt_hip_keyswitch *ks, size_t *bytes);
/* Resident evaluation key (Session::upload_hybrid_key, src/resident.cpp:492-529):
 * host key_b/key_a hold ceil(M/alpha) digits of the FULL context (M Q towers,
 * then the K P towers), digit-major, each (M + K) * n; q_full/psi_full are the
 * M full-context Q towers (their first L must equal the context's Q towers).
 * Active rows use key index i (< L) and i + (M - L) (P towers). */
int ntt_hip_keyswitch_upload_key(ntt_hip_keyswitch *ks, const uint64_t *key_b, const uint64_t *key_a, size_t M,
                                 const uint64_t *q_full);
/* Device workspace bytes for `batch` components. */
int ntt_hip_keyswitch_workspace_bytes(ntt_hip_keyswitch *ks, size_t batch, size_t *bytes);
/* Key-switch `batch` components: in = batch * L * n evaluation residues,
 * out_b / out_a = batch * L * n each (B then A contributions, evaluation Q).
 * All pointers are device pointers; workspace of the size above. */
int ntt_hip_keyswitch_run(ntt_hip_keyswitch *ks, const uint64_t *in, size_t batch, uint64_t *out_b, uint64_t *out_a,
                          void *workspace, void *stream);
/* The same batch with every component at its own device addresses (no
 * counterpart in the reference; for callers whose components live in separate
 * buffers, e.g. a server batching the requests of different clients):
 * in[c] = L * n evaluation residues, out_b[c] / out_a[c] = L * n each. The
 * pointer arrays are host arrays of device pointers, staged behind the
 * workspace on `stream`; the workspace holds
 * ntt_hip_keyswitch_gather_workspace_bytes(ks, batch) bytes. Outputs equal
 * ntt_hip_keyswitch_run on the same components gathered into one buffer.
 * Not for stream capture (the pointer tables are staged from host arrays
 * that live only during the call; ntt_hip_keyswitch_capture records the
 * contiguous form). */
int ntt_hip_keyswitch_gather_workspace_bytes(ntt_hip_keyswitch *ks, size_t batch, size_t *bytes);
int ntt_hip_keyswitch_run_gather(ntt_hip_keyswitch *ks, const uint64_t *const *in, size_t batch,
                                 uint64_t *const *out_b, uint64_t *const *out_a, void *workspace, void *stream);
/* Allocation-free form of ntt_hip_keyswitch_run (no counterpart in the
 * reference). For callers that cannot allocate device memory: code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc. The
 * workspace is a static device array inside the library, 136 MiB per GPU
 * (one component at N = 2^16, L = 24, K = 6, dnum = 4), held from the first
 * load until the process ends; a batch whose workspace
 * (ntt_hip_keyswitch_workspace_bytes) exceeds it is rejected with
 * NTT_HIP_EINVAL. Calls no hipMalloc/hipFree and creates no streams or events
 * (one lane on `stream`). Every call on a GPU shares that GPU's array, so calls
 * that overlap on one GPU (other streams, other threads) race. */
int ntt_hip_keyswitch_run_static(ntt_hip_keyswitch *ks, const uint64_t *in, size_t batch, uint64_t *out_b,
                                 uint64_t *out_a, void *stream);
/* The same batch captured once as a HIP graph on `stream` (fixed buffer
 * addresses; no memset nodes) and replayed by ntt_hip_graph_launch. No
 * reference counterpart: the reference launches every stage per call. */
int ntt_hip_keyswitch_capture(ntt_hip_keyswitch *ks, const uint64_t *in, size_t batch, uint64_t *out_b,
                              uint64_t *out_a, void *workspace, void *stream, void **graph);
int ntt_hip_graph_launch(void *graph, void *stream);
/* Batched ModUp (Session::mod_up, src/resident.cpp:431-447) on the plan's
 * basis: in = batch * L * n (evaluation if in_eval, else coefficient Q),
 * out = batch * (L + K) * n evaluation QP. */
int ntt_hip_modup_workspace_bytes(ntt_hip_keyswitch *ks, size_t batch, size_t *bytes);
int ntt_hip_modup_run(ntt_hip_keyswitch *ks, const uint64_t *in, int in_eval, size_t batch, uint64_t *out,
                      void *workspace, void *stream);
/* Batched ModDown (Session::mod_down, src/resident.cpp:448-457, correction
 * :404-430): in = batch * (L + K) * n evaluation QP, out = batch * L * n. */
int ntt_hip_moddown_workspace_bytes(ntt_hip_keyswitch *ks, size_t batch, size_t *bytes);
int ntt_hip_moddown_run(ntt_hip_keyswitch *ks, const uint64_t *in, size_t batch, uint64_t *out, void *workspace,
                        void *stream);
int ntt_hip_graph_destroy(void *graph);
/* In-place key switch (reference openfhe_hybrid_adapter.hpp:221-248
 * hybrid_key_switch_in_place; docs/openfhe-hybrid-operation-graph.md:350-362):
 * c2 == NULL: two-component ciphertexts (c0, c1): c1 is key-switched,
 *   c0 += B and c1 = A;
 * c2 != NULL: three-component (c0, c1, c2): c2 is key-switched, c0 += B and
 *   c1 += A (c2 is left as it was; the caller drops it).
 * All components are batch * L * n evaluation residues on the device (the
 * caller converts coefficient-form components with the forward NTT first, as
 * the adapter's SetFormat(EVALUATION) does). The additions are fused into the
 * ModDown's last pass. Workspace: ntt_hip_keyswitch_workspace_bytes. The
 * components must be distinct buffers (NTT_HIP_EINVAL otherwise). */
int ntt_hip_keyswitch_in_place(ntt_hip_keyswitch *ks, uint64_t *c0, uint64_t *c1, const uint64_t *c2, size_t batch,
                               void *workspace, void *stream);
/* Relinearization (reference openfhe_hybrid_adapter.hpp:250-286
 * hybrid_relinearize_in_place): component 2 + j (higher[j], j < extra) is
 * key-switched with keys[j] (the j-th relinearization key, OpenFHE's
 * EvalMultKeyVector order) and each contribution, reduced by its own ModDown,
 * is added into c0 (B) and c1 (A). extra == 0 leaves (c0, c1) unchanged. Keys
 * share one context and (L, K, alpha); keys[0] sizes the workspace
 * (ntt_hip_keyswitch_workspace_bytes). Evaluation form, as above. */
int ntt_hip_relinearize(ntt_hip_keyswitch *const *keys, size_t extra, uint64_t *c0, uint64_t *c1,
                        const uint64_t *const *higher, size_t batch, void *workspace, void *stream);
/* Hoisted rotations (reference openfhe_hybrid_adapter.hpp:307-360
 * hybrid_rotate_many over Session::hybrid_digits/hybrid_accumulate/
 * hybrid_rotate_extended/mod_down, src/resident.cpp:458-617): the digits of
 * c1 are computed once for all `rotations`; for rotation r, the key inner
 * product with keys[r]'s resident key, sigma_k(B + P c0) and sigma_k(A) on
 * the QP basis (k = automorphisms[r], odd, < 2n; computed before ModDown as
 * OpenFHE does) and ModDown of both channels into out_b[r] / out_a[r]
 * (batch * L * n each, evaluation Q). c0, c1: batch * L * n evaluation
 * residues. keys[0..rotations) are plans of one context with equal L, K,
 * alpha; keys[0] sizes the workspace (ntt_hip_rotate_workspace_bytes). */
int ntt_hip_rotate_workspace_bytes(ntt_hip_keyswitch *ks, size_t batch, size_t *bytes);
int ntt_hip_keyswitch_rotate(ntt_hip_keyswitch *const *keys, size_t rotations, const uint32_t *automorphisms,
                             const uint64_t *c0, const uint64_t *c1, size_t batch, uint64_t *const *out_b,
                             uint64_t *const *out_a, void *workspace, void *stream);

#pragma GCC visibility pop
#ifdef __cplusplus
}
#endif
#endif /* NTT_HIP_H */
