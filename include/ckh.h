/*
 * ckh: RNS-CKKS key-switching backend for AMD Instinct MI355X (gfx950).
 *
 * A C ABI over a C++ host runtime and offline-compiled HIP kernels. It replaces the device
 * backend behind the reference's `HeOps` trait (crates/ckks-train-ckks/src/ops.rs:198-356),
 * i.e. the role of `CudaEvaluator` / `CudaDevice` (crates/ckks-train-cuda/src/evaluator.rs,
 * src/device.rs). Plain pointers and sizes only; every residue array is u64, limb-major
 * (limb i of a polynomial at offset i * N), NTT form unless stated.
 *
 * Semantics, checks, errors and operation counts follow the CPU `Evaluator`
 * (crates/ckks-train-ckks/src/evaluator.rs); results are bit-identical to the CPU path.
 *
 * Threading: a context may be used from several threads; each thread works on its selected
 * stream (ckh_set_stream). Operations are asynchronous on that stream: a returned handle may
 * be used at once by later calls; downloads and ckh_synchronize wait. A handle used on several
 * streams must be released from the thread whose stream used it last (or after a synchronize).
 * Ownership: handles are reference counted (`retain` = the reference's cheap `Clone`,
 * ops.rs:17-20); device memory goes back to the context's block cache on the last release, in
 * the order of the stream it was allocated on, after the work queued so far on the releasing
 * thread's stream; it is reused from there, not returned to the driver until ckh_trim_pool.
 */
#ifndef CKH_H
#define CKH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes: OK plus one per HekitError variant (crates/ckks-train-core/src/error.rs:7-21). */
typedef enum ckh_status {
    CKH_OK = 0,
    CKH_ERR_INVALID_PARAMS = 1, /* HekitError::InvalidParams */
    CKH_ERR_SCALE_MISMATCH = 2, /* HekitError::ScaleMismatch */
    CKH_ERR_LEVEL_MISMATCH = 3, /* HekitError::LevelMismatch */
    CKH_ERR_NO_LEVELS_LEFT = 4, /* HekitError::NoLevelsLeft */
    CKH_ERR_DECODE = 5,         /* HekitError::Decode */
    CKH_ERR_DEVICE = 6          /* HekitError::Device: never a crash (device.rs:29-50) */
} ckh_status;

typedef struct ckh_context ckh_context;
typedef struct ckh_poly ckh_poly;
typedef struct ckh_ksk ckh_ksk;
typedef struct ckh_ct ckh_ct;
typedef struct ckh_evaluator ckh_evaluator;
typedef struct ckh_server ckh_server;

/* Message of the calling thread's last error (the String payload of HekitError). */
const char *ckh_last_error(void);

/* ------------------------------------------------------------------ parameters */

/* SEAL-style prime hand-out of CkksParams::new (crates/ckks-train-ckks/src/params.rs:87-157):
 * per bit size the c largest NTT primes (descending, primes.rs:67-112) handed out from the
 * smallest in order of appearance. Writes nq primes to q_out and np to p_out. */
ckh_status ckh_params_primes(uint32_t log_n, const uint32_t *q_bits, uint32_t nq,
                             const uint32_t *p_bits, uint32_t np, uint64_t *q_out,
                             uint64_t *p_out);

/* Galois elements (crates/ckks-train-ckks/src/keys.rs:89-107). */
uint64_t ckh_galois_elt(uint32_t log_n, int32_t step);
uint64_t ckh_conjugation_elt(uint32_t log_n);

typedef struct ckh_params {
    uint32_t log_n;     /* 6 ..= 17 */
    const uint64_t *q;  /* data primes q_0..q_L (< 2^60, = 1 mod 2N) */
    uint32_t nq;
    const uint64_t *p;  /* special primes */
    uint32_t np;
    uint32_t dnum;      /* key-switch digits; alpha = nq / dnum, must divide */
} ckh_params;

/* ------------------------------------------------------------------ context */

/* Opens device `ordinal` for one parameter set and builds its tables (CudaDevice::new,
 * crates/ckks-train-cuda/src/device.rs:262-330). n_streams >= 1 HIP streams (default priority)
 * for callers, plus the serving loop's own streams. */
ckh_status ckh_device_open(int ordinal, const ckh_params *params, uint32_t n_streams,
                           ckh_context **out);
void ckh_device_close(ckh_context *ctx);
/* Selects the calling thread's stream (device.rs:131-136,539-574). */
ckh_status ckh_set_stream(ckh_context *ctx, uint32_t index);
/* The calling thread's hipStream_t (for event timing by callers). */
void *ckh_stream_handle(ckh_context *ctx);
/* The project's autotuner (no counterpart in the reference): times candidate launch schedules
 * on this device with synthetic limbs of the context's primes and keeps the fastest (NTT chunks,
 * the batched key-switch schedule, and the small-batch schedule on one top-level switch); writes
 * a JSON report (candidates and choices) to report (may be NULL). */
ckh_status ckh_autotune(ckh_context *ctx, char *report, size_t report_len);
/* The autotuner's key-switch candidates (conversion kernel variant, row-pass fusion, then the
 * inverse-NTT chunk against none) timed on the caller's own batch: n NTT-form polynomials switched under key (no counterpart in the
 * reference). Keeps the fastest; writes the timings as JSON to report (may be NULL). */
ckh_status ckh_autotune_ks(ckh_context *ctx, uint32_t n, const ckh_poly *const *d, const ckh_ksk *key,
                           char *report, size_t report_len);
/* Forces one tuning value: "ntt_chunk_fwd", "ntt_chunk_inv" (limbs per launch, 0 = all);
 * "graphs" (1 = replay repeated launch sequences from captured hipGraphs, the default; 0 = launch
 * directly); "graph_cache" (captured sequences kept per context); "conv_variant" (base conversion:
 * 0 = vector kernel, 1 = vector kernel with exact f64-piece sums for 60-bit sources, 2 = int8
 * matrix-core kernel); "conv_variant_md" (the variant of mod_down's conversions in batched
 * switches, default 1; 3 = conv_variant); "conv_tg" (targets per base-conversion
 * workgroup); "fuse_ip" (1 = row pass fused into the inner product); "fuse_fin" (1, the default:
 * with fuse_ip, mod_down's finish of the Q limbs in the same kernel, so the inner-product sums
 * never reach memory); "conv_yform" (a mask of the conversions whose inverse NTT ends with the
 * conversion's first product, [x + h] (Q/q_j)^-1, folded into its last stage: 1 = ModUp on the
 * vector kernel, 2 = mod_down, 4 = ModUp on the matrix-core kernel; default 7); "small_limbs" (batched
 * switches of at most this many input limbs in all, a single switch for one, use their own
 * schedule: "conv_variant_small", "conv_tg_small", "fuse_ip_small"); "pool" (1 = the context's
 * block cache, the default: freed device blocks stay cached per size class and serve their own
 * stream at once and another stream once their last use has completed on the device; 0 = the
 * device's default stream-ordered pool, its frees ordered after every stream that used the
 * buffer; the switch waits for every stream first); "service_streams" (1..4 streams the serving
 * loop spreads its batches over, default 4); "hoisted_batch" (hoisted-rotation requests the
 * serving loop runs as one batch, default 4); "ntt_fused" (1 = whole transforms of key-switch
 * programs as one persistent launch per direction, default 0), "fused_lag", "fused_wg" (its
 * wait depth and workgroups per CU). Read only (ckh_get_tuning): "xcds", "stat_mallocs",
 * "stat_malloc_us", "stat_oom", "stat_cached_bytes". Test hooks, as the reference's failing
 * backend (crates/ckks-train-cuda/tests/cases/failing.rs:34-67), negative = off:
 * "fail_alloc_after" (the n-th device allocation from now fails with CKH_ERR_DEVICE, as an
 * out-of-memory would); "fail_launch_after" (the n-th launch from now, counted per program,
 * elementwise kernel or automorphism, fails with CKH_ERR_DEVICE before anything of it is queued;
 * the hook disarms itself when it fires); "fail_upload" (1 = every upload of polynomial,
 * ciphertext or key limbs fails with CKH_ERR_DEVICE). Setting a hook keeps captured plans. */
ckh_status ckh_set_tuning(ckh_context *ctx, const char *key, int64_t value);
/* CudaDevice::pool_used_memory / set_pool_release_threshold / trim_pool
 * (crates/ckks-train-cuda/src/device.rs:478-527) for the context's block cache (and the
 * device's default pool): bytes in use; bytes kept cached when freed (default: all, so
 * temporaries are reused; beyond it a freed block goes back to the driver once its last use has
 * finished); and, after every stream has finished, the cache beyond keep_bytes returned to the
 * driver (call after one-off peaks on a shared GPU). */
ckh_status ckh_pool_used_memory(ckh_context *ctx, uint64_t *bytes);
ckh_status ckh_set_pool_release_threshold(ckh_context *ctx, uint64_t bytes);
ckh_status ckh_trim_pool(ckh_context *ctx, uint64_t keep_bytes);
/* Reads one tuning value (the keys of ckh_set_tuning). */
ckh_status ckh_get_tuning(ckh_context *ctx, const char *key, int64_t *value);
/* Captured hipGraphs so far and their replays (no counterpart in the reference). */
ckh_status ckh_graph_stats(ckh_context *ctx, uint64_t *captures, uint64_t *launches);
/* Waits for all work of the context. */
ckh_status ckh_synchronize(ckh_context *ctx);
/* Device memory: bytes in use by live handles, and free device bytes (evaluator.rs:496-519). */
ckh_status ckh_memory_info(ckh_context *ctx, uint64_t *live_bytes, uint64_t *free_bytes);

/* ------------------------------------------------------------------ polynomials */

/* A polynomial over Q_level (level + 1 limbs) followed by p_limbs special-prime limbs
 * (0 or np: a QpPoly, crates/ckks-train-ckks/src/keys.rs:109-150). */
ckh_status ckh_poly_upload(ckh_context *ctx, const uint64_t *limbs, uint32_t level,
                           uint32_t p_limbs, int ntt_form, ckh_poly **out);
ckh_status ckh_poly_download(ckh_context *ctx, const ckh_poly *poly, uint64_t *out);
ckh_status ckh_poly_info(const ckh_poly *poly, uint32_t *level, uint32_t *p_limbs,
                         int *ntt_form);
ckh_poly *ckh_poly_retain(ckh_poly *poly);
void ckh_poly_release(ckh_poly *poly);
/* An order-independent 64-bit digest of every word of the polynomial (sum of a 64-bit mix of
 * (word, index)), for comparing large results on the device. No counterpart in the reference. */
ckh_status ckh_poly_digest(ckh_context *ctx, const ckh_poly *poly, uint64_t *digest);
/* Seeded uniform residues on the device: limb i from generator stream stream0 + i. */
ckh_status ckh_poly_uniform(ckh_context *ctx, uint32_t level, uint32_t p_limbs, int ntt_form,
                            uint64_t seed, uint64_t stream0, ckh_poly **out);
/* RnsPoly::to_ntt / to_coeff (crates/ckks-train-core/src/poly.rs:192-211). */
ckh_status ckh_ntt(ckh_context *ctx, const ckh_poly *in, ckh_poly **out);
ckh_status ckh_intt(ckh_context *ctx, const ckh_poly *in, ckh_poly **out);
/* RnsPoly::automorphism (poly.rs:242-283): the NTT-form permutation or, for coefficient-form
 * input, the signed permutation X -> X^g. */
ckh_status ckh_automorphism(ckh_context *ctx, const ckh_poly *in, uint64_t galois_elt,
                            ckh_poly **out);
/* Measurement hook (no counterpart in the reference): transforms every limb of the n polys in
 * place (forward or inverse), as one batched launch, reps times between two HIP events on the
 * calling thread's stream, and writes the mean milliseconds per launch. Overwrites the
 * polynomials' contents (residues stay reduced). */
ckh_status ckh_time_ntt(ckh_context *ctx, ckh_poly *const *polys, uint32_t n, int forward,
                        uint32_t reps, double *ms_per_launch);
/* Measurement hook (no counterpart in the reference): a streaming device copy of bytes (a
 * multiple of 16) between two fresh buffers, 16-byte loads and stores, reps launches between two
 * HIP events on the calling thread's stream; writes the mean milliseconds per launch. */
ckh_status ckh_time_copy(ckh_context *ctx, uint64_t bytes, uint32_t reps, double *ms_per_launch);
/* RnsPoly::rescale_by_last (crates/ckks-train-core/src/baseconv.rs:258-291). */
ckh_status ckh_rescale_poly(ckh_context *ctx, const ckh_poly *in, ckh_poly **out);
/* RnsPoly::mod_down of a QpPoly (baseconv.rs:459-511): Q part level l -> level l. */
ckh_status ckh_mod_down(ckh_context *ctx, const ckh_poly *qp, ckh_poly **out);
/* RnsPoly::raise_from_level0 (crates/ckks-train-core/src/poly.rs:298-345; device form
 * crates/ckks-train-cuda/src/basis.rs:513-533): a level-0 Q polynomial to level top, limb 0 as it
 * is and limbs 1..=top the centered lift of its coefficients (same form as the input). */
ckh_status ckh_raise_from_level0(ckh_context *ctx, const ckh_poly *in, uint32_t top, ckh_poly **out);

/* ------------------------------------------------------------------ key switching */

/* A key-switching key in the reference's device order (crates/ckks-train-cuda/src/ks.rs:76-116):
 * parts j < digits(max_level), each b_j then a_j, each as Q limbs 0..=max_level then the np P
 * limbs, NTT form. galois_elt 0 = relinearization key. */
ckh_status ckh_ksk_upload(ckh_context *ctx, const uint64_t *limbs, uint32_t max_level,
                          uint64_t galois_elt, ckh_ksk **out);
ckh_status ckh_ksk_uniform(ckh_context *ctx, uint32_t max_level, uint64_t galois_elt,
                           uint64_t seed, ckh_ksk **out);
ckh_status ckh_ksk_info(const ckh_ksk *key, uint32_t *max_level, uint64_t *galois_elt,
                        uint64_t *bytes);
void ckh_ksk_release(ckh_ksk *key);

/* keyswitch::decompose (crates/ckks-train-ckks/src/keyswitch.rs:82-124): the exact centered
 * lift of every digit of d (NTT form, level l) to Q_l * P. digits must hold dnum handles. */
ckh_status ckh_decompose(ckh_context *ctx, const ckh_poly *d, ckh_poly **digits,
                         uint32_t *n_digits);
/* keyswitch::apply_key (keyswitch.rs:230-309): inner product and mod_down of both sums. */
ckh_status ckh_apply_key(ckh_context *ctx, ckh_poly *const *digits, uint32_t n_digits,
                         const ckh_ksk *key, ckh_poly **kb, ckh_poly **ka);
/* keyswitch::key_switch (keyswitch.rs:317-319): one ciphertext per launch sequence. */
ckh_status ckh_key_switch(ckh_context *ctx, const ckh_poly *d, const ckh_ksk *key,
                          ckh_poly **kb, ckh_poly **ka);
/* n key switches under one shared key, polynomials at any levels, in batched launches
 * (no counterpart in the reference; bit-identical to n calls of ckh_key_switch). */
ckh_status ckh_key_switch_batch(ckh_context *ctx, uint32_t n, const ckh_poly *const *d,
                                const ckh_ksk *key, ckh_poly **kb, ckh_poly **ka);

/* Device bytes of the fixed workspace of ckh_key_switch_batch_static (one per GPU, held from the
 * first context on that GPU until the process ends). */
#define CKH_STATIC_SCRATCH_BYTES (256u << 20)
/* ckh_key_switch_batch without any device allocation. No counterpart in the reference: it is for
 * callers that cannot allocate (code under stream capture, serving loops that forbid hipMalloc).
 * The caller passes the outputs: kb[i] and ka[i] are existing polynomials of this context at the
 * level of d[i] (NTT form, Q limbs only), overwritten in place (every handle sharing them sees
 * the new words); they may not overlap any input or each other. Intermediates live in the
 * GPU's fixed workspace of CKH_STATIC_SCRATCH_BYTES (descriptors included; a 24-limb switch at
 * N = 2^16 with 6 special primes and dnum 4 needs about 132 MiB); a batch needing more is
 * rejected with CKH_ERR_INVALID_PARAMS before any work. Launches go directly on the calling
 * thread's stream, with no graph capture. Every call on one GPU shares that workspace: calls that
 * overlap in time on one GPU (other streams, other contexts) race and must be ordered by the
 * caller; calls on one stream are ordered by it. Results equal ckh_key_switch_batch's. */
ckh_status ckh_key_switch_batch_static(ckh_context *ctx, uint32_t n, const ckh_poly *const *d,
                                       const ckh_ksk *key, ckh_poly *const *kb,
                                       ckh_poly *const *ka);

/* ------------------------------------------------------------------ ciphertexts */

/* n_comp polynomials of level + 1 limbs each, back to back, NTT form; scale is the f64 label. */
ckh_status ckh_ct_upload(ckh_context *ctx, const uint64_t *polys, uint32_t n_comp,
                         uint32_t level, double scale, ckh_ct **out);
ckh_status ckh_ct_from_polys(ckh_context *ctx, ckh_poly *const *polys, uint32_t n_comp,
                             double scale, ckh_ct **out);
ckh_status ckh_ct_download(ckh_context *ctx, const ckh_ct *ct, uint64_t *out);
ckh_status ckh_ct_info(const ckh_ct *ct, uint32_t *n_comp, uint32_t *level, double *scale);
/* Borrowed component handle (retain it to keep it beyond the ciphertext). */
ckh_poly *ckh_ct_component(const ckh_ct *ct, uint32_t index);
ckh_ct *ckh_ct_retain(ckh_ct *ct);
void ckh_ct_release(ckh_ct *ct);

/* ------------------------------------------------------------------ evaluator (HeOps) */

/* OpCounter (crates/ckks-train-ckks/src/counter.rs:42-79). */
typedef struct ckh_op_counter {
    uint64_t rotations, relins, ct_ct_tensor, ct_pt_mults, rescales, sse_switches, bootstraps;
} ckh_op_counter;

/* Evaluator::new(params, relin, galois) (evaluator.rs:179-186); keys are shared, not copied. */
ckh_status ckh_evaluator_create(ckh_context *ctx, const ckh_ksk *relin,
                                const ckh_ksk *const *galois, uint32_t n_galois,
                                ckh_evaluator **out);
/* HeOps::fork (ops.rs:243-246): same keys and configuration, own counter from zero. */
ckh_status ckh_evaluator_fork(ckh_evaluator *ev, ckh_evaluator **out);
void ckh_evaluator_destroy(ckh_evaluator *ev);
/* Evaluator::with_config / config (EvalConfig, ops.rs:111-134; default all 1): hoist = one
 * shared decomposition per hoisted rotation (0: individual rotations, same words and counts);
 * bsgs = one giant rotation per BSGS block (0: one per product, d - 1 rotations); lazy_relin =
 * mul_lazy is mul_no_relin and relin_lazy relinearizes (0: mul, and relin_lazy passes through). */
ckh_status ckh_evaluator_set_config(ckh_evaluator *ev, uint32_t hoist, uint32_t bsgs, uint32_t lazy_relin);
ckh_status ckh_evaluator_get_config(ckh_evaluator *ev, uint32_t *hoist, uint32_t *bsgs, uint32_t *lazy_relin);

ckh_status ckh_counter(ckh_evaluator *ev, ckh_op_counter *out);
ckh_status ckh_reset_counter(ckh_evaluator *ev);
ckh_status ckh_add_counts(ckh_evaluator *ev, const ckh_op_counter *counts);
int ckh_has_relin_key(ckh_evaluator *ev);
int ckh_has_galois_key(ckh_evaluator *ev, int32_t step);
int ckh_owns(ckh_evaluator *ev, const ckh_ct *ct);

/* HeOps::{add, sub, negate} (evaluator.rs add/sub/negate). */
ckh_status ckh_add(ckh_evaluator *ev, const ckh_ct *a, const ckh_ct *b, ckh_ct **out);
ckh_status ckh_sub(ckh_evaluator *ev, const ckh_ct *a, const ckh_ct *b, ckh_ct **out);
ckh_status ckh_negate(ckh_evaluator *ev, const ckh_ct *a, ckh_ct **out);
/* HeOps::add_plain / mul_plain (evaluator.rs:287-308): the plaintext is an NTT-form Q polynomial
 * at the ciphertext's level with f64 scale pt_scale. */
ckh_status ckh_add_plain(ckh_evaluator *ev, const ckh_ct *a, const ckh_poly *pt, double pt_scale,
                         ckh_ct **out);
ckh_status ckh_mul_plain(ckh_evaluator *ev, const ckh_ct *a, const ckh_poly *pt, double pt_scale,
                         ckh_ct **out);
/* HeOps::mul_scalar (evaluator.rs:369-383): ct times the integer round(value * L / scale),
 * labelled L = exact_product_scale(level, target_scale), so a rescale lands on target_scale. */
ckh_status ckh_mul_scalar(ckh_evaluator *ev, const ckh_ct *ct, double value, double target_scale,
                          ckh_ct **out);
/* HeOps::mul_plain_to_scale (evaluator.rs:399-418): pt_scale must equal
 * target_plain_scale(a, target_scale) exactly; the product is labelled L. */
ckh_status ckh_mul_plain_to_scale(ckh_evaluator *ev, const ckh_ct *a, const ckh_poly *pt, double pt_scale,
                                  double target_scale, ckh_ct **out);
/* HeOps::exact_product_scale / target_plain_scale (ops.rs:322-337; evaluator.rs:330-345,
 * 738-772): the label L with L / q_level == target_scale exactly, and L / ct scale. */
ckh_status ckh_exact_product_scale(ckh_evaluator *ev, uint32_t level, double target_scale, double *out);
ckh_status ckh_target_plain_scale(ckh_evaluator *ev, const ckh_ct *ct, double target_scale, double *out);
/* Evaluator::mod_raise (crates/ckks-train-ckks/src/bootstrap/modraise.rs:53-69): a level-0
 * ciphertext switched to the ephemeral key at level 0 (to_sparse: one part at level 0), both
 * components raised to max_level, and switched back (to_dense: dnum parts at max_level), the
 * SseKeys of keys.rs:425-467 in the key layout of ckh_ksk_upload; sse_switches += 2.
 * mod_raise_main_key (modraise.rs:77-80): the raise alone. */
ckh_status ckh_mod_raise(ckh_evaluator *ev, const ckh_ct *ct, const ckh_ksk *to_sparse, const ckh_ksk *to_dense,
                         ckh_ct **out);
ckh_status ckh_mod_raise_main_key(ckh_evaluator *ev, const ckh_ct *ct, ckh_ct **out);
/* HeOps::mul_lazy / relin_lazy (ops.rs:339-356), following the lazy_relin flag. */
ckh_status ckh_mul_lazy(ckh_evaluator *ev, const ckh_ct *a, const ckh_ct *b, ckh_ct **out);
ckh_status ckh_relin_lazy(ckh_evaluator *ev, const ckh_ct *sum, ckh_ct **out);
/* HeOps::mul_no_relin (evaluator.rs:422-438), relinearize (441-458), mul (463-467). */
ckh_status ckh_mul_no_relin(ckh_evaluator *ev, const ckh_ct *a, const ckh_ct *b, ckh_ct **out);
ckh_status ckh_relinearize(ckh_evaluator *ev, const ckh_ct *ct, ckh_ct **out);
ckh_status ckh_mul(ckh_evaluator *ev, const ckh_ct *a, const ckh_ct *b, ckh_ct **out);
/* HeOps::rescale (evaluator.rs:498-509), mod_switch_to_level (513-525), relabel (ops.rs:300-306). */
ckh_status ckh_rescale(ckh_evaluator *ev, const ckh_ct *ct, ckh_ct **out);
ckh_status ckh_mod_switch_to_level(ckh_evaluator *ev, const ckh_ct *ct, uint32_t level,
                                   ckh_ct **out);
ckh_status ckh_relabel(ckh_evaluator *ev, const ckh_ct *ct, double scale, ckh_ct **out);
/* HeOps::rotate (evaluator.rs:534-540), conjugate (547-550), hoisted_rotate (572-611). */
ckh_status ckh_rotate(ckh_evaluator *ev, const ckh_ct *ct, int32_t steps, ckh_ct **out);
ckh_status ckh_conjugate(ckh_evaluator *ev, const ckh_ct *ct, ckh_ct **out);
ckh_status ckh_hoisted_rotate(ckh_evaluator *ev, const ckh_ct *ct, const int32_t *steps,
                              uint32_t n_steps, ckh_ct **outs);
/* Evaluator::matvec_plain's BSGS product (crates/ckks-train-ckks/src/linalg.rs:608-621,800-888)
 * for Diagonals::from_prerotated_with_offset (linalg.rs:230-260): entries[k1 * baby + k2] is
 * rot(diag_{offset + baby k1 + k2}, -(offset + baby k1) stride) (NULL = a zero entry, skipped),
 * all at x's level with scale pt_scale; -baby * giant < offset <= 0. The baby rotations
 * k2 * stride share one decomposition; block k1 is rotated by (offset + baby k1) stride (the
 * giant rotations of all blocks are switched in one batch). Returns the unrescaled sum. */
ckh_status ckh_matvec_plain(ckh_evaluator *ev, const ckh_ct *x, const ckh_poly *const *entries,
                            double pt_scale, uint32_t baby, uint32_t giant, uint32_t stride,
                            int32_t offset, ckh_ct **out);
/* Evaluator::matvec_enc (linalg.rs:638-665): the same product with encrypted 2-component
 * diagonals; lazy_relin: per block the tensor products are summed and relinearized once. */
ckh_status ckh_matvec_enc(ckh_evaluator *ev, const ckh_ct *x, const ckh_ct *const *entries, uint32_t baby,
                          uint32_t giant, uint32_t stride, int32_t offset, ckh_ct **out);
/* Evaluator::matvec_plain_sum_sparse (crates/ckks-train-ckks/src/linalg.rs:729-790): sum over
 * n_terms of W_t x_t, one shape for every term, entries[t][k1 * baby + k2] pre-rotated as for
 * ckh_matvec_plain (NULL = zero, skipped with the rotations only it would use), every entry at
 * the terms' common level with scale pt_scale; the giant rotations are shared by all terms. */
ckh_status ckh_matvec_plain_sum(ckh_evaluator *ev, uint32_t n_terms, const ckh_ct *const *xs,
                                const ckh_poly *const *const *entries, double pt_scale, uint32_t baby,
                                uint32_t giant, uint32_t stride, int32_t offset, ckh_ct **out);
/* CudaEvaluator::matvec_enc_batch (crates/ckks-train-cuda/src/evaluator.rs:412-519): n
 * independent products (job i: xs[i], entries[i][0..baby * giant), shapes[i]); every job is
 * checked first, in order; a product whose working set exceeds the device memory available now
 * (free plus the pool's unused reserve) fails with CKH_ERR_DEVICE before any work; the products
 * then run one after the other on the calling thread's stream, and their counts are merged
 * only when every job succeeded. */
typedef struct ckh_bsgs_shape {
    uint32_t baby, giant, stride;
    int32_t offset;
} ckh_bsgs_shape;
/* One stage of CoeffToSlot / SlotToCoeff (LinearTransform, crates/ckks-train-ckks/src/bootstrap/
 * linear.rs:467-502): baby * giant plaintext diagonals at `level` (NULL = zero), pre-rotated as
 * for ckh_matvec_plain, with the common scale pt_scale; `second` (NULL: a one-input stage) the
 * diagonals applied to the second input, in their own shape and scale, which must equal the
 * first's; `scale` the exact ciphertext scale after the stage's rescale (to level - 1). */
typedef struct ckh_linear_stage {
    int32_t slot_to_coeff; /* Direction: 0 CoeffToSlot, 1 SlotToCoeff */
    uint32_t level;
    double scale;
    ckh_bsgs_shape shape;
    const ckh_poly *const *diags;
    double pt_scale;
    ckh_bsgs_shape second_shape;
    const ckh_poly *const *second;
    double second_pt_scale;
} ckh_linear_stage;
/* Evaluator::coeff_to_slot (linear.rs:776-806): the stages in order, the last one (and only it)
 * also applied to the conjugate of its input; then re = x + conj(x), im = (conj(x) - x) X^(N/2).
 * Every check (directions, stage order, levels descending from ct's, diagonal scales against
 * each stage's exact pre-rescale label, layouts, rotation keys, the conjugation key) is made
 * before any work. */
ckh_status ckh_coeff_to_slot(ckh_evaluator *ev, const ckh_ct *ct, uint32_t n_stages, const ckh_linear_stage *stages,
                             ckh_ct **re, ckh_ct **im);
/* Evaluator::slot_to_coeff (linear.rs:824-859): the first stage (and only it) takes re and im,
 * the others the previous output; re and im at one level with exactly one scale. */
ckh_status ckh_slot_to_coeff(ckh_evaluator *ev, const ckh_ct *re, const ckh_ct *im, uint32_t n_stages,
                             const ckh_linear_stage *stages, ckh_ct **out);
ckh_status ckh_matvec_enc_batch(ckh_evaluator *ev, uint32_t n, const ckh_ct *const *xs,
                                const ckh_ct *const *const *entries, const ckh_bsgs_shape *shapes,
                                ckh_ct **outs);
/* CudaEvaluator::matvec_enc_workspace_bytes (evaluator.rs:403-410) for this engine. */
ckh_status ckh_matvec_enc_workspace_bytes(ckh_evaluator *ev, uint32_t level, uint32_t baby, uint32_t giant,
                                          uint64_t *bytes);
/* n relinearizations in batched launches (bit-identical to n ckh_relinearize calls). */
ckh_status ckh_relinearize_batch(ckh_evaluator *ev, uint32_t n, const ckh_ct *const *in,
                                 ckh_ct **out);

/* ------------------------------------------------------------------ serving */

/* No counterpart in the reference (its trainers call HeOps synchronously). A scheduler thread
 * keeps evaluators and their keys resident, admits requests in arrival order while the bytes
 * of the admitted work stay within hbm_budget_bytes, and batches whatever is admitted at each
 * turn: relinearizations of one evaluator as one batched key switch, mul+relin+rescale with its
 * products in one launch, its relinearization batched likewise and its rescales in one program,
 * hoisted rotations in groups of "hoisted_batch" requests. Batches run round-robin on
 * service streams of each context that no caller thread can select (tuning "service_streams"),
 * so work admitted while other batches run overlaps with them and callers may keep working on
 * their own streams meanwhile. Results are bit-identical to the same operations called alone. */
typedef enum ckh_op_kind {
    CKH_OP_RELINEARIZE = 0,       /* a: 3 components -> 1 output */
    CKH_OP_HOISTED_ROTATE = 1,    /* a: 2 components, steps -> n_steps outputs */
    CKH_OP_MUL_RELIN_RESCALE = 2  /* a, b: 2 components -> rescale(relinearize(a * b)) */
} ckh_op_kind;

typedef struct ckh_server_config {
    uint64_t hbm_budget_bytes; /* bound on the device bytes of admitted, unfinished requests */
    uint32_t max_batch;        /* requests admitted per scheduler turn */
    uint32_t max_inflight;     /* batches launched and unfinished at once (0: 4); while that many
                                * run, arrivals wait unadmitted (no device memory held) */
} ckh_server_config;

typedef struct ckh_request {
    uint32_t kind;           /* ckh_op_kind */
    ckh_evaluator *ev;       /* its context's streams run the work */
    const ckh_ct *a, *b;
    const int32_t *steps;    /* CKH_OP_HOISTED_ROTATE */
    uint32_t n_steps;
    double arrival_s;        /* arrival on the server clock (ckh_server_now); < 0: now */
    uint32_t inject_failure; /* test hook: the request fails with CKH_ERR_DEVICE */
} ckh_request;

typedef struct ckh_server_stats {
    uint64_t submitted, completed, failed, rejected, batches, peak_bytes;
    double p50_ms, p99_ms, mean_ms, max_ms; /* latency: completion - arrival */
    double makespan_s, ops_per_s;           /* first arrival to last completion */
    double launch_ms, launch_max_ms;        /* scheduler host time enqueueing batches: total, longest */
} ckh_server_stats;

ckh_status ckh_server_create(const ckh_server_config *cfg, ckh_server **out);
/* Waits for the requests in flight, then stops the scheduler. */
void ckh_server_destroy(ckh_server *srv);
double ckh_server_now(ckh_server *srv);
/* The request's inputs must have been produced on the calling thread's stream of their context
 * (the scheduler waits for that point before using them). */
ckh_status ckh_server_submit(ckh_server *srv, const ckh_request *req, uint64_t *ticket);
/* *finished = 1 once the request is done (wait != 0 blocks until then); then op_status holds
 * its result (its message in ckh_last_error), outs[0..n_outs) its ciphertexts (released by the
 * caller) and latency_ms its completion minus arrival. A finished ticket is forgotten. */
ckh_status ckh_server_poll(ckh_server *srv, uint64_t ticket, int wait, int *finished, ckh_ct **outs,
                           uint32_t max_outs, uint32_t *n_outs, ckh_status *op_status,
                           double *latency_ms);
ckh_status ckh_server_stats_get(ckh_server *srv, ckh_server_stats *out);

#ifdef __cplusplus
}
#endif
#endif /* CKH_H */
