/*
 * CPU oracle for the RNS-CKKS key-switching path: a plain-C restatement of the Rust CPU
 * path of ckks-train (crates ckks-train-core and ckks-train-ckks).
 *
 * TEST INFRASTRUCTURE ONLY. This library is the correctness checker and the timed CPU
 * baseline. Only the test suite, the smoke check and the CPU-baseline leg of the benchmark
 * load it; the product library never links, calls or falls back to it.
 *
 * Layout conventions (the same as the device library's):
 *   - a polynomial at level l is (l + 1) limbs of N u64 residues, limb i at a + i * N;
 *   - a Q*P element ("QpPoly") is its Q limbs 0..=l followed by its k P limbs;
 *   - a key-switching key with parts j < parts is, per part, b_j then a_j, each as Q limbs
 *     0..=key_max_level followed by the k P limbs (crates/ckks-train-cuda/src/ks.rs:76-116).
 *
 * Every function returns results that are bit-identical to the Rust functions it cites;
 * tests/test_oracle.py pins each of them against closed forms.
 */
#ifndef CKKS_ORACLE_H
#define CKKS_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef uint64_t u64;
typedef unsigned __int128 u128;

/* ---- modulus.rs ---- */
typedef struct {
    u64 q;
    u64 bar_lo, bar_hi; /* floor(2^128 / q) */
} or_mod;

int or_mod_new(u64 q, or_mod *m);
u64 or_mod_add(const or_mod *m, u64 a, u64 b);
u64 or_mod_sub(const or_mod *m, u64 a, u64 b);
u64 or_mod_mul(const or_mod *m, u64 a, u64 b);
u64 or_mod_pow(const or_mod *m, u64 a, u64 e);
/* Returns 1 and the inverse in *out, or 0 when gcd(a, q) != 1. */
int or_mod_inv(const or_mod *m, u64 a, u64 *out);
u64 or_shoup(const or_mod *m, u64 w);
u64 or_mul_shoup_lazy(const or_mod *m, u64 a, u64 w, u64 ws);
u64 or_mul_shoup(const or_mod *m, u64 a, u64 w, u64 ws);
/* Barrett constant as two words (for the tests). */
void or_mod_barrett(u64 q, u64 *lo, u64 *hi);

/* ---- primes.rs ---- */
int or_is_prime(u64 n);
/* 0 on success, -1 on InvalidParams. */
int or_ntt_primes_skip(unsigned bits, u64 n, u64 skip, u64 count, u64 *out);

/* ---- ntt.rs ---- */
typedef struct {
    u64 n;
    unsigned log_n;
    or_mod m;
    u64 psi;
    u64 *roots, *roots_s;   /* roots[bitrev(i)] = psi^i and Shoup companions */
    u64 *iroots, *iroots_s; /* psi^-i likewise */
    u64 n_inv, n_inv_s;
    u64 w_n_inv, w_n_inv_s;
} or_ntt;

int or_ntt_new(u64 n, u64 q, or_ntt *t);
void or_ntt_free(or_ntt *t);
void or_ntt_forward(const or_ntt *t, u64 *a);
void or_ntt_inverse(const or_ntt *t, u64 *a);
/* Standalone helpers for ctypes tests: build the table, transform `limbs` limbs in place. */
int or_ntt_forward_limbs(u64 n, const u64 *q, u64 limbs, u64 *a);
int or_ntt_inverse_limbs(u64 n, const u64 *q, u64 limbs, u64 *a);
u64 or_ntt_psi(u64 n, u64 q);

/* ---- params.rs (+ the bases of rns.rs) ---- */
typedef struct {
    unsigned log_n;
    u64 n;
    u64 nq, np, dnum, alpha;
    u64 *q, *p;
    or_mod *mq, *mp;
    or_ntt *tq, *tp;
} or_params;

/* SEAL-style prime assignment (params.rs:87-157). NULL on InvalidParams. */
or_params *or_params_new(unsigned log_n, const unsigned *q_bits, u64 nq,
                         const unsigned *p_bits, u64 np, u64 dnum);
/* Explicit primes (for tests that pick their own). */
or_params *or_params_from_primes(unsigned log_n, const u64 *q, u64 nq, const u64 *p,
                                 u64 np, u64 dnum);
void or_params_free(or_params *p);
u64 or_params_q(const or_params *p, u64 i);
u64 or_params_p(const or_params *p, u64 i);

/* ---- rns.rs / keys.rs: Galois elements and permutations ---- */
u64 or_galois_elt(u64 n, int step);
u64 or_conjugation_elt(u64 n);
void or_galois_perm_ntt(u64 n, u64 g, uint32_t *perm);
/* out[i] = in[perm[i]] for every limb (NTT form), poly.rs:249-254. */
void or_automorphism_ntt(u64 n, u64 g, const u64 *in, u64 *out, u64 limbs);
/* Signed scatter in coefficient form, poly.rs:255-276; primes[i] is limb i's prime. */
void or_automorphism_coeff(u64 n, u64 g, const u64 *primes, const u64 *in, u64 *out,
                           u64 limbs);

/* ---- baseconv.rs ---- */
/* BaseConverter::new(from, to).convert{_exact}(src): src k limbs, out t limbs. */
void or_convert(const u64 *from, u64 k, const u64 *to, u64 t, u64 n, const u64 *src,
                u64 *out, int exact);
/* In place: poly has level + 1 limbs; afterwards its first `level` limbs are the result. */
void or_rescale_by_last(const or_params *p, u64 *poly, u64 level, int ntt_form);
/* In place on the Q part (level + 1 limbs); p_part has `np_active` limbs. */
void or_mod_down(const or_params *p, u64 *q_part, u64 level, const u64 *p_part,
                 u64 np_active, int ntt_form);

/* ---- keyswitch.rs ---- */
/* Returns the digit count; starts/ends receive the ranges (may be NULL). */
u64 or_digit_ranges(const or_params *p, u64 level, u64 *starts, u64 *ends);
/* digits: ndig x (level + 1 + np) limbs (QpPoly layout), NTT form. */
void or_decompose(const or_params *p, const u64 *poly, u64 level, u64 *digits);
void or_apply_key(const or_params *p, const u64 *digits, u64 ndig, u64 level,
                  const u64 *key, u64 key_max_level, u64 *kb, u64 *ka);
void or_key_switch(const or_params *p, const u64 *poly, u64 level, const u64 *key,
                   u64 key_max_level, u64 *kb, u64 *ka);

/* ---- evaluator.rs ---- */
/* c0, c1, c2 at level l (NTT form) -> out0 = c0 + kb, out1 = c1 + ka. */
void or_relinearize(const or_params *p, const u64 *c0, const u64 *c1, const u64 *c2,
                    u64 level, const u64 *key, u64 key_max_level, u64 *out0, u64 *out1);
/* (sigma_g(c0) + kb, ka) with (kb, ka) = key_switch(sigma_g(c1)). */
void or_apply_galois(const or_params *p, const u64 *c0, const u64 *c1, u64 level, u64 g,
                     const u64 *key, u64 key_max_level, u64 *out0, u64 *out1);
/* Hoisted: one decomposition of c1, one key per element; outs: nsteps x 2 polys. */
void or_hoisted_galois(const or_params *p, const u64 *c0, const u64 *c1, u64 level,
                       u64 nsteps, const u64 *elts, const u64 *const *keys,
                       u64 key_max_level, u64 *outs);
/* Both components rescaled (in place; 2 x (level + 1) limbs in, first `level` limbs of each
 * component valid afterwards, component stride (level + 1) * N). */
void or_rescale_ct(const or_params *p, u64 *ct, u64 ncomp, u64 level);
/* (a0 b0, a0 b1 + a1 b0, a1 b1), NTT form. */
void or_mul_no_relin(const or_params *p, const u64 *a0, const u64 *a1, const u64 *b0,
                     const u64 *b1, u64 level, u64 *d0, u64 *d1, u64 *d2);
/* Pointwise ops on `limbs` limbs of the Q basis starting at limb 0. */
void or_poly_add(const or_params *p, u64 *a, const u64 *b, u64 limbs);
void or_poly_mul(const or_params *p, u64 *a, const u64 *b, u64 limbs);
void or_poly_mul_scalar(const or_params *p, u64 *a, u64 limbs, u64 k, int neg);
void or_raise_from_level0(const or_params *p, const u64 *a, u64 top, int ntt, u64 *out);
void or_params_ntt(const or_params *p, u64 *a, u64 limbs, int forward);

/* ---- seeded inputs (shared definition with the device generator) ---- */
u64 or_uniform_word(u64 seed, u64 stream, u64 index, u64 q);
void or_fill_uniform(u64 *out, u64 n, u64 q, u64 seed, u64 stream);

/* ---- threading (par.rs shape) ---- */
void or_set_threads(int n);
int or_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
