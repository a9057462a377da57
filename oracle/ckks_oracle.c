/*
 * CPU oracle: plain-C restatement of the ckks-train CPU key-switching path.
 * TEST INFRASTRUCTURE ONLY (see ckks_oracle.h). Each function cites the Rust function it
 * restates; paths are relative to the reference repository root.
 *
 * Floating point: the only f64 step is the overflow estimate of convert_exact
 * (crates/ckks-train-core/src/baseconv.rs:198-206). It is computed with separate IEEE-754
 * round-to-nearest multiplies and adds in source order; this file is compiled with
 * -ffp-contract=off so no FMA contraction can change a rounding.
 */
#include "ckks_oracle.h"

#include <stdlib.h>
#include <string.h>
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ======================= modulus.rs ======================= */

/* High 128 bits of the 256-bit product x * y (modulus.rs:36-46). */
static u128 mulhi_u128(u128 x, u128 y) {
    u128 x0 = (u64)x, x1 = x >> 64, y0 = (u64)y, y1 = y >> 64;
    u128 ll = x0 * y0, lh = x0 * y1, hl = x1 * y0, hh = x1 * y1;
    u128 mid = (ll >> 64) + (u128)(u64)lh + (u128)(u64)hl;
    return hh + (lh >> 64) + (hl >> 64) + (mid >> 64);
}

void or_mod_barrett(u64 q, u64 *lo, u64 *hi) {
    /* floor(2^128 / q) = floor((2^128 - 1) / q), plus one when q | 2^128 (modulus.rs:56-61) */
    u128 maxv = ~(u128)0;
    u128 ratio = maxv / q;
    if (maxv % q + 1 == q) ratio += 1;
    *lo = (u64)ratio;
    *hi = (u64)(ratio >> 64);
}

/* Modulus::new (modulus.rs:50-66): 2 <= q < 2^60. */
int or_mod_new(u64 q, or_mod *m) {
    if (q < 2 || q >= (1ULL << 60)) return -1;
    m->q = q;
    or_mod_barrett(q, &m->bar_lo, &m->bar_hi);
    return 0;
}

/* Modulus::reduce_wide (modulus.rs:79-89). */
static u64 reduce_wide(const or_mod *m, u128 x) {
    u128 ratio = (u128)m->bar_lo | ((u128)m->bar_hi << 64);
    u128 est = mulhi_u128(x, ratio);
    u64 r = (u64)(x - est * (u128)m->q);
    if (r >= m->q) r -= m->q;
    return r;
}

u64 or_mod_add(const or_mod *m, u64 a, u64 b) { /* modulus.rs:92-96 */
    u64 s = a + b;
    return s >= m->q ? s - m->q : s;
}

u64 or_mod_sub(const or_mod *m, u64 a, u64 b) { /* modulus.rs:99-102 */
    return a >= b ? a - b : a + m->q - b;
}

u64 or_mod_mul(const or_mod *m, u64 a, u64 b) { /* modulus.rs:106-109 */
    return reduce_wide(m, (u128)a * b);
}

u64 or_mod_pow(const or_mod *m, u64 a, u64 e) { /* modulus.rs:112-123 */
    u64 base = a % m->q, acc = 1 % m->q;
    while (e > 0) {
        if (e & 1) acc = or_mod_mul(m, acc, base);
        base = or_mod_mul(m, base, base);
        e >>= 1;
    }
    return acc;
}

int or_mod_inv(const or_mod *m, u64 a, u64 *out) { /* modulus.rs:127-139 */
    __int128 r0 = (__int128)m->q, r1 = (__int128)(a % m->q);
    __int128 t0 = 0, t1 = 1;
    while (r1 != 0) {
        __int128 quo = r0 / r1;
        __int128 tmp = r0 - quo * r1;
        r0 = r1;
        r1 = tmp;
        tmp = t0 - quo * t1;
        t0 = t1;
        t1 = tmp;
    }
    if (r0 != 1) return 0;
    __int128 qq = (__int128)m->q;
    __int128 v = t0 % qq;
    if (v < 0) v += qq;
    *out = (u64)v;
    return 1;
}

u64 or_shoup(const or_mod *m, u64 w) { /* Modulus::shoup, modulus.rs:142-148 */
    w %= m->q;
    return (u64)(((u128)w << 64) / m->q);
}

/* mul_shoup_lazy (modulus.rs:152-155): [0, 2q), any u64 a. */
u64 or_mul_shoup_lazy(const or_mod *m, u64 a, u64 w, u64 ws) {
    u64 qhat = (u64)(((u128)a * ws) >> 64);
    return a * w - qhat * m->q;
}

u64 or_mul_shoup(const or_mod *m, u64 a, u64 w, u64 ws) { /* modulus.rs:158-161 */
    u64 r = or_mul_shoup_lazy(m, a, w, ws);
    return r >= m->q ? r - m->q : r;
}

/* A (value, companion) pair for a reduced multiplicand. */
typedef struct {
    u64 w, ws;
} shoup_t;

static shoup_t mk_shoup(const or_mod *m, u64 w) {
    shoup_t s;
    s.w = w % m->q;
    s.ws = or_shoup(m, s.w);
    return s;
}

static inline u64 msh(const or_mod *m, u64 a, shoup_t s) { return or_mul_shoup(m, a, s.w, s.ws); }

/* ======================= primes.rs ======================= */

static u64 mul_mod_plain(u64 a, u64 b, u64 m) { return (u64)((u128)a * b % m); }

static u64 pow_mod_plain(u64 a, u64 e, u64 m) { /* primes.rs:9-20 */
    u64 acc = 1 % m;
    a %= m;
    while (e > 0) {
        if (e & 1) acc = mul_mod_plain(acc, a, m);
        a = mul_mod_plain(a, a, m);
        e >>= 1;
    }
    return acc;
}

/* Deterministic Miller-Rabin with the first twelve primes as bases (primes.rs:24-53). */
int or_is_prime(u64 n) {
    static const u64 bases[12] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    if (n < 2) return 0;
    for (int i = 0; i < 12; ++i) {
        if (n == bases[i]) return 1;
        if (n % bases[i] == 0) return 0;
    }
    unsigned s = (unsigned)__builtin_ctzll(n - 1);
    u64 d = (n - 1) >> s;
    for (int i = 0; i < 12; ++i) {
        u64 x = pow_mod_plain(bases[i], d, n);
        if (x == 1 || x == n - 1) continue;
        int witness = 1;
        for (unsigned r = 1; r < s; ++r) {
            x = mul_mod_plain(x, x, n);
            if (x == n - 1) {
                witness = 0;
                break;
            }
        }
        if (witness) return 0;
    }
    return 1;
}

/* ntt_primes_skip (primes.rs:67-112): descending primes 2^(bits-1) < q < 2^bits, q = 1 mod 2n. */
int or_ntt_primes_skip(unsigned bits, u64 n, u64 skip, u64 count, u64 *out) {
    if (bits < 2 || bits > 60) return -1;
    if (n == 0 || (n & (n - 1)) != 0) return -1;
    u128 step = (u128)n * 2;
    u128 lo = (u128)1 << (bits - 1), hi = (u128)1 << bits;
    if (step >= hi) return count == 0 ? 0 : -1;
    u128 q = (hi - 2) / step * step + 1;
    u64 seen = 0, got = 0;
    while (got < count) {
        if (q <= lo) return -1;
        if (or_is_prime((u64)q)) {
            if (seen >= skip) out[got++] = (u64)q;
            seen++;
        }
        if (q < step) return -1;
        q -= step;
    }
    return 0;
}

/* ======================= ntt.rs ======================= */

static u64 bitrev_u(u64 i, unsigned bits) {
    if (bits == 0) return 0;
    u64 r = 0;
    for (unsigned b = 0; b < bits; ++b) r |= ((i >> b) & 1) << (bits - 1 - b);
    return r;
}

/* minimal_primitive_root (ntt.rs:66-82). */
static u64 minimal_primitive_root(u64 n, const or_mod *m) {
    u64 q = m->q, e = (q - 1) / (2 * n), c = 0;
    for (u64 g = 2;; ++g) {
        c = or_mod_pow(m, g, e);
        if (or_mod_pow(m, c, n) == q - 1) break;
    }
    u64 c2 = or_mod_mul(m, c, c), cur = c, best = c;
    for (u64 i = 1; i < n; ++i) {
        cur = or_mod_mul(m, cur, c2);
        if (cur < best) best = cur;
    }
    return best;
}

/* bitrev_power_table (ntt.rs:54-63). */
static void bitrev_power_table(u64 n, unsigned bits, const or_mod *m, u64 base, u64 *tab,
                               u64 *tab_s) {
    u64 p = 1;
    for (u64 i = 0; i < n; ++i) {
        u64 j = bitrev_u(i, bits);
        tab[j] = p;
        tab_s[j] = or_shoup(m, p);
        p = or_mod_mul(m, p, base);
    }
}

/* NttTable::new (ntt.rs:89-122). */
int or_ntt_new(u64 n, u64 q, or_ntt *t) {
    memset(t, 0, sizeof(*t));
    if (n < 2 || (n & (n - 1)) != 0) return -1;
    if (or_mod_new(q, &t->m) != 0) return -1;
    if (n >= q || (q - 1) % (2 * n) != 0) return -1;
    if (!or_is_prime(q)) return -1;
    t->n = n;
    t->log_n = (unsigned)__builtin_ctzll(n);
    t->psi = minimal_primitive_root(n, &t->m);
    u64 psi_inv = 0, n_inv = 0;
    or_mod_inv(&t->m, t->psi, &psi_inv);
    or_mod_inv(&t->m, n, &n_inv);
    t->roots = malloc(n * sizeof(u64));
    t->roots_s = malloc(n * sizeof(u64));
    t->iroots = malloc(n * sizeof(u64));
    t->iroots_s = malloc(n * sizeof(u64));
    bitrev_power_table(n, t->log_n, &t->m, t->psi, t->roots, t->roots_s);
    bitrev_power_table(n, t->log_n, &t->m, psi_inv, t->iroots, t->iroots_s);
    t->n_inv = n_inv;
    t->n_inv_s = or_shoup(&t->m, n_inv);
    t->w_n_inv = or_mod_mul(&t->m, t->iroots[1], n_inv);
    t->w_n_inv_s = or_shoup(&t->m, t->w_n_inv);
    return 0;
}

void or_ntt_free(or_ntt *t) {
    free(t->roots);
    free(t->roots_s);
    free(t->iroots);
    free(t->iroots_s);
    memset(t, 0, sizeof(*t));
}

/* NttTable::forward (ntt.rs:141-179): Cooley-Tukey, natural in, bit-reversed out, [0, q). */
void or_ntt_forward(const or_ntt *t, u64 *a) {
    const or_mod *m = &t->m;
    u64 q = m->q, two_q = 2 * q, n = t->n, tt = n, groups = 1;
    while (groups < n / 2) {
        tt >>= 1;
        for (u64 g = 0; g < groups; ++g) {
            u64 w = t->roots[groups + g], ws = t->roots_s[groups + g];
            u64 *lo = a + g * 2 * tt, *hi = lo + tt;
            for (u64 j = 0; j < tt; ++j) {
                u64 x = lo[j];
                u64 xr = x >= two_q ? x - two_q : x;
                u64 v = or_mul_shoup_lazy(m, hi[j], w, ws);
                lo[j] = xr + v;
                hi[j] = xr + two_q - v;
            }
        }
        groups <<= 1;
    }
    for (u64 g = 0; g < n / 2; ++g) {
        u64 w = t->roots[n / 2 + g], ws = t->roots_s[n / 2 + g];
        u64 x = a[2 * g];
        u64 xr = x >= two_q ? x - two_q : x;
        u64 v = or_mul_shoup_lazy(m, a[2 * g + 1], w, ws);
        u64 r0 = xr + v, r1 = xr + two_q - v;
        r0 = r0 >= two_q ? r0 - two_q : r0;
        r0 = r0 >= q ? r0 - q : r0;
        r1 = r1 >= two_q ? r1 - two_q : r1;
        r1 = r1 >= q ? r1 - q : r1;
        a[2 * g] = r0;
        a[2 * g + 1] = r1;
    }
}

/* NttTable::inverse (ntt.rs:184-232): Gentleman-Sande, bit-reversed in, natural out. */
void or_ntt_inverse(const or_ntt *t, u64 *a) {
    const or_mod *m = &t->m;
    u64 q = m->q, two_q = 2 * q, tt = 1, h = t->n >> 1;
    if (h > 1) {
        for (u64 g = 0; g < h; ++g) {
            u64 x = a[2 * g], y = a[2 * g + 1];
            u64 s = x + y;
            a[2 * g] = s >= two_q ? s - two_q : s;
            a[2 * g + 1] = or_mul_shoup_lazy(m, x + two_q - y, t->iroots[h + g], t->iroots_s[h + g]);
        }
        tt = 2;
        h >>= 1;
    }
    while (h > 1) {
        for (u64 g = 0; g < h; ++g) {
            u64 w = t->iroots[h + g], ws = t->iroots_s[h + g];
            u64 *lo = a + g * 2 * tt, *hi = lo + tt;
            for (u64 j = 0; j < tt; ++j) {
                u64 s = lo[j] + hi[j], d = lo[j] + two_q - hi[j];
                lo[j] = s >= two_q ? s - two_q : s;
                hi[j] = or_mul_shoup_lazy(m, d, w, ws);
            }
        }
        tt <<= 1;
        h >>= 1;
    }
    u64 half = t->n / 2;
    for (u64 j = 0; j < half; ++j) {
        u64 x = a[j], y = a[j + half];
        u64 s = x + y, d = x + two_q - y;
        u64 r0 = or_mul_shoup_lazy(m, s, t->n_inv, t->n_inv_s);
        u64 r1 = or_mul_shoup_lazy(m, d, t->w_n_inv, t->w_n_inv_s);
        a[j] = r0 >= q ? r0 - q : r0;
        a[j + half] = r1 >= q ? r1 - q : r1;
    }
}

static int ntt_limbs(u64 n, const u64 *q, u64 limbs, u64 *a, int fwd) {
    for (u64 i = 0; i < limbs; ++i) {
        or_ntt t;
        if (or_ntt_new(n, q[i], &t) != 0) return -1;
        if (fwd)
            or_ntt_forward(&t, a + i * n);
        else
            or_ntt_inverse(&t, a + i * n);
        or_ntt_free(&t);
    }
    return 0;
}

int or_ntt_forward_limbs(u64 n, const u64 *q, u64 limbs, u64 *a) { return ntt_limbs(n, q, limbs, a, 1); }
int or_ntt_inverse_limbs(u64 n, const u64 *q, u64 limbs, u64 *a) { return ntt_limbs(n, q, limbs, a, 0); }

u64 or_ntt_psi(u64 n, u64 q) {
    or_ntt t;
    if (or_ntt_new(n, q, &t) != 0) return 0;
    u64 psi = t.psi;
    or_ntt_free(&t);
    return psi;
}

/* ======================= params.rs ======================= */

static or_params *params_alloc(unsigned log_n, const u64 *q, u64 nq, const u64 *p, u64 np,
                               u64 dnum) {
    if (log_n < 1 || log_n > 17 || nq == 0 || np == 0 || dnum == 0 || nq % dnum != 0) return NULL;
    or_params *pp = calloc(1, sizeof(or_params));
    pp->log_n = log_n;
    pp->n = 1ULL << log_n;
    pp->nq = nq;
    pp->np = np;
    pp->dnum = dnum;
    pp->alpha = nq / dnum;
    pp->q = malloc(nq * sizeof(u64));
    pp->p = malloc(np * sizeof(u64));
    memcpy(pp->q, q, nq * sizeof(u64));
    memcpy(pp->p, p, np * sizeof(u64));
    pp->mq = calloc(nq, sizeof(or_mod));
    pp->mp = calloc(np, sizeof(or_mod));
    pp->tq = calloc(nq, sizeof(or_ntt));
    pp->tp = calloc(np, sizeof(or_ntt));
    int bad = 0;
    for (u64 i = 0; i < nq && !bad; ++i)
        bad |= or_mod_new(q[i], &pp->mq[i]) != 0 || or_ntt_new(pp->n, q[i], &pp->tq[i]) != 0;
    for (u64 i = 0; i < np && !bad; ++i)
        bad |= or_mod_new(p[i], &pp->mp[i]) != 0 || or_ntt_new(pp->n, p[i], &pp->tp[i]) != 0;
    if (bad) {
        or_params_free(pp);
        return NULL;
    }
    return pp;
}

or_params *or_params_from_primes(unsigned log_n, const u64 *q, u64 nq, const u64 *p, u64 np,
                                 u64 dnum) {
    return params_alloc(log_n, q, nq, p, np, dnum);
}

/* CkksParams::new (params.rs:87-157): per bit size, the c largest primes descending, handed
 * out from the smallest in order of appearance (SEAL's pop_back). */
or_params *or_params_new(unsigned log_n, const unsigned *q_bits, u64 nq, const unsigned *p_bits,
                         u64 np, u64 dnum) {
    if (log_n < 1 || log_n > 17 || nq == 0 || np == 0) return NULL;
    u64 tot = nq + np, n = 1ULL << log_n;
    unsigned *all = malloc(tot * sizeof(unsigned));
    for (u64 i = 0; i < nq; ++i) all[i] = q_bits[i];
    for (u64 i = 0; i < np; ++i) all[nq + i] = p_bits[i];
    u64 *primes = malloc(tot * sizeof(u64));
    int ok = 1;
    for (u64 i = 0; i < tot && ok; ++i) {
        if (all[i] < 2 || all[i] > 60) ok = 0;
    }
    for (unsigned b = 2; b <= 60 && ok; ++b) {
        u64 count = 0;
        for (u64 i = 0; i < tot; ++i) count += all[i] == b;
        if (count == 0) continue;
        u64 *pool = malloc(count * sizeof(u64));
        if (or_ntt_primes_skip(b, n, 0, count, pool) != 0) {
            ok = 0;
        } else {
            u64 next = 0;
            for (u64 i = 0; i < tot; ++i)
                if (all[i] == b) primes[i] = pool[count - 1 - next++];
        }
        free(pool);
    }
    or_params *pp = ok ? params_alloc(log_n, primes, nq, primes + nq, np, dnum) : NULL;
    free(all);
    free(primes);
    return pp;
}

void or_params_free(or_params *pp) {
    if (!pp) return;
    for (u64 i = 0; i < pp->nq; ++i) or_ntt_free(&pp->tq[i]);
    for (u64 i = 0; i < pp->np; ++i) or_ntt_free(&pp->tp[i]);
    free(pp->tq);
    free(pp->tp);
    free(pp->mq);
    free(pp->mp);
    free(pp->q);
    free(pp->p);
    free(pp);
}

u64 or_params_q(const or_params *p, u64 i) { return p->q[i]; }
u64 or_params_p(const or_params *p, u64 i) { return p->p[i]; }

/* ======================= rns.rs / keys.rs: Galois ======================= */

/* keys.rs:89-103, generator 3 (encoder.rs GALOIS_GENERATOR). */
u64 or_galois_elt(u64 n, int step) {
    u64 m = 2 * n;
    long long slots = (long long)(n / 2 > 0 ? n / 2 : 1);
    long long k0 = (long long)step % slots;
    if (k0 < 0) k0 += slots;
    u64 k = (u64)k0, base = 3 % m, acc = 1 % m;
    while (k > 0) {
        if (k & 1) acc = acc * base % m;
        base = base * base % m;
        k >>= 1;
    }
    return acc;
}

u64 or_conjugation_elt(u64 n) { return 2 * n - 1; } /* keys.rs:106-108 */

/* galois_permutation_ntt (rns.rs:147-157). */
void or_galois_perm_ntt(u64 n, u64 g, uint32_t *perm) {
    unsigned log_n = (unsigned)__builtin_ctzll(n);
    u64 mask = n - 1;
    for (u64 i = n; i < 2 * n; ++i) {
        u64 reversed = bitrev_u(i, log_n + 1);
        u64 index_raw = ((g * reversed) >> 1) & mask;
        perm[i - n] = (uint32_t)bitrev_u(index_raw, log_n);
    }
}

void or_automorphism_ntt(u64 n, u64 g, const u64 *in, u64 *out, u64 limbs) {
    uint32_t *perm = malloc(n * sizeof(uint32_t));
    or_galois_perm_ntt(n, g % (2 * n), perm);
    for (u64 l = 0; l < limbs; ++l)
        for (u64 i = 0; i < n; ++i) out[l * n + i] = in[l * n + perm[i]];
    free(perm);
}

void or_automorphism_coeff(u64 n, u64 g, const u64 *primes, const u64 *in, u64 *out,
                           u64 limbs) {
    u64 two_n = 2 * n;
    g %= two_n;
    for (u64 l = 0; l < limbs; ++l) {
        u64 q = primes[l], e = 0;
        for (u64 i = 0; i < n; ++i) {
            u64 c = in[l * n + i];
            if (e < n)
                out[l * n + e] = c;
            else
                out[l * n + e - n] = c == 0 ? 0 : q - c;
            e += g;
            if (e >= two_n) e -= two_n;
        }
    }
}

/* ======================= baseconv.rs ======================= */

typedef struct {
    u64 k, t;
    or_mod *from, *to;
    u64 *q_hat_inv;   /* k */
    u64 *q_hat_mod_p; /* t x k */
} conv_t;

/* (Q / q_j) mod m (baseconv.rs:104-111). */
static u64 q_hat_mod(const or_mod *from, u64 k, u64 j, const or_mod *m) {
    u64 acc = 1 % m->q;
    for (u64 i = 0; i < k; ++i)
        if (i != j) acc = or_mod_mul(m, acc, from[i].q % m->q);
    return acc;
}

/* BaseConverter::new (baseconv.rs:95-130). */
static void conv_new(conv_t *c, const u64 *from, u64 k, const u64 *to, u64 t) {
    c->k = k;
    c->t = t;
    c->from = malloc(k * sizeof(or_mod));
    c->to = malloc((t ? t : 1) * sizeof(or_mod));
    for (u64 j = 0; j < k; ++j) or_mod_new(from[j], &c->from[j]);
    for (u64 i = 0; i < t; ++i) or_mod_new(to[i], &c->to[i]);
    c->q_hat_inv = malloc(k * sizeof(u64));
    for (u64 j = 0; j < k; ++j) or_mod_inv(&c->from[j], q_hat_mod(c->from, k, j, &c->from[j]), &c->q_hat_inv[j]);
    c->q_hat_mod_p = malloc((t ? t : 1) * k * sizeof(u64));
    for (u64 i = 0; i < t; ++i)
        for (u64 j = 0; j < k; ++j) c->q_hat_mod_p[i * k + j] = q_hat_mod(c->from, k, j, &c->to[i]);
}

static void conv_free(conv_t *c) {
    free(c->from);
    free(c->to);
    free(c->q_hat_inv);
    free(c->q_hat_mod_p);
}

/* BaseConverter::convert_inner (baseconv.rs:171-233). */
static void conv_run(const conv_t *c, u64 n, const u64 *src, u64 *out, int exact) {
    u64 k = c->k;
    u64 *y = malloc(k * n * sizeof(u64));
    for (u64 j = 0; j < k; ++j) {
        shoup_t s = mk_shoup(&c->from[j], c->q_hat_inv[j]);
        for (u64 i = 0; i < n; ++i) y[j * n + i] = msh(&c->from[j], src[j * n + i], s);
    }
    u64 *e = NULL;
    if (exact && k > 1) {
        /* e = floor(sum_j f64(y_j) * (1.0 / f64(q_j))), RN products and sums in order. */
        double *inv = malloc(k * sizeof(double));
        for (u64 j = 0; j < k; ++j) inv[j] = 1.0 / (double)c->from[j].q;
        e = malloc(n * sizeof(u64));
        for (u64 i = 0; i < n; ++i) {
            double f = 0.0;
            for (u64 j = 0; j < k; ++j) {
                double prod = (double)y[j * n + i] * inv[j];
                f = f + prod;
            }
            e[i] = (u64)floor(f);
        }
        free(inv);
    }
    for (u64 t = 0; t < c->t; ++t) {
        const or_mod *p = &c->to[t];
        u64 *o = out + t * n;
        for (u64 i = 0; i < n; ++i) o[i] = 0;
        for (u64 j = 0; j < k; ++j) {
            shoup_t s = mk_shoup(p, c->q_hat_mod_p[t * k + j]);
            for (u64 i = 0; i < n; ++i) o[i] = or_mod_add(p, o[i], msh(p, y[j * n + i], s));
        }
        if (e) {
            u64 q_mod_p = 1 % p->q;
            for (u64 j = 0; j < k; ++j) q_mod_p = or_mod_mul(p, q_mod_p, c->from[j].q % p->q);
            shoup_t s = mk_shoup(p, q_mod_p);
            for (u64 i = 0; i < n; ++i) o[i] = or_mod_sub(p, o[i], msh(p, e[i], s));
        }
    }
    free(e);
    free(y);
}

void or_convert(const u64 *from, u64 k, const u64 *to, u64 t, u64 n, const u64 *src, u64 *out,
                int exact) {
    conv_t c;
    conv_new(&c, from, k, to, t);
    conv_run(&c, n, src, out, exact);
    conv_free(&c);
}

/* half_mod (baseconv.rs:238-243; keyswitch.rs:213-218): (D - 1)/2 mod m. */
static u64 half_mod(const or_mod *factors, u64 k, const or_mod *m) {
    u64 d = 1 % m->q;
    for (u64 i = 0; i < k; ++i) d = or_mod_mul(m, d, factors[i].q % m->q);
    return or_mod_mul(m, or_mod_sub(m, d, 1 % m->q), (m->q + 1) / 2);
}

/* RnsPoly::rescale_by_last (baseconv.rs:258-291). */
void or_rescale_by_last(const or_params *pp, u64 *poly, u64 l, int ntt) {
    u64 n = pp->n;
    const or_mod *ql = &pp->mq[l];
    u64 half = ql->q >> 1;
    u64 *last = malloc(n * sizeof(u64));
    memcpy(last, poly + l * n, n * sizeof(u64));
    if (ntt) or_ntt_inverse(&pp->tq[l], last);
    for (u64 c = 0; c < n; ++c) last[c] = or_mod_add(ql, last[c], half);
    u64 *t = malloc(n * sizeof(u64));
    for (u64 i = 0; i < l; ++i) {
        const or_mod *m = &pp->mq[i];
        shoup_t one = mk_shoup(m, 1);
        u64 half_i = msh(m, half, one);
        for (u64 c = 0; c < n; ++c) t[c] = or_mod_sub(m, msh(m, last[c], one), half_i);
        if (ntt) or_ntt_forward(&pp->tq[i], t);
        u64 inv = 0;
        or_mod_inv(m, ql->q % m->q, &inv);
        shoup_t s = mk_shoup(m, inv);
        u64 *a = poly + i * n;
        for (u64 c = 0; c < n; ++c) a[c] = msh(m, or_mod_sub(m, a[c], t[c]), s);
    }
    free(t);
    free(last);
}

/* RnsPoly::mod_down (baseconv.rs:459-511). */
void or_mod_down(const or_params *pp, u64 *qpart, u64 l, const u64 *ppart, u64 np, int ntt) {
    u64 n = pp->n;
    u64 *src = malloc(np * n * sizeof(u64));
    memcpy(src, ppart, np * n * sizeof(u64));
    for (u64 t = 0; t < np; ++t) {
        if (ntt) or_ntt_inverse(&pp->tp[t], src + t * n);
        u64 h = pp->mp[t].q >> 1;
        for (u64 c = 0; c < n; ++c) src[t * n + c] = or_mod_add(&pp->mp[t], src[t * n + c], h);
    }
    u64 *conv = malloc((l + 1) * n * sizeof(u64));
    or_convert(pp->p, np, pp->q, l + 1, n, src, conv, 1);
    for (u64 i = 0; i <= l; ++i) {
        const or_mod *m = &pp->mq[i];
        u64 p_mod = 1 % m->q;
        for (u64 t = 0; t < np; ++t) p_mod = or_mod_mul(m, p_mod, pp->p[t] % m->q);
        u64 p_inv = 0;
        or_mod_inv(m, p_mod, &p_inv);
        u64 half = half_mod(pp->mp, np, m);
        u64 *tt = conv + i * n;
        for (u64 c = 0; c < n; ++c) tt[c] = or_mod_sub(m, tt[c], half);
        if (ntt) or_ntt_forward(&pp->tq[i], tt);
        shoup_t s = mk_shoup(m, p_inv);
        u64 *a = qpart + i * n;
        for (u64 c = 0; c < n; ++c) a[c] = msh(m, or_mod_sub(m, a[c], tt[c]), s);
    }
    free(conv);
    free(src);
}

/* ======================= keyswitch.rs ======================= */

/* digit_ranges (keyswitch.rs:68-74). */
u64 or_digit_ranges(const or_params *pp, u64 level, u64 *starts, u64 *ends) {
    u64 cnt = 0;
    for (u64 j = 0; j < pp->dnum; ++j) {
        u64 s = j * pp->alpha, e = (j + 1) * pp->alpha;
        if (e > level + 1) e = level + 1;
        if (s > level) break;
        if (starts) starts[cnt] = s;
        if (ends) ends[cnt] = e;
        cnt++;
    }
    return cnt;
}

/* lift_digit_exact (keyswitch.rs:140-210): writes one QpPoly (l + 1 + np limbs). */
static void lift_digit_exact(const or_params *pp, const u64 *poly, u64 l, u64 ds, u64 de,
                             u64 *out) {
    u64 n = pp->n, np = pp->np, dl = de - ds;
    u64 *src = malloc(dl * n * sizeof(u64));
    for (u64 j = 0; j < dl; ++j) {
        const or_mod *m = &pp->mq[ds + j];
        memcpy(src + j * n, poly + (ds + j) * n, n * sizeof(u64));
        or_ntt_inverse(&pp->tq[ds + j], src + j * n);
        u64 h = m->q >> 1;
        for (u64 c = 0; c < n; ++c) src[j * n + c] = or_mod_add(m, src[j * n + c], h);
    }
    /* targets: Q limbs outside the digit (increasing), then every P prime */
    u64 nt = (l + 1 - dl) + np;
    const or_mod **tm = malloc(nt * sizeof(or_mod *));
    const or_ntt **tt = malloc(nt * sizeof(or_ntt *));
    u64 *dst_limb = malloc(nt * sizeof(u64));
    u64 ti = 0;
    for (u64 i = 0; i <= l; ++i) {
        if (i >= ds && i < de) continue;
        tm[ti] = &pp->mq[i];
        tt[ti] = &pp->tq[i];
        dst_limb[ti++] = i;
    }
    for (u64 t = 0; t < np; ++t) {
        tm[ti] = &pp->mp[t];
        tt[ti] = &pp->tp[t];
        dst_limb[ti++] = l + 1 + t;
    }
    const or_mod *dm = &pp->mq[ds];
    if (dl == 1) {
        for (u64 t = 0; t < nt; ++t) {
            const or_mod *m = tm[t];
            shoup_t one = mk_shoup(m, 1);
            u64 h = half_mod(dm, 1, m);
            u64 *a = out + dst_limb[t] * n;
            for (u64 c = 0; c < n; ++c) a[c] = or_mod_sub(m, msh(m, src[c], one), h);
            or_ntt_forward(tt[t], a);
        }
    } else {
        u64 *to = malloc(nt * sizeof(u64));
        for (u64 t = 0; t < nt; ++t) to[t] = tm[t]->q;
        u64 *conv = malloc(nt * n * sizeof(u64));
        or_convert(pp->q + ds, dl, to, nt, n, src, conv, 1);
        for (u64 t = 0; t < nt; ++t) {
            const or_mod *m = tm[t];
            u64 h = half_mod(dm, dl, m);
            u64 *a = out + dst_limb[t] * n;
            for (u64 c = 0; c < n; ++c) a[c] = or_mod_sub(m, conv[t * n + c], h);
            or_ntt_forward(tt[t], a);
        }
        free(conv);
        free(to);
    }
    for (u64 i = ds; i < de; ++i) memcpy(out + i * n, poly + i * n, n * sizeof(u64));
    free(dst_limb);
    free(tt);
    free(tm);
    free(src);
}

/* decompose (keyswitch.rs:82-124): one task per digit (parallel over digits, par.rs). */
void or_decompose(const or_params *pp, const u64 *poly, u64 l, u64 *digits) {
    u64 starts[64], ends[64];
    u64 nd = or_digit_ranges(pp, l, starts, ends);
    u64 stride = (l + 1 + pp->np) * pp->n;
#pragma omp parallel for schedule(dynamic, 1)
    for (long long j = 0; j < (long long)nd; ++j)
        lift_digit_exact(pp, poly, l, starts[j], ends[j], digits + (u64)j * stride);
}

/* apply_key (keyswitch.rs:230-309): lazy u128 inner product per limb (parallel over limbs),
 * then mod_down of both accumulators. */
void or_apply_key(const or_params *pp, const u64 *digits, u64 nd, u64 l, const u64 *key,
                  u64 kml, u64 *kb, u64 *ka) {
    u64 n = pp->n, np = pp->np;
    u64 dstride = (l + 1 + np) * n;        /* one lifted digit */
    u64 kpoly = (kml + 1 + np) * n;        /* one key polynomial b_j or a_j */
    u64 total = l + 1 + np;
    u64 *acc_b = malloc(total * n * sizeof(u64));
    u64 *acc_a = malloc(total * n * sizeof(u64));
#pragma omp parallel for schedule(dynamic, 1)
    for (long long ii = 0; ii < (long long)total; ++ii) {
        u64 i = (u64)ii;
        int is_p = i > l;
        const or_mod *m = is_p ? &pp->mp[i - l - 1] : &pp->mq[i];
        u64 klimb = is_p ? (kml + 1 + (i - l - 1)) : i; /* limb index inside a key polynomial */
        u128 *sb = calloc(n, sizeof(u128)), *sa = calloc(n, sizeof(u128));
        for (u64 j = 0; j < nd; ++j) {
            const u64 *d = digits + j * dstride + i * n;
            const u64 *b = key + (2 * j) * kpoly + klimb * n;
            const u64 *a = key + (2 * j + 1) * kpoly + klimb * n;
            for (u64 c = 0; c < n; ++c) {
                sb[c] += (u128)d[c] * b[c];
                sa[c] += (u128)d[c] * a[c];
            }
        }
        u64 two64 = (u64)(((u128)1 << 64) % m->q);
        shoup_t s64 = mk_shoup(m, two64), one = mk_shoup(m, 1);
        for (u64 c = 0; c < n; ++c) {
            acc_b[i * n + c] = or_mod_add(m, msh(m, (u64)(sb[c] >> 64), s64), msh(m, (u64)sb[c], one));
            acc_a[i * n + c] = or_mod_add(m, msh(m, (u64)(sa[c] >> 64), s64), msh(m, (u64)sa[c], one));
        }
        free(sb);
        free(sa);
    }
    or_mod_down(pp, acc_b, l, acc_b + (l + 1) * n, np, 1);
    or_mod_down(pp, acc_a, l, acc_a + (l + 1) * n, np, 1);
    memcpy(kb, acc_b, (l + 1) * n * sizeof(u64));
    memcpy(ka, acc_a, (l + 1) * n * sizeof(u64));
    free(acc_b);
    free(acc_a);
}

/* key_switch (keyswitch.rs:317-319). */
void or_key_switch(const or_params *pp, const u64 *poly, u64 l, const u64 *key, u64 kml,
                   u64 *kb, u64 *ka) {
    u64 nd = or_digit_ranges(pp, l, NULL, NULL);
    u64 *digits = malloc(nd * (l + 1 + pp->np) * pp->n * sizeof(u64));
    or_decompose(pp, poly, l, digits);
    or_apply_key(pp, digits, nd, l, key, kml, kb, ka);
    free(digits);
}

/* ======================= evaluator.rs ======================= */

void or_poly_add(const or_params *pp, u64 *a, const u64 *b, u64 limbs) {
    for (u64 i = 0; i < limbs; ++i)
        for (u64 c = 0; c < pp->n; ++c) a[i * pp->n + c] = or_mod_add(&pp->mq[i], a[i * pp->n + c], b[i * pp->n + c]);
}

void or_poly_mul(const or_params *pp, u64 *a, const u64 *b, u64 limbs) {
    for (u64 i = 0; i < limbs; ++i)
        for (u64 c = 0; c < pp->n; ++c) a[i * pp->n + c] = or_mod_mul(&pp->mq[i], a[i * pp->n + c], b[i * pp->n + c]);
}

/* RnsPoly::to_ntt / to_coeff (poly.rs:191-215) on Q limbs 0..limbs with the parameters' tables. */
void or_params_ntt(const or_params *pp, u64 *a, u64 limbs, int forward) {
    for (u64 i = 0; i < limbs; ++i) {
        if (forward)
            or_ntt_forward(&pp->tq[i], a + i * pp->n);
        else
            or_ntt_inverse(&pp->tq[i], a + i * pp->n);
    }
}

/* RnsPoly::raise_from_level0 (poly.rs:298-345): limb 0 (level-0 input `a`, NTT form when ntt)
 * copied; limb i in 1..=top the coefficient x of limb 0 (inverse-transformed first) lifted to
 * x or -(q_0 - x) by x > (q_0 - 1) / 2, reduced mod q_i, transformed forward again. */
void or_raise_from_level0(const or_params *pp, const u64 *a, u64 top, int ntt, u64 *out) {
    const u64 n = pp->n, q0 = pp->mq[0].q, half = q0 / 2;
    u64 *src = malloc(n * sizeof(u64));
    memcpy(src, a, n * sizeof(u64));
    if (ntt) or_ntt_inverse(&pp->tq[0], src);
    memcpy(out, a, n * sizeof(u64));
    for (u64 i = 1; i <= top; ++i) {
        const u64 q = pp->mq[i].q;
        u64 *o = out + i * n;
        for (u64 c = 0; c < n; ++c) {
            const u64 x = src[c];
            if (x > half) {
                const u64 r = (q0 - x) % q;
                o[c] = r == 0 ? 0 : q - r;
            } else {
                o[c] = x % q;
            }
        }
        if (ntt) or_ntt_forward(&pp->tq[i], o);
    }
    free(src);
}

/* RnsPoly::mul_scalar_assign (poly.rs:184-189): limb i times k mod q_i as a Shoup product;
 * then RnsPoly::neg (poly.rs:157-161), x -> 0 - x mod q_i, when neg is set. */
void or_poly_mul_scalar(const or_params *pp, u64 *a, u64 limbs, u64 k, int neg) {
    for (u64 i = 0; i < limbs; ++i) {
        const or_mod *m = &pp->mq[i];
        shoup_t s = mk_shoup(m, k);
        for (u64 c = 0; c < pp->n; ++c) {
            u64 v = msh(m, a[i * pp->n + c], s);
            a[i * pp->n + c] = neg ? or_mod_sub(m, 0, v) : v;
        }
    }
}

/* Evaluator::relinearize (evaluator.rs:441-458). */
void or_relinearize(const or_params *pp, const u64 *c0, const u64 *c1, const u64 *c2, u64 l,
                    const u64 *key, u64 kml, u64 *out0, u64 *out1) {
    u64 sz = (l + 1) * pp->n;
    u64 *kb = malloc(sz * sizeof(u64)), *ka = malloc(sz * sizeof(u64));
    or_key_switch(pp, c2, l, key, kml, kb, ka);
    memcpy(out0, c0, sz * sizeof(u64));
    memcpy(out1, c1, sz * sizeof(u64));
    or_poly_add(pp, out0, kb, l + 1);
    or_poly_add(pp, out1, ka, l + 1);
    free(kb);
    free(ka);
}

/* finish_galois (evaluator.rs:800-804). */
static void finish_galois(const or_params *pp, const u64 *c0, u64 l, u64 g, const u64 *kb,
                          const u64 *ka, u64 *out0, u64 *out1) {
    u64 sz = (l + 1) * pp->n;
    or_automorphism_ntt(pp->n, g, c0, out0, l + 1);
    or_poly_add(pp, out0, kb, l + 1);
    memcpy(out1, ka, sz * sizeof(u64));
}

/* Evaluator::apply_galois (evaluator.rs:640-647). */
void or_apply_galois(const or_params *pp, const u64 *c0, const u64 *c1, u64 l, u64 g,
                     const u64 *key, u64 kml, u64 *out0, u64 *out1) {
    u64 sz = (l + 1) * pp->n;
    u64 *d = malloc(sz * sizeof(u64)), *kb = malloc(sz * sizeof(u64)), *ka = malloc(sz * sizeof(u64));
    or_automorphism_ntt(pp->n, g, c1, d, l + 1);
    or_key_switch(pp, d, l, key, kml, kb, ka);
    finish_galois(pp, c0, l, g, kb, ka, out0, out1);
    free(d);
    free(kb);
    free(ka);
}

/* Evaluator::hoisted_rotate (evaluator.rs:572-611): decompose c1 once, then per step the
 * automorphism of both parts of every digit, apply_key, finish_galois (parallel over steps). */
void or_hoisted_galois(const or_params *pp, const u64 *c0, const u64 *c1, u64 l, u64 nsteps,
                       const u64 *elts, const u64 *const *keys, u64 kml, u64 *outs) {
    u64 n = pp->n, np = pp->np, sz = (l + 1) * n;
    u64 nd = or_digit_ranges(pp, l, NULL, NULL);
    u64 dsz = (l + 1 + np) * n;
    u64 *digits = malloc(nd * dsz * sizeof(u64));
    or_decompose(pp, c1, l, digits);
#pragma omp parallel for schedule(dynamic, 1)
    for (long long s = 0; s < (long long)nsteps; ++s) {
        u64 *rot = malloc(nd * dsz * sizeof(u64));
        for (u64 j = 0; j < nd; ++j) or_automorphism_ntt(n, elts[s], digits + j * dsz, rot + j * dsz, l + 1 + np);
        u64 *kb = malloc(sz * sizeof(u64)), *ka = malloc(sz * sizeof(u64));
        or_apply_key(pp, rot, nd, l, keys[s], kml, kb, ka);
        finish_galois(pp, c0, l, elts[s], kb, ka, outs + (2 * (u64)s) * sz, outs + (2 * (u64)s + 1) * sz);
        free(kb);
        free(ka);
        free(rot);
    }
    free(digits);
}

/* Evaluator::rescale (evaluator.rs:498-509): rescale_by_last on every component. */
void or_rescale_ct(const or_params *pp, u64 *ct, u64 ncomp, u64 l) {
    for (u64 c = 0; c < ncomp; ++c) or_rescale_by_last(pp, ct + c * (l + 1) * pp->n, l, 1);
}

/* Evaluator::mul_no_relin (evaluator.rs:422-438). */
void or_mul_no_relin(const or_params *pp, const u64 *a0, const u64 *a1, const u64 *b0,
                     const u64 *b1, u64 l, u64 *d0, u64 *d1, u64 *d2) {
    u64 sz = (l + 1) * pp->n;
    u64 *t = malloc(sz * sizeof(u64));
    memcpy(d0, a0, sz * sizeof(u64));
    or_poly_mul(pp, d0, b0, l + 1);
    memcpy(d1, a0, sz * sizeof(u64));
    or_poly_mul(pp, d1, b1, l + 1);
    memcpy(t, a1, sz * sizeof(u64));
    or_poly_mul(pp, t, b0, l + 1);
    or_poly_add(pp, d1, t, l + 1);
    memcpy(d2, a1, sz * sizeof(u64));
    or_poly_mul(pp, d2, b1, l + 1);
    free(t);
}

/* ======================= seeded inputs ======================= */

static inline u64 mix64(u64 z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

/* Counter-based uniform residue below q by rejection (no modulo bias, as Rng::below,
 * crates/ckks-train-core/src/sampler.rs:73-82, but with a stateless generator so a device can
 * produce the same words in parallel). */
u64 or_uniform_word(u64 seed, u64 stream, u64 index, u64 q) {
    u64 key = mix64(seed ^ mix64(stream + 0x9e3779b97f4a7c15ULL));
    u64 zone = (UINT64_MAX / q) * q;
    u64 x = 0;
    for (u64 a = 0; a < 64; ++a) {
        x = mix64(key + (index * 64 + a) * 0x9e3779b97f4a7c15ULL);
        if (x < zone) break;
    }
    return x % q;
}

void or_fill_uniform(u64 *out, u64 n, u64 q, u64 seed, u64 stream) {
    for (u64 i = 0; i < n; ++i) out[i] = or_uniform_word(seed, stream, i, q);
}

/* ======================= threading ======================= */

void or_set_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n > 0 ? n : 1);
#else
    (void)n;
#endif
}

int or_get_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
