/*
 * ffrs_oracle.c -- CPU restatement of libffrs' RS256 / RSi16 encode and repair.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path (see ffrs_oracle.h).
 * It is written from the reference's behaviour, step for step where the
 * reference's exact integer sequence decides the result (RSi16 residue
 * arithmetic, Sugiyama with cyclic NTT products, BM status rules), and in
 * textbook form where the result is fixed by the mathematics (RS256 parity).
 * Paths cited below are relative to the reference repository root.
 */
#include "ffrs_oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* =====================================================================
 * GF(2^8)
 * ===================================================================== */

/* galois.hpp:624-651: shift-and-add multiplication modulo the polynomial. */
static uint8_t gf256_mul_serial(uint8_t a, uint8_t b, uint8_t poly1)
{
    uint8_t r = 0;
    for (int i = 7; i >= 0; --i) {
        r = (r & 0x80) ? (uint8_t)((r << 1) ^ poly1) : (uint8_t)(r << 1);
        if (a & (1u << i))
            r ^= b;
    }
    return r;
}

int orc_gf256_tables(unsigned primitive, unsigned polynomial, uint8_t exp_t[256], uint8_t log_t[256])
{
    /* galois.hpp:133-144 */
    uint8_t poly1 = (uint8_t)(polynomial & 0xff);
    memset(log_t, 0, 256);
    uint8_t x = 1;
    for (int i = 0; i < 256; ++i) {
        exp_t[i] = x;
        log_t[x] = (uint8_t)i;
        x = gf256_mul_serial(x, (uint8_t)primitive, poly1);
    }
    /* generator check: alpha^i distinct for i < 255 */
    uint8_t seen[256] = {0};
    for (int i = 0; i < 255; ++i) {
        if (seen[exp_t[i]] || exp_t[i] == 0)
            return -1;
        seen[exp_t[i]] = 1;
    }
    return 0;
}

struct orc_rs256 {
    unsigned ecc;
    uint8_t poly1;
    uint8_t exp_t[256];
    uint8_t log_t[256];
    uint8_t gen[256];   /* generator, highest degree first, ecc+1 entries */
    uint8_t roots[256]; /* generator_roots[i] = alpha^i */
};

/* galois.hpp:653-667 */
static inline uint8_t rs_mul(const orc_rs256 *rs, uint8_t a, uint8_t b)
{
    if (a == 0 || b == 0)
        return 0;
    unsigned r = (unsigned)rs->log_t[a] + rs->log_t[b];
    if (r >= 255)
        r -= 255;
    return rs->exp_t[r];
}

/* galois.hpp:98-101: inv(a) = exp[255 - log a]; inv(0) = exp[255] = 1. */
static inline uint8_t rs_inv(const orc_rs256 *rs, uint8_t a)
{
    return rs->exp_t[255 - rs->log_t[a]];
}

/* galois.hpp:103-114: div(0, b) = 0; div(a, 0) = a. */
static inline uint8_t rs_div(const orc_rs256 *rs, uint8_t a, uint8_t b)
{
    if (a == 0)
        return 0;
    unsigned r = (unsigned)rs->log_t[a] + 255 - rs->log_t[b];
    if (r >= 255)
        r -= 255;
    return rs->exp_t[r];
}

orc_rs256 *orc_rs256_new(unsigned ecc_len, unsigned primitive, unsigned polynomial)
{
    if (ecc_len < 1 || ecc_len > 254)
        return NULL;
    orc_rs256 *rs = calloc(1, sizeof(*rs));
    rs->ecc = ecc_len;
    rs->poly1 = (uint8_t)(polynomial & 0xff);
    if (orc_gf256_tables(primitive, polynomial, rs->exp_t, rs->log_t) != 0 || (rs->poly1 & 0x80)) {
        free(rs);
        return NULL;
    }
    /* reed_solomon.hpp:65-86: g(x) = prod_{i<ecc} (x - alpha^i), highest first */
    uint8_t g[256] = {0};
    size_t len = 1;
    g[0] = 1;
    for (unsigned i = 0; i < ecc_len; ++i) {
        uint8_t root = rs->exp_t[i];
        rs->roots[i] = root;
        uint8_t ng[256] = {0};
        for (size_t a = 0; a < len; ++a) {
            ng[a] ^= g[a];
            ng[a + 1] ^= rs_mul(rs, g[a], root);
        }
        len += 1;
        memcpy(g, ng, len);
    }
    memcpy(rs->gen, g, ecc_len + 1);
    return rs;
}

void orc_rs256_free(orc_rs256 *rs) { free(rs); }

void orc_rs256_generator(const orc_rs256 *rs, uint8_t *generator)
{
    memcpy(generator, rs->gen, rs->ecc + 1);
}

void orc_rs256_encode_block(const orc_rs256 *rs, const uint8_t *msg, size_t len, uint8_t *ecc_out)
{
    /* Textbook LFSR: remainder of x^ecc * M(x) by g(x), rem[0] = coeff of x^(ecc-1). */
    unsigned e = rs->ecc;
    uint8_t rem[256] = {0};
    for (size_t i = 0; i < len; ++i) {
        uint8_t f = rem[0] ^ msg[i];
        memmove(rem, rem + 1, e - 1);
        rem[e - 1] = 0;
        if (f)
            for (unsigned j = 0; j < e; ++j)
                rem[j] ^= rs_mul(rs, f, rs->gen[j + 1]);
    }
    memcpy(ecc_out, rem, e);
}

size_t orc_rs256_encode(const orc_rs256 *rs, size_t message_len, const uint8_t *buf, size_t size, uint8_t *out)
{
    /* pyrs256.hpp:99-127 */
    if (size == 0 || message_len == 0)
        return 0;
    size_t full = size / message_len;
    size_t tail = size - full * message_len;
    for (size_t b = 0; b < full; ++b)
        orc_rs256_encode_block(rs, buf + b * message_len, message_len, out + b * rs->ecc);
    if (tail)
        orc_rs256_encode_block(rs, buf + size - tail, tail, out + full * rs->ecc);
    return (full + (tail ? 1 : 0)) * rs->ecc;
}

void orc_rs256_synds(const orc_rs256 *rs, const uint8_t *msg, size_t k, const uint8_t *ecc, uint8_t *synds)
{
    /* reed_solomon.hpp:424-431: Horner over message bytes then parity bytes. */
    for (unsigned i = 0; i < rs->ecc; ++i) {
        uint8_t x = rs->roots[i], r = 0;
        for (size_t j = 0; j < k; ++j)
            r = rs_mul(rs, r, x) ^ msg[j];
        for (unsigned j = 0; j < rs->ecc; ++j)
            r = rs_mul(rs, r, x) ^ ecc[j];
        synds[i] = r;
    }
}

/* galois.hpp:501-507, Horner, coefficients highest first. */
static uint8_t rs_poly_eval(const orc_rs256 *rs, const uint8_t *p, size_t n, uint8_t x)
{
    uint8_t r = 0;
    for (size_t i = 0; i < n; ++i)
        r = rs_mul(rs, r, x) ^ p[i];
    return r;
}

/* reed_solomon.hpp:655-702. err_poly is reversed with capacity ecc:
 * err_poly[ecc-1-i] = Lambda_i; shifting drops terms of degree >= ecc. */
static size_t rs_berlekamp_massey(const orc_rs256 *rs, const uint8_t *synds, uint8_t *err_poly)
{
    size_t E = rs->ecc;
    uint8_t prev[256] = {0}, temp[256];
    memset(err_poly, 0, E);
    prev[E - 1] = 1;
    err_poly[E - 1] = 1;
    size_t L = 0, m = 1;
    uint8_t b = 1;
    for (size_t n = 0; n < E; ++n) {
        uint8_t d = synds[n];
        for (size_t i = 1; i < L + 1; ++i)
            d ^= rs_mul(rs, err_poly[E - 1 - i], synds[n - i]);
        if (d == 0) {
            m += 1;
        } else if (2 * L <= n) {
            memcpy(temp, err_poly, E);
            /* poly_shift(prev, E, m), poly_scale, poly_sub */
            for (size_t i = 0; i < E; ++i)
                prev[i] = (i + m < E) ? prev[i + m] : 0;
            uint8_t s = rs_div(rs, d, b);
            for (size_t i = 0; i < E; ++i)
                err_poly[i] ^= rs_mul(rs, s, prev[i]);
            L = n + 1 - L;
            memcpy(prev, temp, E);
            b = d;
            m = 1;
        } else {
            for (size_t i = 0; i < E; ++i)
                temp[i] = (i + m < E) ? prev[i + m] : 0;
            uint8_t s = rs_div(rs, d, b);
            for (size_t i = 0; i < E; ++i)
                err_poly[i] ^= rs_mul(rs, s, temp[i]);
            m += 1;
        }
    }
    return L;
}

int orc_rs256_decode(const orc_rs256 *rs, uint8_t *msg, size_t k, uint8_t *ecc)
{
    size_t E = rs->ecc, n = k + E;
    uint8_t synds[256];
    orc_rs256_synds(rs, msg, k, ecc, synds);
    int any = 0;
    for (size_t i = 0; i < E; ++i)
        any |= synds[i];
    if (!any)
        return 1;

    uint8_t err_poly[256];
    size_t L = rs_berlekamp_massey(rs, synds, err_poly);
    if (L > E / 2)
        return 0;

    /* reed_solomon.hpp:516-538: Chien search over positions [0, n), ascending. */
    const uint8_t *lam = &err_poly[E - L - 1]; /* Lambda_L .. Lambda_0 */
    uint8_t pos[256];
    size_t count = 0;
    for (size_t p = 0; p < n; ++p) {
        uint8_t x = rs_inv(rs, rs->exp_t[p]);
        if (rs_poly_eval(rs, lam, L + 1, x) == 0)
            pos[count++] = (uint8_t)p;
    }
    if (count != L)
        return 0;

    /* reed_solomon.hpp:704-743: Omega = S(x) Lambda(x) mod x^ecc (both highest
     * first), leading zeros stripped; Lambda' keeps the odd terms. */
    uint8_t prod[512] = {0};
    for (size_t i = 0; i < E; ++i)          /* synds reversed: S_{E-1} .. S_0 */
        for (size_t j = 0; j < L + 1; ++j)
            prod[i + j] ^= rs_mul(rs, synds[E - 1 - i], lam[j]);
    size_t psize = E + L;
    size_t begin = psize - (E + 1) + 1; /* ex_synth_div quotient length */
    while (prod[begin] == 0)
        begin++;
    const uint8_t *omega = &prod[begin];
    size_t omega_len = psize - begin;

    uint8_t lam_d[256];
    memcpy(lam_d, lam, L + 1);
    for (size_t i = 1; i < L + 1; ++i) /* gf_poly_deriv_pw2, galois.hpp:725-734 */
        lam_d[L + 1 - i] = (i & 1) ? lam_d[L - i] : 0;
    const uint8_t *deriv = lam_d + 1; /* L coefficients */

    uint8_t mag[256];
    for (size_t i = 0; i < L; ++i) {
        uint8_t xi = rs->exp_t[pos[i]];
        uint8_t xi_inv = rs_inv(rs, xi);
        uint8_t nv = rs_poly_eval(rs, omega, omega_len, xi_inv);
        uint8_t dv = rs_poly_eval(rs, deriv, L, xi_inv);
        mag[i] = rs_mul(rs, xi, rs_div(rs, nv, dv));
    }
    for (size_t i = 0; i < L; ++i) {
        size_t p = n - 1 - pos[i];
        if (p < k)
            msg[p] ^= mag[i];
        else
            ecc[p - k] ^= mag[i];
    }
    return 1;
}

/* ---------------------------------------------------------------- threads */

typedef struct {
    void (*fn)(void *ctx, size_t begin, size_t end);
    void *ctx;
    size_t begin, end;
} orc_job;

static void *orc_job_run(void *p)
{
    orc_job *j = p;
    j->fn(j->ctx, j->begin, j->end);
    return NULL;
}

static void orc_parallel(size_t count, int threads, void (*fn)(void *, size_t, size_t), void *ctx)
{
    if (threads < 1)
        threads = 1;
    if (threads > 256)
        threads = 256;
    if ((size_t)threads > count)
        threads = count ? (int)count : 1;
    if (threads == 1) {
        fn(ctx, 0, count);
        return;
    }
    pthread_t tid[256];
    orc_job jobs[256];
    size_t per = (count + threads - 1) / threads;
    for (int t = 0; t < threads; ++t) {
        jobs[t].fn = fn;
        jobs[t].ctx = ctx;
        jobs[t].begin = t * per < count ? t * per : count;
        jobs[t].end = (t + 1) * per < count ? (t + 1) * per : count;
        pthread_create(&tid[t], NULL, orc_job_run, &jobs[t]);
    }
    for (int t = 0; t < threads; ++t)
        pthread_join(tid[t], NULL);
}

typedef struct {
    const orc_rs256 *rs;
    uint8_t *msgs;
    uint8_t *eccs;
    size_t k;
    uint8_t *status;
} rs256_batch;

static void rs256_enc_range(void *p, size_t b, size_t e)
{
    rs256_batch *c = p;
    for (size_t i = b; i < e; ++i)
        orc_rs256_encode_block(c->rs, c->msgs + i * c->k, c->k, c->eccs + i * c->rs->ecc);
}

static void rs256_dec_range(void *p, size_t b, size_t e)
{
    rs256_batch *c = p;
    for (size_t i = b; i < e; ++i)
        c->status[i] = (uint8_t)orc_rs256_decode(c->rs, c->msgs + i * c->k, c->k, c->eccs + i * c->rs->ecc);
}

void orc_rs256_encode_batch(const orc_rs256 *rs, const uint8_t *msgs, size_t count, size_t k,
                            uint8_t *ecc_out, int threads)
{
    rs256_batch c = {rs, (uint8_t *)msgs, ecc_out, k, NULL};
    orc_parallel(count, threads, rs256_enc_range, &c);
}

void orc_rs256_decode_batch(const orc_rs256 *rs, uint8_t *msgs, uint8_t *eccs, size_t count, size_t k,
                            uint8_t *status, int threads)
{
    rs256_batch c = {rs, msgs, eccs, k, status};
    orc_parallel(count, threads, rs256_dec_range, &c);
}

/* =====================================================================
 * GF(65537)
 * ===================================================================== */

#define P16 0x10001u

/* --- scalar arithmetic, the integral branches of galois.hpp:258-395 --- */
static inline uint32_t s_add(uint32_t a, uint32_t b)
{
    uint32_t r = a + b;
    if (r >= P16)
        r -= P16;
    return r;
}
static inline uint32_t s_sub(uint32_t a, uint32_t b)
{
    uint32_t r = a - b;
    if (r >= 0x80000000u)
        r += P16;
    return r;
}
static inline uint32_t s_neg(uint32_t a) { return a == 0 ? 0 : P16 - a; }
static inline uint32_t s_mul(uint32_t a, uint32_t b)
{
    uint32_t r = a * b;
    if (r == 0 && a && b)
        return 1;
    r = (r & 0xffff) - (r >> 16);
    if (r >= 0x80000000u)
        r += P16;
    return r;
}
static inline uint32_t s_mul_residue(uint32_t a, uint32_t b)
{
    uint32_t r = a * b;
    if (r == 0 && a && b)
        return 1;
    return (r & 0xffff) - (r >> 16);
}
static inline uint32_t s_mod_p(uint32_t x)
{
    uint32_t r = (x & 0xffff) - (x >> 16) - (x >> 31);
    if (r >= 0x80000000u)
        r += P16;
    return r;
}

/* --- the vector branches of the same routines (W = 4/8/16 lanes) --- */
static inline uint32_t v_add(uint32_t a, uint32_t b)
{
    uint32_t r = a + b;
    return r - ((r >= P16) ? P16 : 0);
}
static inline uint32_t v_sub(uint32_t a, uint32_t b) { return a - b + ((b > a) ? P16 : 0); }
static inline uint32_t v_mul(uint32_t a, uint32_t b)
{
    uint32_t r = a * b;
    uint32_t ovf = (r == 0 && (a & b) != 0) ? 1u : 0u;
    uint32_t lo = r & 0xffff, hi = r >> 16;
    return lo - hi + ovf + ((hi > lo) ? P16 : 0);
}
static inline uint32_t v_mul_residue(uint32_t a, uint32_t b)
{
    uint32_t r = a * b;
    uint32_t ovf = (r == 0 && (a & b) != 0) ? 1u : 0u;
    return (r & 0xffff) - (r >> 16) + ovf;
}
static inline uint32_t v_mod_p(uint32_t x)
{
    uint32_t lo = x & 0xffff, hi = x >> 16;
    return lo - hi - (x >> 31) + ((hi >= lo) ? P16 : 0);
}

static uint16_t rbo16(uint16_t b)
{
    b = (uint16_t)(((b >> 1) & 0x5555) | ((b & 0x5555) << 1));
    b = (uint16_t)(((b >> 2) & 0x3333) | ((b & 0x3333) << 2));
    b = (uint16_t)(((b >> 4) & 0x0f0f) | ((b & 0x0f0f) << 4));
    b = (uint16_t)((b >> 8) | (b << 8));
    return b;
}

struct orc_rsi16 {
    size_t n, ecc, k, N;
    unsigned rbo_shift;
    uint32_t primitive, root;
    uint32_t N_inv, ecc_inv, ecc2_inv;
    uint32_t *exp_t, *log_t;       /* 65537 entries each */
    uint32_t *roots_ntt, *roots_i_ntt;
    uint32_t *roots_ecc, *roots_i_ecc;
    uint32_t *roots_ecc2, *roots_i_ecc2;
    uint32_t *pntt_shift, *pintt_shift, *rbo_ecc;
    uint32_t *mix;
};

/* galois.hpp:98-127 over the 65537-entry tables */
static inline uint32_t g_inv(const orc_rsi16 *rs, uint32_t a) { return rs->exp_t[65536 - rs->log_t[a]]; }
static inline uint32_t g_div(const orc_rsi16 *rs, uint32_t a, uint32_t b)
{
    if (a == 0)
        return 0;
    uint32_t r = rs->log_t[a] + 65536 - rs->log_t[b];
    if (r >= 65536)
        r -= 65536;
    return rs->exp_t[r];
}
static inline uint32_t g_pow(const orc_rsi16 *rs, uint32_t a, uint32_t e)
{
    return rs->exp_t[(uint32_t)(rs->log_t[a] * e) % 65536u];
}

static uint16_t rs_rbo(const orc_rsi16 *rs, uint32_t b) { return (uint16_t)(rbo16((uint16_t)b) >> rs->rbo_shift); }

void orc_rsi16_free(orc_rsi16 *rs)
{
    if (!rs)
        return;
    free(rs->exp_t);
    free(rs->log_t);
    free(rs->roots_ntt);
    free(rs->roots_i_ntt);
    free(rs->roots_ecc);
    free(rs->roots_i_ecc);
    free(rs->roots_ecc2);
    free(rs->roots_i_ecc2);
    free(rs->pntt_shift);
    free(rs->pintt_shift);
    free(rs->rbo_ecc);
    free(rs->mix);
    free(rs);
}

orc_rsi16 *orc_rsi16_new(size_t block_len, size_t ecc_len, uint32_t primitive)
{
    if (ecc_len < 2 || (ecc_len & (ecc_len - 1)) || block_len % ecc_len || block_len <= ecc_len ||
        block_len > 65536)
        return NULL;
    orc_rsi16 *rs = calloc(1, sizeof(*rs));
    rs->n = block_len;
    rs->ecc = ecc_len;
    rs->k = block_len - ecc_len;
    rs->primitive = primitive;

    /* galois.hpp:133-144 with gf_mul_mod (galois.hpp:239-250) */
    rs->exp_t = calloc(65537, 4);
    rs->log_t = calloc(65537, 4);
    uint32_t x = 1;
    for (uint32_t i = 0; i < 65537; ++i) {
        rs->exp_t[i] = x;
        rs->log_t[x] = i;
        uint32_t r;
        if (__builtin_mul_overflow(x, primitive, &r))
            r = 1;
        x = r % P16;
    }
    /* pygfi16.hpp:46-49; a primitive above 65536 has no log-table entry */
    if (primitive > 65536 || !(g_pow(rs, primitive, 32768) == 65536 && g_pow(rs, primitive, 65536) == 1))
        goto fail;

    /* ntt.hpp:55-123 */
    {
        size_t N = 2;
        while (N < block_len)
            N <<= 1;
        rs->N = N;
        rs->rbo_shift = 16 - (unsigned)__builtin_ctzl(N);
        rs->N_inv = g_inv(rs, (uint32_t)N);
        rs->ecc_inv = g_inv(rs, (uint32_t)ecc_len);
        rs->ecc2_inv = g_inv(rs, (uint32_t)(ecc_len * 2));
        rs->root = g_pow(rs, primitive, (uint32_t)(65536 / N));
        if (g_pow(rs, rs->root, (uint32_t)N) != 1 || g_pow(rs, rs->root, (uint32_t)(N / 2)) != 65536)
            goto fail;
        uint32_t root_i = g_inv(rs, rs->root);
        rs->roots_ntt = malloc(N * 4);
        rs->roots_i_ntt = malloc(N * 4);
        for (size_t i = 0; i < N; ++i) {
            rs->roots_ntt[i] = g_pow(rs, rs->root, (uint32_t)i);
            rs->roots_i_ntt[i] = g_pow(rs, root_i, (uint32_t)i);
        }
        rs->rbo_ecc = malloc(ecc_len * 4);
        for (size_t i = 0; i < ecc_len; ++i)
            rs->rbo_ecc[i] = rs_rbo(rs, (uint32_t)(i * (N / ecc_len)));
        uint32_t er = g_pow(rs, primitive, (uint32_t)(65536 / ecc_len));
        if (g_pow(rs, er, (uint32_t)ecc_len) != 1 || g_pow(rs, er, (uint32_t)(ecc_len / 2)) != 65536)
            goto fail;
        uint32_t er_i = g_inv(rs, er);
        rs->roots_ecc = malloc(ecc_len * 4);
        rs->roots_i_ecc = malloc(ecc_len * 4);
        for (size_t i = 0; i < ecc_len; ++i) {
            rs->roots_ecc[i] = g_pow(rs, er, (uint32_t)i);
            rs->roots_i_ecc[i] = g_pow(rs, er_i, (uint32_t)i);
        }
        uint32_t e2 = g_pow(rs, primitive, (uint32_t)(65536 / (ecc_len * 2)));
        if (g_pow(rs, e2, (uint32_t)(ecc_len * 2)) != 1 || g_pow(rs, e2, (uint32_t)ecc_len) != 65536)
            goto fail;
        uint32_t e2_i = g_inv(rs, e2);
        rs->roots_ecc2 = malloc(ecc_len * 8);
        rs->roots_i_ecc2 = malloc(ecc_len * 8);
        for (size_t i = 0; i < ecc_len * 2; ++i) {
            rs->roots_ecc2[i] = g_pow(rs, e2, (uint32_t)i);
            rs->roots_i_ecc2[i] = g_pow(rs, e2_i, (uint32_t)i);
        }
        rs->pntt_shift = malloc(block_len * 4);
        for (size_t i = 0; i < block_len; ++i) {
            size_t blk = (i / ecc_len) * ecc_len, j = i - blk;
            rs->pntt_shift[i] = g_pow(rs, rs->root, (uint32_t)(j * rs_rbo(rs, (uint32_t)blk)));
        }
        rs->pintt_shift = malloc(N * 4);
        for (size_t i = 0; i < N; ++i) {
            size_t blk = (i / ecc_len) * ecc_len, j = i - blk;
            rs->pintt_shift[i] =
                g_div(rs, g_pow(rs, root_i, (uint32_t)(j * rs_rbo(rs, (uint32_t)blk))), (uint32_t)N);
        }
        /* rsi16v_impl.hpp:68-86 */
        rs->mix = malloc(ecc_len * 4);
        uint32_t mw = rs->roots_ntt[rs_rbo(rs, (uint32_t)(block_len - ecc_len))];
        uint32_t mw_i = g_inv(rs, mw);
        for (size_t i = 0; i < ecc_len; ++i)
            rs->mix[i] = s_neg(g_div(rs, g_pow(rs, mw_i, (uint32_t)i), (uint32_t)ecc_len));
    }
    return rs;
fail:
    orc_rsi16_free(rs);
    return NULL;
}

uint32_t orc_rsi16_root(const orc_rsi16 *rs) { return rs->root; }
size_t orc_rsi16_ntt_len(const orc_rsi16 *rs) { return rs->N; }

size_t orc_rsi16_table(const orc_rsi16 *rs, int which, uint32_t *out, size_t cap)
{
    const uint32_t *src = NULL;
    size_t len = 0;
    switch (which) {
    case 0: src = rs->exp_t; len = 65537; break;
    case 1: src = rs->log_t; len = 65537; break;
    case 2: src = rs->roots_ntt; len = rs->N; break;
    case 3: src = rs->pntt_shift; len = rs->n; break;
    case 4: src = rs->pintt_shift; len = rs->N; break;
    case 5: src = rs->mix; len = rs->ecc; break;
    case 6: src = rs->roots_ecc; len = rs->ecc; break;
    case 7: src = rs->roots_i_ecc; len = rs->ecc; break;
    default: return 0;
    }
    if (out)
        memcpy(out, src, (len < cap ? len : cap) * 4);
    return len;
}

/* ---------------------------------------------------------------- transforms */

/* ntt.hpp:225-249 (RBO in, natural out), full reduction, scalar arithmetic */
static void ct_butterfly(const uint32_t *roots, uint32_t *blk, size_t len)
{
    for (size_t stride = 1, ef = len >> 1; stride < len; stride *= 2, ef >>= 1)
        for (size_t start = 0; start < len; start += stride * 2) {
            uint32_t a = blk[start], b = blk[start + stride];
            blk[start] = s_add(a, b);
            blk[start + stride] = s_sub(a, b);
            for (size_t i = start + 1; i < start + stride; ++i) {
                uint32_t w = roots[ef * (i - start)];
                uint32_t a2 = blk[i], m = s_mul(blk[i + stride], w);
                blk[i] = s_add(a2, m);
                blk[i + stride] = s_sub(a2, m);
            }
        }
}

/* ntt.hpp:280-302 (natural in, RBO out); `vec` picks the SIMD-lane arithmetic. */
static void gs_butterfly(const uint32_t *roots, uint32_t *blk, size_t len, int vec)
{
    for (size_t stride = len / 2, ef = 0; stride > 0; stride /= 2, ef += 1)
        for (size_t start = 0; start < len; start += stride * 2) {
            uint32_t a = blk[start], b = blk[start + stride];
            blk[start] = vec ? v_add(a, b) : s_add(a, b);
            blk[start + stride] = vec ? v_sub(a, b) : s_sub(a, b);
            for (size_t i = start + 1; i < start + stride; ++i) {
                uint32_t w = roots[(i - start) << ef];
                uint32_t a2 = blk[i], b2 = blk[i + stride];
                blk[i] = vec ? v_add(a2, b2) : s_add(a2, b2);
                blk[i + stride] = vec ? v_mul(v_sub(a2, b2), w) : s_mul(s_sub(a2, b2), w);
            }
        }
}

/* ntt.hpp:252-275, lazy residue arithmetic */
static void ct_butterfly_residue(const uint32_t *roots, uint32_t *blk, size_t len, int vec)
{
    for (size_t stride = 1, ef = len >> 1; stride < len; stride *= 2, ef >>= 1)
        for (size_t start = 0; start < len; start += stride * 2) {
            uint32_t a = blk[start], b = blk[start + stride];
            blk[start] = a + b;
            blk[start + stride] = a - b;
            for (size_t i = start + 1; i < start + stride; ++i) {
                uint32_t w = roots[ef * (i - start)];
                uint32_t a2 = blk[i], b2 = blk[i + stride];
                uint32_t m = vec ? v_mul_residue(v_mod_p(b2), w) : s_mul_residue(s_mod_p(b2), w);
                blk[i] = a2 + m;
                blk[i + stride] = a2 - m;
            }
        }
}

/* ntt.hpp:569-579 */
static void nttr(const orc_rsi16 *rs, uint32_t *b) { gs_butterfly(rs->roots_ecc, b, rs->ecc, 0); }
static void inttr(const orc_rsi16 *rs, uint32_t *b)
{
    ct_butterfly(rs->roots_i_ecc, b, rs->ecc);
    for (size_t i = 0; i < rs->ecc; ++i)
        b[i] = s_mul(b[i], rs->ecc_inv);
}

/* ---------------------------------------------------------------- encode */

static void rsi16_encode_one(const orc_rsi16 *rs, const uint16_t *msg, uint16_t *out, int vec, uint32_t *blk)
{
    size_t E = rs->ecc, nb = rs->n / E;
    for (size_t i = 0; i < rs->k; ++i)
        blk[i] = msg[i];
    /* NTT::pntt_message_residue, ntt.hpp:180-199 */
    ct_butterfly_residue(rs->roots_ecc, blk, E, vec);
    for (size_t j = 0; j < E; ++j)
        blk[j] = vec ? v_mul_residue(v_mod_p(blk[j]), rs->pntt_shift[j])
                     : s_mul_residue(s_mod_p(blk[j]), rs->pntt_shift[j]);
    for (size_t i = 1; i + 1 < nb; ++i) {
        uint32_t *p = blk + i * E;
        ct_butterfly_residue(rs->roots_ecc, p, E, vec);
        for (size_t j = 0; j < E; ++j)
            p[j] = vec ? v_mul_residue(v_mod_p(p[j]), rs->pntt_shift[i * E + j])
                       : s_mul_residue(s_mod_p(p[j]), rs->pntt_shift[i * E + j]);
        for (size_t j = 0; j < E; ++j)
            blk[j] = blk[j] + p[j];
    }
    /* RSi16v::_mix_ecc_residue, rsi16v_impl.hpp:108-119 */
    for (size_t j = 0; j < E; ++j)
        blk[j] = vec ? v_mul(v_mod_p(blk[j]), rs->mix[j]) : s_mul(s_mod_p(blk[j]), rs->mix[j]);
    gs_butterfly(rs->roots_i_ecc, blk, E, vec);
    for (size_t j = 0; j < E; ++j)
        out[j] = (uint16_t)(blk[j] & 0xffff);
}

typedef struct {
    const orc_rsi16 *rs;
    uint16_t *msgs, *eccs;
    uint8_t *status;
    int vec;
} rsi16_batch;

static void rsi16_enc_range(void *p, size_t b, size_t e)
{
    rsi16_batch *c = p;
    uint32_t *blk = malloc(c->rs->n * 4);
    for (size_t i = b; i < e; ++i)
        rsi16_encode_one(c->rs, c->msgs + i * c->rs->k, c->eccs + i * c->rs->ecc, c->vec, blk);
    free(blk);
}

void orc_rsi16_encode(const orc_rsi16 *rs, const uint16_t *msgs, size_t count, uint16_t *ecc_out,
                      int vector_lanes, int threads)
{
    rsi16_batch c = {rs, (uint16_t *)msgs, ecc_out, NULL, vector_lanes != 0};
    orc_parallel(count, threads, rsi16_enc_range, &c);
}

/* ---------------------------------------------------------------- repair */

/* NTT::pntt, ntt.hpp:133-152, in place over blk[0..n); result in blk[0..ecc). */
static void pntt(const orc_rsi16 *rs, uint32_t *blk)
{
    size_t E = rs->ecc, nb = rs->n / E;
    ct_butterfly(rs->roots_ecc, blk, E);
    for (size_t j = 0; j < E; ++j)
        blk[j] = s_mul(blk[j], rs->pntt_shift[j]);
    for (size_t i = 1; i < nb; ++i) {
        uint32_t *p = blk + i * E;
        ct_butterfly(rs->roots_ecc, p, E);
        for (size_t j = 0; j < E; ++j)
            p[j] = s_mul(p[j], rs->pntt_shift[i * E + j]);
        for (size_t j = 0; j < E; ++j)
            blk[j] = s_add(blk[j], p[j]);
    }
}

void orc_rsi16_synds(const orc_rsi16 *rs, const uint16_t *msg, const uint16_t *ecc, uint32_t *synds)
{
    uint32_t *blk = malloc(rs->n * 4);
    for (size_t i = 0; i < rs->k; ++i)
        blk[i] = msg[i];
    for (size_t i = 0; i < rs->ecc; ++i)
        blk[rs->k + i] = ecc[i];
    pntt(rs, blk);
    memcpy(synds, blk, rs->ecc * 4);
    free(blk);
}

/* simd.hpp:234-242: degree + 1 over the first max entries */
static uint32_t poly_len(const uint32_t *r, uint32_t max)
{
    while (max > 0 && r[max - 1] == 0)
        --max;
    return max;
}

/* NTT::_reverse_ntt_vec, ntt.hpp:344-354 */
static void reverse_ntt(const orc_rsi16 *rs, uint32_t *v, uint32_t shift)
{
    size_t E = rs->ecc;
    uint32_t mask = (uint32_t)E - 1;
    for (size_t i = 1; i < E; ++i) {
        uint32_t j = rs->rbo_ecc[i];
        uint32_t k = rs->rbo_ecc[E - j];
        if (i < k) {
            uint32_t t = v[i];
            v[i] = v[k];
            v[k] = t;
        }
        v[i] = s_mul(v[i], rs->roots_i_ecc[j]);
        v[i] = s_mul(v[i], rs->roots_ecc[(shift * j) & mask]);
    }
}

/* NTT::poly_inv_mod_xn_rev, ntt.hpp:363-469 (scalar instantiation). */
static uint32_t poly_inv_mod_xn_rev(const orc_rsi16 *rs, uint32_t *g, uint32_t g_len, uint32_t g0,
                                    uint32_t a_len, uint32_t *a, int *err)
{
    size_t E = rs->ecc;
    uint32_t a0 = g_inv(rs, g0);
    for (size_t i = 0; i < E; ++i)
        a[i] = a0;
    uint32_t l = 2;
    uint32_t a_max = a_len;
    while (l < a_len) {
        if (g_len + l - 2 > E) { /* py_assert at ntt.hpp:388-391 */
            *err = 1;
            return 0;
        }
        for (size_t i = 0; i < E; ++i) {
            uint32_t ai = a[i];
            uint32_t ag = s_mul(ai, g[i]);
            a[i] = s_mul(ai, s_sub(2, ag));
        }
        inttr(rs, a);
        for (size_t i = l; i < E; ++i)
            a[i] = 0;
        nttr(rs, a);
        l *= 2;
    }
    for (size_t i = 0; i < E; ++i) {
        uint32_t ag = s_mul(a[i], g[i]);
        g[i] = s_sub(2, ag);
    }
    for (size_t i = 0; i < E; ++i)
        g[i] = s_mul(a[i], g[i]);
    inttr(rs, g); /* inttr(a) of the reference is overwritten below */
    memcpy(a, g, E * 4);
    for (size_t i = l; i < E; ++i)
        a[i] = 0;
    for (uint32_t i = 0, j = a_max - 1; i < a_max / 2; ++i, --j) {
        uint32_t t = a[i];
        a[i] = a[j];
        a[j] = t;
    }
    for (size_t i = a_len; i < E; ++i)
        a[i] = 0;
    nttr(rs, a);
    return a_len;
}

/* NTT::poly_div_ntt, ntt.hpp:471-564 (scalar instantiation). */
static uint32_t poly_div_ntt(const orc_rsi16 *rs, const uint32_t *f, uint32_t f_len, const uint32_t *g,
                             uint32_t g_len, uint32_t f0, uint32_t g0, uint32_t *t, uint32_t *q, int *err)
{
    size_t E = rs->ecc;
    uint32_t q_len = f_len - g_len + 1;
    if (g_len == 0) { /* py_assert(g_len > 0) */
        *err = 1;
        return 0;
    }
    if (f_len < g_len) {
        memset(q, 0, E * 4);
        return 0;
    }
    memcpy(t, g, E * 4);
    reverse_ntt(rs, t, g_len);
    poly_inv_mod_xn_rev(rs, t, g_len, g0, q_len, q, err);
    if (*err)
        return 0;
    memcpy(t, f, E * 4);
    for (size_t j = 0; j < E; ++j)
        t[j] = s_sub(t[j], f0);
    for (size_t j = 0; j < E; ++j)
        t[j] = s_mul(t[j], q[j]);
    inttr(rs, t);
    uint32_t right_n = q_len < (uint32_t)E - g_len ? q_len : (uint32_t)E - g_len;
    for (uint32_t i = 0; i < right_n; ++i)
        q[i] = t[g_len + i];
    for (uint32_t i = 0; i < q_len - right_n; ++i)
        q[right_n + i] = t[i];
    for (size_t i = q_len; i < E; ++i)
        q[i] = 0;
    nttr(rs, q);
    return q_len;
}

/* RSi16v<1>::_sugiyama, rsi16v_impl.hpp:271-440. r1 holds the syndromes on entry
 * and the (normalised) evaluator on exit; temp has 4*ecc entries. */
static uint32_t sugiyama(const orc_rsi16 *rs, uint32_t *a1, uint32_t *r1, uint32_t *temp, int *err)
{
    size_t E = rs->ecc;
    uint32_t mask = (uint32_t)E - 1;
    memset(temp, 0, 4 * E * 4);
    memset(a1, 0, E * 4);
    uint32_t *q = temp, *t = temp + E, *a2 = temp + 2 * E, *r2 = temp + 3 * E;
    uint32_t q_len = 0, t_len = 0, a2_len = 1, a1_len;
    for (size_t i = 0; i < E; ++i)
        a2[i] = 1;
    uint32_t r1_len = poly_len(r1, (uint32_t)E);
    memcpy(r2, r1, E * 4);
    uint32_t r2_len = r1_len;
    uint32_t r2l = r2[0];
    uint32_t lanes = (uint32_t)-(int32_t)(r1_len != 0);

    {
        uint32_t r1h = r1[(r1_len - 1) & mask];
        uint32_t r1h2 = r1[(r1_len - 2) & mask];
        a1[1] = s_neg(g_inv(rs, r1h));
        a1[0] = s_mul(r1h2, s_mul(a1[1], a1[1]));
        a1_len = 2 & lanes;
        for (size_t i = 0; i < E; ++i)
            r1[i] = s_mul(a1[0], r1[i]);
        for (size_t i = 1; i < E; ++i)
            r1[i] = s_add(r1[i], s_mul(a1[1], r2[i - 1]));
        r1_len = poly_len(r1, (uint32_t)E - 1) & lanes;
        for (size_t i = r1_len; i < E; ++i)
            r1[i] = 0;
        nttr(rs, a1);
        nttr(rs, r2);
    }

    uint32_t cond = (uint32_t)-(int32_t)(r1_len >= (uint32_t)(E / 2));
    if (cond) {
        for (size_t it = 1; it < E / 2; ++it) {
            uint32_t r1h = r1[(r1_len - 1) & mask];
            uint32_t r1l = r1[0];
            nttr(rs, r1);
            q_len = poly_div_ntt(rs, r2, r2_len & cond, r1, r1_len, r2l, r1h, t, q, err);
            if (*err)
                return 0;
            for (size_t i = 0; i < E; ++i) /* poly_mul_ntt, ntt.hpp:356-361 */
                t[i] = s_mul(q[i], a1[i]);
            t_len = q_len + a1_len - 1;
            for (size_t i = 0; i < E; ++i)
                q[i] = s_mul(q[i], r1[i]);
            q_len = q_len + r1_len - 1;
            for (size_t i = 0; i < E; ++i) /* _vec_sub, rsi16v_impl.hpp:578-584 */
                t[i] = s_sub(a2[i], t[i]);
            t_len = a2_len > t_len ? a2_len : t_len;
            for (size_t i = 0; i < E; ++i)
                q[i] = s_sub(r2[i], q[i]);
            q_len = r2_len > q_len ? r2_len : q_len;
            memcpy(a2, a1, E * 4);
            a2_len = a1_len;
            memcpy(r2, r1, E * 4);
            r2_len = r1_len;
            r2l = r1l;
            cond = (uint32_t)-(int32_t)(r1_len >= (uint32_t)(E / 2));
            if (!cond) {
                inttr(rs, r1);
                break;
            }
            memcpy(a1, t, E * 4);
            memcpy(r1, q, E * 4);
            a1_len = t_len;
            r1_len = q_len;
            inttr(rs, r1);
            r1_len = poly_len(r1, r1_len & cond);
        }
    }
    inttr(rs, a1);
    a1_len = poly_len(a1, a1_len & lanes);
    uint32_t inv0 = g_inv(rs, a1[0]);
    for (size_t i = 0; i < E; ++i)
        a1[i] = s_mul(a1[i], inv0);
    for (size_t i = 0; i < E; ++i)
        r1[i] = s_mul(r1[i], inv0);
    return a1_len & lanes;
}

uint32_t orc_rsi16_sugiyama(const orc_rsi16 *rs, const uint32_t *synds, uint32_t *locator, uint32_t *evaluator)
{
    size_t E = rs->ecc;
    uint32_t *temp = malloc(4 * E * 4);
    int err = 0;
    memcpy(evaluator, synds, E * 4);
    uint32_t L = sugiyama(rs, locator, evaluator, temp, &err);
    free(temp);
    return err ? (uint32_t)-1 : L;
}

/* RSi16v::_eval, rsi16v_impl.hpp:628-636 */
static uint32_t poly_eval(const uint32_t *r, uint32_t len, uint32_t x)
{
    uint32_t s = 0;
    for (uint32_t i = len - 1; i < len; --i)
        s = s_add(s_mul(s, x), r[i]);
    return s;
}

/* RSi16v::_forney, rsi16v_impl.hpp:500-512 */
static uint32_t forney(const orc_rsi16 *rs, const uint32_t *deriv, uint32_t dlen, const uint32_t *eval,
                       uint32_t elen, uint32_t x)
{
    uint32_t xi = g_inv(rs, x);
    uint32_t num = poly_eval(eval, elen, xi);
    uint32_t den = poly_eval(deriv, dlen, xi);
    return s_mul(g_div(rs, num, den), x);
}

/* RSi16v::_repair (unknown positions), rsi16v_impl.hpp:122-168, + _find_roots_ntt :443-497 */
static int rsi16_repair_block(const orc_rsi16 *rs, uint32_t *blk, uint32_t *temp)
{
    size_t E = rs->ecc, n = rs->n, N = rs->N;
    uint32_t *synds = temp;
    memcpy(synds, blk, n * 4);
    pntt(rs, synds);
    int any = 0;
    for (size_t i = 0; i < E; ++i)
        any |= synds[i] != 0;
    if (!any)
        return ORC_NO_ERRORS;

    uint32_t *eval = synds;
    uint32_t *loc = temp + E;
    int err = 0;
    uint32_t L = sugiyama(rs, loc, eval, temp + 2 * E, &err);
    if (err)
        return ORC_RSI16_ASSERT;
    uint32_t eval_len = (uint32_t)E - L + 1;

    uint32_t *deriv = temp + 2 * E;
    memcpy(deriv, loc, E * 4);
    for (uint32_t i = 0; i < E - 1; ++i) /* _deriv, rsi16v_impl.hpp:608-616 */
        deriv[i] = s_mul(deriv[i + 1], i + 1);
    deriv[E - 1] = 0;
    uint32_t deriv_len = poly_len(deriv, (uint32_t)E);

    uint32_t *roots = temp + 3 * E;
    memset(roots, 0, N * 4);
    memcpy(roots, loc, E * 4);
    if (L > E) /* py_assert at rsi16v_impl.hpp:454 */
        return ORC_RSI16_ASSERT;
    /* NTT::pintt_ecc, ntt.hpp:206-220 */
    for (size_t i = 1; i < n / E; ++i) {
        memcpy(roots + i * E, roots, E * 4);
        for (size_t j = 0; j < E; ++j)
            roots[i * E + j] = s_mul(roots[i * E + j], rs->pintt_shift[i * E + j]);
        gs_butterfly(rs->roots_i_ecc, roots + i * E, E, 0);
    }
    for (size_t j = 0; j < E; ++j)
        roots[j] = s_mul(roots[j], rs->pintt_shift[j]);
    gs_butterfly(rs->roots_i_ecc, roots, E, 0);

    uint32_t lanes = (uint32_t)-(int32_t)(L != 0);
    uint32_t dmax = deriv_len & lanes, emax = eval_len & lanes;
    uint32_t root_count = 0;
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t x = rs->roots_ntt[rs_rbo(rs, i)];
        uint32_t hit = (roots[i] == 0) & lanes;
        root_count += 1 & hit;
        if (hit)
            blk[i] = s_add(blk[i], forney(rs, deriv, dmax, eval, emax, x));
    }
    uint32_t no_root = (uint32_t)-(int32_t)((root_count == 0) & lanes);
    uint32_t found = (uint32_t)-(int32_t)(root_count > 0);
    return (int)((ORC_REPAIR_OK & found) | (ORC_ERROR_LOCATION_FAIL & no_root));
}

int orc_rsi16_repair(const orc_rsi16 *rs, uint16_t *msg, uint16_t *ecc)
{
    size_t n = rs->n, k = rs->k, E = rs->ecc;
    uint32_t *blk = malloc(n * 4);
    uint32_t *temp = malloc((rs->N + 6 * E) * 4);
    for (size_t i = 0; i < k; ++i)
        blk[i] = msg[i];
    for (size_t i = 0; i < E; ++i)
        blk[k + i] = ecc[i];
    int st = rsi16_repair_block(rs, blk, temp);
    if (st != ORC_RSI16_ASSERT) {
        for (size_t i = 0; i < k; ++i)
            msg[i] = (uint16_t)blk[i];
        for (size_t i = 0; i < E; ++i)
            ecc[i] = (uint16_t)blk[k + i];
    }
    free(blk);
    free(temp);
    return st;
}

static void rsi16_rep_range(void *p, size_t b, size_t e)
{
    rsi16_batch *c = p;
    for (size_t i = b; i < e; ++i)
        c->status[i] = (uint8_t)orc_rsi16_repair(c->rs, c->msgs + i * c->rs->k, c->eccs + i * c->rs->ecc);
}

void orc_rsi16_repair_batch(const orc_rsi16 *rs, uint16_t *msgs, uint16_t *eccs, size_t count,
                            uint8_t *status, int threads)
{
    rsi16_batch c = {rs, msgs, eccs, status, 0};
    orc_parallel(count, threads, rsi16_rep_range, &c);
}

/* RSi16v::_repair (known positions), rsi16v_impl.hpp:171-231, with _error_locator
 * (:532-542) and _error_evaluator (:546-563). */
int orc_rsi16_repair_known(const orc_rsi16 *rs, uint16_t *msg, uint16_t *ecc, const size_t *pos, size_t npos)
{
    size_t n = rs->n, k = rs->k, E = rs->ecc, N = rs->N;
    if (npos == 0)
        return ORC_NO_ERRORS;
    if (npos > E)
        return ORC_RSI16_ASSERT;
    for (size_t i = 0; i < npos; ++i)
        if (pos[i] >= n)
            return ORC_RSI16_ASSERT;
    uint32_t *blk = malloc(n * 4);
    uint32_t *temp = malloc((N + 6 * E) * 4);
    for (size_t i = 0; i < k; ++i)
        blk[i] = msg[i];
    for (size_t i = 0; i < E; ++i)
        blk[k + i] = ecc[i];
    int st;
    uint32_t *synds = temp;
    memcpy(synds, blk, n * 4);
    pntt(rs, synds);
    int any = 0;
    for (size_t i = 0; i < E; ++i)
        any |= synds[i] != 0;
    if (!any) {
        st = ORC_NO_ERRORS;
    } else {
        uint32_t *ev = synds; /* 2*ecc */
        uint32_t *loc = temp + 2 * E;
        memset(loc, 0, E * 4);
        for (size_t j = 0; j < npos; ++j) {
            uint32_t x = rs->roots_ntt[rs_rbo(rs, (uint32_t)pos[j])];
            for (size_t i = npos - 1; i > 0; --i)
                loc[i] = s_sub(loc[i], s_mul(loc[i - 1], x));
            loc[0] = s_sub(loc[0], x);
        }
        uint32_t *tmp = temp + 4 * E;
        for (size_t i = E; i < 2 * E; ++i)
            ev[i] = 0;
        gs_butterfly(rs->roots_ecc2, ev, 2 * E, 0);
        tmp[0] = 1;
        for (size_t i = 0; i < npos; ++i)
            tmp[1 + i] = loc[i];
        for (size_t i = npos + 1; i < 2 * E; ++i)
            tmp[i] = 0;
        gs_butterfly(rs->roots_ecc2, tmp, 2 * E, 0);
        for (size_t i = 0; i < 2 * E; ++i)
            ev[i] = s_mul(ev[i], tmp[i]);
        ct_butterfly(rs->roots_i_ecc2, ev, 2 * E);
        for (size_t i = 0; i < 2 * E; ++i)
            ev[i] = s_mul(ev[i], rs->ecc2_inv);
        for (size_t i = E; i < 2 * E; ++i)
            ev[i] = 0;
        uint32_t *deriv = temp + E;
        memcpy(deriv, loc, E * 4);
        for (uint32_t i = 0; i < E; ++i) /* _deriv_shifted, rsi16v_impl.hpp:619-626 */
            deriv[i] = s_mul(deriv[i], i + 1);
        uint32_t dlen = poly_len(deriv, (uint32_t)E);
        for (size_t i = 0; i < npos; ++i) {
            uint32_t p = (uint32_t)pos[i];
            uint32_t x = rs->roots_ntt[rs_rbo(rs, p)];
            blk[p] = s_add(blk[p], forney(rs, deriv, dlen, ev, (uint32_t)E, x));
        }
        st = ORC_REPAIR_OK;
    }
    for (size_t i = 0; i < k; ++i)
        msg[i] = (uint16_t)blk[i];
    for (size_t i = 0; i < E; ++i)
        ecc[i] = (uint16_t)blk[k + i];
    free(blk);
    free(temp);
    return st;
}

/* =====================================================================
 * RSi16 repair with the SIMD words: W = 4, 8 or 16 columns in lockstep, as the vector
 * instantiations RSi16v<W>::_repair (rsi16v_impl.hpp:122-168), _sugiyama (:271-440) and
 * _find_roots_ntt (:443-497), NTT::poly_div_ntt / poly_inv_mod_xn_rev (ntt.hpp:363-564)
 * and the vector helpers of simd.hpp:60-254 run them for _repair_interleaved
 * (pyrsi16.hpp:389-453). Lane j's polynomial of E coefficients is at x + j * E; per-lane
 * lengths are arrays of W. Field arithmetic on canonical values is the same in the vector
 * and scalar words, so the per-lane work reuses the scalar routines; what the W lanes share
 * is every loop bound and branch the reference takes on a vector: any() / max() over the
 * lanes, masked copies, and lengths assigned to all lanes.
 * ===================================================================== */

#define VW_MAX 16

static uint32_t vl_max(const uint32_t *a, unsigned W)
{
    uint32_t m = 0;
    for (unsigned j = 0; j < W; ++j)
        m = a[j] > m ? a[j] : m;
    return m;
}

static int vl_any(const uint32_t *a, unsigned W)
{
    for (unsigned j = 0; j < W; ++j)
        if (a[j])
            return 1;
    return 0;
}

static inline uint32_t vl_cond(int c) { return c ? 0xffffffffu : 0u; }

/* simd.hpp:228-254 with a shared bound (clamped to E, where the reference would read past
 * the polynomial) */
static uint32_t vl_poly_len(const uint32_t *r, uint32_t bound, size_t E)
{
    return poly_len(r, bound < (uint32_t)E ? bound : (uint32_t)E);
}

/* ntt.hpp:363-469, vector instantiation. The Newton step's vec::assign_masked honours its
 * condition for W = 4 and 8 (rsi16v_sse2.cpp:47-51, rsi16v_avx2.cpp) and ignores it for
 * W = 16 (rsi16v_avx512.cpp: `condition[j] || true`), where every lane is updated. */
static void vl_poly_inv_mod_xn_rev(const orc_rsi16 *rs, unsigned W, uint32_t *g, const uint32_t *g_len,
                                   const uint32_t *g0, const uint32_t *a_len, uint32_t *a, int *err)
{
    size_t E = rs->ecc;
    uint32_t cond[VW_MAX];
    for (unsigned j = 0; j < W; ++j) {
        uint32_t a0 = g_inv(rs, g0[j]);
        for (size_t i = 0; i < E; ++i)
            a[j * E + i] = a0;
    }
    uint32_t l = 2;
    uint32_t a_max = vl_max(a_len, W);
    for (unsigned j = 0; j < W; ++j)
        cond[j] = vl_cond(l < a_len[j]);
    while (vl_any(cond, W)) {
        uint32_t gm[VW_MAX];
        for (unsigned j = 0; j < W; ++j)
            gm[j] = g_len[j] & cond[j];
        if (vl_max(gm, W) + l - 2 > E) { /* py_assert at ntt.hpp:388-391 */
            *err = 1;
            return;
        }
        for (unsigned j = 0; j < W; ++j) {
            uint32_t *aj = a + j * E, *gj = g + j * E;
            if (cond[j] || W == 16)
                for (size_t i = 0; i < E; ++i)
                    aj[i] = s_mul(aj[i], s_sub(2, s_mul(aj[i], gj[i])));
            inttr(rs, aj);
            for (size_t i = l; i < E; ++i)
                aj[i] = 0;
            nttr(rs, aj);
        }
        l *= 2;
        for (unsigned j = 0; j < W; ++j)
            cond[j] = vl_cond(l < a_len[j]);
    }
    for (unsigned j = 0; j < W; ++j) {
        uint32_t *aj = a + j * E, *gj = g + j * E;
        for (size_t i = 0; i < E; ++i)
            gj[i] = s_sub(2, s_mul(aj[i], gj[i]));
        for (size_t i = 0; i < E; ++i)
            gj[i] = s_mul(aj[i], gj[i]);
        inttr(rs, gj);
        memcpy(aj, gj, E * 4);
        for (size_t i = l; i < E; ++i)
            aj[i] = 0;
        for (uint32_t i = 0, m = a_max - 1; i < a_max / 2; ++i, --m) { /* std::reverse(&a[0], &a[a_max]) */
            uint32_t t = aj[i];
            aj[i] = aj[m];
            aj[m] = t;
        }
        for (uint32_t i = 0; i < a_len[j]; ++i) /* vec::copy_n(&a[0], a_max - a_len, a_len, &a[0], 0) */
            aj[i] = aj[a_max - a_len[j] + i];
        for (size_t i = a_len[j]; i < E; ++i)
            aj[i] = 0;
        nttr(rs, aj);
    }
}

/* ntt.hpp:471-564, vector instantiation; q_len[] receives the quotient lengths */
static void vl_poly_div_ntt(const orc_rsi16 *rs, unsigned W, const uint32_t *f, const uint32_t *f_len,
                            const uint32_t *g, const uint32_t *g_len, const uint32_t *f0, const uint32_t *g0,
                            uint32_t *t, uint32_t *q, uint32_t *q_len, int *err)
{
    size_t E = rs->ecc;
    for (unsigned j = 0; j < W; ++j)
        if (g_len[j] == 0 && f_len[j] > 0) { /* py_assert(!any((g_len == 0) && (f_len > 0))) */
            *err = 1;
            return;
        }
    for (unsigned j = 0; j < W; ++j)
        q_len[j] = (f_len[j] - g_len[j] + 1) & vl_cond(g_len[j] != 0 && f_len[j] >= g_len[j]);
    for (unsigned j = 0; j < W; ++j) {
        memcpy(t + j * E, g + j * E, E * 4);
        reverse_ntt(rs, t + j * E, g_len[j]);
    }
    vl_poly_inv_mod_xn_rev(rs, W, t, g_len, g0, q_len, q, err);
    if (*err)
        return;
    for (unsigned j = 0; j < W; ++j) {
        uint32_t *tj = t + j * E, *qj = q + j * E;
        const uint32_t *fj = f + j * E;
        for (size_t i = 0; i < E; ++i)
            tj[i] = s_mul(s_sub(fj[i], f0[j]), qj[i]);
        inttr(rs, tj);
        uint32_t room = (uint32_t)E - g_len[j];
        uint32_t right_n = q_len[j] < room ? q_len[j] : room;
        for (uint32_t i = 0; i < right_n; ++i)
            qj[i] = tj[g_len[j] + i];
        for (uint32_t i = 0; i < q_len[j] - right_n; ++i)
            qj[right_n + i] = tj[i];
        for (size_t i = q_len[j]; i < E; ++i)
            qj[i] = 0;
        nttr(rs, qj);
    }
}

/* rsi16v_impl.hpp:271-440, vector instantiation. r1 holds the W syndrome vectors on entry
 * and the evaluators on exit; temp has 4 * E * W entries. L[] receives the locator
 * lengths. */
static void vl_sugiyama(const orc_rsi16 *rs, unsigned W, uint32_t *a1, uint32_t *r1, uint32_t *temp, uint32_t *L,
                        int *err)
{
    size_t E = rs->ecc, EW = E * W;
    uint32_t mask = (uint32_t)E - 1;
    memset(temp, 0, 4 * EW * 4);
    memset(a1, 0, EW * 4);
    uint32_t *q = temp, *t = temp + EW, *a2 = temp + 2 * EW, *r2 = temp + 3 * EW;
    uint32_t q_len[VW_MAX], t_len[VW_MAX], a2_len[VW_MAX], a1_len[VW_MAX], r1_len[VW_MAX], r2_len[VW_MAX];
    uint32_t r2l[VW_MAX], lanes[VW_MAX], cond[VW_MAX], r1h[VW_MAX], r1l[VW_MAX], f_len[VW_MAX], bound[VW_MAX];
    for (unsigned j = 0; j < W; ++j) {
        uint32_t *a1j = a1 + j * E, *r1j = r1 + j * E, *a2j = a2 + j * E, *r2j = r2 + j * E;
        q_len[j] = 0;
        t_len[j] = 0;
        a2_len[j] = 1;
        for (size_t i = 0; i < E; ++i)
            a2j[i] = 1;
        r1_len[j] = poly_len(r1j, (uint32_t)E);
        memcpy(r2j, r1j, E * 4);
        r2_len[j] = r1_len[j];
        r2l[j] = r2j[0];
        lanes[j] = vl_cond(r1_len[j] != 0);
        /* the degree-1 first step (:308-335) */
        uint32_t h = r1j[(r1_len[j] - 1) & mask];
        uint32_t h2 = r1j[(r1_len[j] - 2) & mask];
        a1j[1] = s_neg(g_inv(rs, h));
        a1j[0] = s_mul(h2, s_mul(a1j[1], a1j[1]));
        a1_len[j] = 2 & lanes[j];
        for (size_t i = 0; i < E; ++i)
            r1j[i] = s_mul(a1j[0], r1j[i]);
        for (size_t i = 1; i < E; ++i)
            r1j[i] = s_add(r1j[i], s_mul(a1j[1], r2j[i - 1]));
        r1_len[j] = poly_len(r1j, (uint32_t)E - 1) & lanes[j];
        for (size_t i = r1_len[j]; i < E; ++i)
            r1j[i] = 0;
        nttr(rs, a1j);
        nttr(rs, r2j);
        cond[j] = vl_cond(r1_len[j] >= (uint32_t)(E / 2));
    }
    if (vl_any(cond, W)) {
        for (size_t it = 1; it < E / 2; ++it) {
            for (unsigned j = 0; j < W; ++j) {
                uint32_t *r1j = r1 + j * E;
                r1h[j] = r1j[(r1_len[j] - 1) & mask];
                r1l[j] = r1j[0];
                nttr(rs, r1j);
                f_len[j] = r2_len[j] & cond[j];
            }
            vl_poly_div_ntt(rs, W, r2, f_len, r1, r1_len, r2l, r1h, t, q, q_len, err);
            if (*err)
                return;
            for (unsigned j = 0; j < W; ++j) {
                uint32_t *qj = q + j * E, *tj = t + j * E, *a1j = a1 + j * E, *r1j = r1 + j * E;
                uint32_t *a2j = a2 + j * E, *r2j = r2 + j * E;
                for (size_t i = 0; i < E; ++i) /* poly_mul_ntt, ntt.hpp:356-361 */
                    tj[i] = s_mul(qj[i], a1j[i]);
                t_len[j] = q_len[j] + a1_len[j] - 1;
                for (size_t i = 0; i < E; ++i)
                    qj[i] = s_mul(qj[i], r1j[i]);
                q_len[j] = q_len[j] + r1_len[j] - 1;
                for (size_t i = 0; i < E; ++i) /* _vec_sub, rsi16v_impl.hpp:578-584 */
                    tj[i] = s_sub(a2j[i], tj[i]);
                t_len[j] = a2_len[j] > t_len[j] ? a2_len[j] : t_len[j];
                for (size_t i = 0; i < E; ++i)
                    qj[i] = s_sub(r2j[i], qj[i]);
                q_len[j] = r2_len[j] > q_len[j] ? r2_len[j] : q_len[j];
                memcpy(a2j, a1j, E * 4);
                a2_len[j] = a1_len[j];
                memcpy(r2j, r1j, E * 4);
                r2_len[j] = r1_len[j];
                r2l[j] = r1l[j];
                cond[j] = vl_cond(r1_len[j] >= (uint32_t)(E / 2));
            }
            if (!vl_any(cond, W)) {
                for (unsigned j = 0; j < W; ++j)
                    inttr(rs, r1 + j * E);
                break;
            }
            for (unsigned j = 0; j < W; ++j) {
                if (cond[j]) { /* vec::copy_n_masked */
                    memcpy(a1 + j * E, t + j * E, E * 4);
                    memcpy(r1 + j * E, q + j * E, E * 4);
                }
                a1_len[j] = t_len[j]; /* assigned in every lane */
                r1_len[j] = q_len[j];
                inttr(rs, r1 + j * E);
                bound[j] = r1_len[j] & cond[j];
            }
            uint32_t b = vl_max(bound, W);
            for (unsigned j = 0; j < W; ++j)
                r1_len[j] = vl_poly_len(r1 + j * E, b, E);
        }
    }
    for (unsigned j = 0; j < W; ++j) {
        inttr(rs, a1 + j * E);
        bound[j] = a1_len[j] & lanes[j];
    }
    uint32_t b = vl_max(bound, W);
    for (unsigned j = 0; j < W; ++j) {
        uint32_t *a1j = a1 + j * E, *r1j = r1 + j * E;
        a1_len[j] = vl_poly_len(a1j, b, E);
        uint32_t inv0 = g_inv(rs, a1j[0]);
        for (size_t i = 0; i < E; ++i)
            a1j[i] = s_mul(a1j[i], inv0);
        for (size_t i = 0; i < E; ++i)
            r1j[i] = s_mul(r1j[i], inv0);
        L[j] = a1_len[j] & lanes[j];
    }
}

/* rsi16v_impl.hpp:122-168 + 443-497 for W lanes: blk holds lane j's codeword at
 * blk + j * n. Returns the group's status (the maximum over the lanes) or ORC_RSI16_ASSERT. */
static int vl_repair_group(const orc_rsi16 *rs, unsigned W, uint32_t *blk)
{
    size_t E = rs->ecc, n = rs->n, N = rs->N, EW = E * W;
    uint32_t *synds = malloc(N * 4);
    uint32_t *ev = malloc(EW * 4), *loc = malloc(EW * 4), *der = malloc(EW * 4), *tmp = malloc(4 * EW * 4);
    uint32_t *roots = malloc(N * 4);
    uint32_t L[VW_MAX], elen[VW_MAX], dlen[VW_MAX], lanes[VW_MAX];
    int any = 0, err = 0, st = ORC_NO_ERRORS;
    for (unsigned j = 0; j < W; ++j) {
        memcpy(synds, blk + j * n, n * 4);
        pntt(rs, synds);
        memcpy(ev + j * E, synds, E * 4);
        for (size_t i = 0; i < E; ++i)
            any |= synds[i] != 0;
    }
    if (!any)
        goto done;
    vl_sugiyama(rs, W, loc, ev, tmp, L, &err);
    if (err) {
        st = ORC_RSI16_ASSERT;
        goto done;
    }
    for (unsigned j = 0; j < W; ++j) {
        uint32_t *dj = der + j * E;
        lanes[j] = vl_cond(L[j] != 0);
        elen[j] = ((uint32_t)E - L[j] + 1) & lanes[j];
        memcpy(dj, loc + j * E, E * 4);
        for (uint32_t i = 0; i < E - 1; ++i) /* _deriv, rsi16v_impl.hpp:608-616 */
            dj[i] = s_mul(dj[i + 1], i + 1);
        dj[E - 1] = 0;
        dlen[j] = poly_len(dj, (uint32_t)E);
    }
    if (vl_max(L, W) > E) { /* py_assert at rsi16v_impl.hpp:454 */
        st = ORC_RSI16_ASSERT;
        goto done;
    }
    uint32_t dm[VW_MAX], em[VW_MAX];
    for (unsigned j = 0; j < W; ++j) {
        dm[j] = dlen[j] & lanes[j];
        em[j] = elen[j] & lanes[j];
    }
    uint32_t dmax = vl_max(dm, W), emax = vl_max(em, W);
    for (unsigned j = 0; j < W; ++j) {
        /* NTT::pintt_ecc, ntt.hpp:206-220 */
        memset(roots, 0, N * 4);
        memcpy(roots, loc + j * E, E * 4);
        for (size_t i = 1; i < n / E; ++i) {
            memcpy(roots + i * E, roots, E * 4);
            for (size_t m = 0; m < E; ++m)
                roots[i * E + m] = s_mul(roots[i * E + m], rs->pintt_shift[i * E + m]);
            gs_butterfly(rs->roots_i_ecc, roots + i * E, E, 0);
        }
        for (size_t m = 0; m < E; ++m)
            roots[m] = s_mul(roots[m], rs->pintt_shift[m]);
        gs_butterfly(rs->roots_i_ecc, roots, E, 0);
        uint32_t count = 0;
        for (uint32_t i = 0; i < n; ++i) {
            if (roots[i] == 0 && lanes[j]) {
                count++;
                uint32_t x = rs->roots_ntt[rs_rbo(rs, i)];
                blk[j * n + i] = s_add(blk[j * n + i], forney(rs, der + j * E, dmax, ev + j * E, emax, x));
            }
        }
        int lst = count > 0 ? ORC_REPAIR_OK : (lanes[j] ? ORC_ERROR_LOCATION_FAIL : ORC_NO_ERRORS);
        st = lst > st ? lst : st;
    }
done:
    free(synds);
    free(ev);
    free(loc);
    free(der);
    free(tmp);
    free(roots);
    return st;
}

/* PyRSi16::_repair_interleaved<W> (pyrsi16.hpp:389-453) without error positions: groups of W
 * columns (column c is symbols i * interleave + c), a last group zero-filled past the
 * columns; each group is written back after its repair and the call returns the largest
 * status; a group whose repair raises ends the call (ORC_RSI16_ASSERT) with the groups
 * before it written back. */
int orc_rsi16_repair_interleaved_simd(const orc_rsi16 *rs, unsigned W, uint16_t *msg, uint16_t *ecc,
                                      size_t interleave)
{
    size_t n = rs->n, k = rs->k, E = rs->ecc;
    if (W != 4 && W != 8 && W != 16)
        return ORC_RSI16_ASSERT;
    uint32_t *blk = malloc(n * W * 4);
    int res = ORC_NO_ERRORS;
    for (size_t c0 = 0; c0 < interleave; c0 += W) {
        size_t cols = interleave - c0 < W ? interleave - c0 : W;
        memset(blk, 0, n * W * 4);
        for (size_t j = 0; j < cols; ++j) {
            for (size_t i = 0; i < k; ++i)
                blk[j * n + i] = msg[i * interleave + c0 + j];
            for (size_t i = 0; i < E; ++i)
                blk[j * n + k + i] = ecc[i * interleave + c0 + j];
        }
        int st = vl_repair_group(rs, W, blk);
        if (st == ORC_RSI16_ASSERT) {
            res = st;
            break;
        }
        for (size_t j = 0; j < cols; ++j) {
            for (size_t i = 0; i < k; ++i)
                msg[i * interleave + c0 + j] = (uint16_t)blk[j * n + i];
            for (size_t i = 0; i < E; ++i)
                ecc[i * interleave + c0 + j] = (uint16_t)blk[j * n + k + i];
        }
        res = st > res ? st : res;
    }
    free(blk);
    return res;
}
