/*
 * CPU oracle for the Circle-PCS commit path -- TEST INFRASTRUCTURE ONLY.
 *
 * Plain-C restatement of the reference's CPU oracles.  Each function names the
 * reference file:line it follows (paths relative to the reference repository
 * root, Sources/AppleZKProver/ unless stated).  This code is the checker: only
 * the tests, the smoke check and the bench's CPU-baseline leg load it.
 */
#include "azk_oracle.h"

#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------- */
/* Keccak-f[1600]                                  SHA3Oracle.swift:100-135,242-280 */
/* ------------------------------------------------------------------------- */

static const uint64_t KRC[24] = {
    0x0000000000000001ull, 0x0000000000008082ull, 0x800000000000808Aull, 0x8000000080008000ull,
    0x000000000000808Bull, 0x0000000080000001ull, 0x8000000080008081ull, 0x8000000000008009ull,
    0x000000000000008Aull, 0x0000000000000088ull, 0x0000000080008009ull, 0x000000008000000Aull,
    0x000000008000808Bull, 0x800000000000008Bull, 0x8000000000008089ull, 0x8000000000008003ull,
    0x8000000000008002ull, 0x8000000000000080ull, 0x000000000000800Aull, 0x800000008000000Aull,
    0x8000000080008081ull, 0x8000000000008080ull, 0x0000000080000001ull, 0x8000000080008008ull,
};

/* rotation offset r[x][y] of lane (x, y), index x + 5y (FIPS 202 table 2). */
static const int KROT[25] = {
    0, 1, 62, 28, 27,
    36, 44, 6, 55, 20,
    3, 10, 43, 25, 39,
    41, 45, 15, 21, 8,
    18, 2, 61, 56, 14,
};

static inline uint64_t rol64(uint64_t v, int n) { return n ? (v << n) | (v >> (64 - n)) : v; }

void ora_keccak_f1600(uint64_t a[25])
{
    for (int round = 0; round < 24; ++round) {
        uint64_t c[5], b[25];
        for (int x = 0; x < 5; ++x) c[x] = a[x] ^ a[x + 5] ^ a[x + 10] ^ a[x + 15] ^ a[x + 20];
        for (int x = 0; x < 5; ++x) {
            uint64_t d = c[(x + 4) % 5] ^ rol64(c[(x + 1) % 5], 1);
            for (int y = 0; y < 25; y += 5) a[x + y] ^= d;
        }
        /* rho + pi: B[y, 2x+3y] = rot(A[x, y], r[x, y]) */
        for (int x = 0; x < 5; ++x)
            for (int y = 0; y < 5; ++y)
                b[y + 5 * ((2 * x + 3 * y) % 5)] = rol64(a[x + 5 * y], KROT[x + 5 * y]);
        /* chi */
        for (int y = 0; y < 25; y += 5)
            for (int x = 0; x < 5; ++x) a[x + y] = b[x + y] ^ (~b[(x + 1) % 5 + y] & b[(x + 2) % 5 + y]);
        a[0] ^= KRC[round];
    }
}

static void absorb_block(uint64_t st[25], const uint8_t *p, int rate)
{
    for (int lane = 0; lane < rate / 8; ++lane) {
        uint64_t v = 0;
        for (int k = 0; k < 8; ++k) v |= (uint64_t)p[lane * 8 + k] << (8 * k);
        st[lane] ^= v;
    }
}

/* Generic pad10*1 sponge.                                  SHA3Oracle.swift:181-219 */
int ora_sponge(const uint8_t *data, size_t len, int rate, uint8_t suffix, uint8_t *out, size_t outlen)
{
    if (rate <= 0 || rate > 200 || rate % 8) return -1;
    uint64_t st[25] = {0};
    size_t off = 0;
    while (off + (size_t)rate <= len) {
        absorb_block(st, data + off, rate);
        ora_keccak_f1600(st);
        off += (size_t)rate;
    }
    uint8_t tail[200] = {0};
    size_t tc = len - off;
    if (tc) memcpy(tail, data + off, tc);
    tail[tc] ^= suffix;
    tail[rate - 1] ^= 0x80;
    absorb_block(st, tail, rate);
    ora_keccak_f1600(st);
    size_t produced = 0;
    while (produced < outlen) {
        size_t take = outlen - produced < (size_t)rate ? outlen - produced : (size_t)rate;
        for (size_t i = 0; i < take; ++i) out[produced + i] = (uint8_t)(st[i >> 3] >> (8 * (i & 7)));
        produced += take;
        if (produced < outlen) ora_keccak_f1600(st);
    }
    return 0;
}

/* SHA3Oracle.swift:137-139 */
void ora_sha3_256(const uint8_t *data, size_t len, uint8_t out[32]) { ora_sponge(data, len, 136, 0x06, out, 32); }
/* SHA3Oracle.swift:289-292 */
void ora_keccak_256(const uint8_t *data, size_t len, uint8_t out[32]) { ora_sponge(data, len, 136, 0x01, out, 32); }

/* Strided permutation batch.                KeccakF1600PermutationBatcher.swift:245-272 */
int ora_keccakf_batch(const uint8_t *in, size_t in_stride, uint8_t *out, size_t out_stride, size_t count)
{
    if (in_stride < 200 || out_stride < 200 || in_stride % 8 || out_stride % 8) return -1;
    for (size_t i = 0; i < count; ++i) {
        uint64_t st[25];
        memcpy(st, in + i * in_stride, 200);
        ora_keccak_f1600(st);
        memcpy(out + i * out_stride, st, 200);
        memset(out + i * out_stride + 200, 0, out_stride - 200);
    }
    return 0;
}

/* One-block (<= 136 B) SHA3-256 / Keccak-256 batch with zeroed output padding.
 * SHA3BatchHasher.swift:170-209,246-263; Keccak256BatchHasher.swift:162-…,238-… */
int ora_hash_oneblock_batch(const uint8_t *in, size_t stride, size_t len, uint8_t *out,
                            size_t out_stride, size_t count, uint8_t suffix)
{
    if (len > 136 || stride < len || out_stride < 32) return -1;
    for (size_t i = 0; i < count; ++i) {
        ora_sponge(in + i * stride, len, 136, suffix, out + i * out_stride, 32);
        memset(out + i * out_stride + 32, 0, out_stride - 32);
    }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Transcript sponge                                       SHA3Oracle.swift:8-98 */
/* ------------------------------------------------------------------------- */

void ora_transcript_init(ora_transcript *t) { memset(t, 0, sizeof *t); }

/* Each call absorbs one SHA3-padded message into the running state.  :19-39 */
void ora_transcript_absorb(ora_transcript *t, const uint8_t *data, size_t len, uint8_t suffix)
{
    size_t off = 0;
    while (off + 136 <= len) {
        absorb_block(t->s, data + off, 136);
        ora_keccak_f1600(t->s);
        off += 136;
    }
    uint8_t tail[136] = {0};
    size_t tc = len - off;
    if (tc) memcpy(tail, data + off, tc);
    tail[tc] ^= suffix;
    tail[135] ^= 0x80;
    absorb_block(t->s, tail, 136);
    ora_keccak_f1600(t->s);
}

/* Rejection-sampled u32 squeeze from a copy of the state, 34 words per block.  :41-70 */
int ora_transcript_squeeze_u32(const ora_transcript *t, uint32_t count, uint32_t modulus, uint32_t *out)
{
    if (modulus == 0) return -1;
    uint64_t space = 1ull << 32;
    uint64_t limit = space - space % modulus;
    uint64_t st[25];
    memcpy(st, t->s, sizeof st);
    uint32_t got = 0;
    uint64_t cand = 0;
    while (got < count) {
        if (cand > 0 && cand % 34 == 0) ora_keccak_f1600(st);
        uint32_t wi = (uint32_t)(cand % 34);
        uint64_t w = (uint32_t)(st[wi >> 1] >> ((wi & 1) * 32));
        ++cand;
        if (w < limit) out[got++] = (uint32_t)(w % modulus);
    }
    return 0;
}

/* :72-97 */
void ora_transcript_squeeze_bytes(const ora_transcript *t, size_t count, uint8_t *out)
{
    uint64_t st[25];
    memcpy(st, t->s, sizeof st);
    for (size_t i = 0; i < count; ++i) {
        if (i > 0 && i % 136 == 0) ora_keccak_f1600(st);
        size_t r = i % 136;
        out[i] = (uint8_t)(st[r >> 3] >> (8 * (r & 7)));
    }
}

/* ------------------------------------------------------------------------- */
/* Fields: M31 (SumcheckOracle.swift:3-147), CM31 (CM31Field.swift:41-96),    */
/* QM31 = CM31[u]/(u^2 - (2+i)) (QM31Field.swift:38-89,237-251 of its file)   */
/* ------------------------------------------------------------------------- */

static inline uint32_t m31_reduce(uint64_t v)
{
    uint64_t r = (v & ORA_P) + (v >> 31);
    r = (r & ORA_P) + (r >> 31);
    return (uint32_t)(r >= ORA_P ? r - ORA_P : r);
}
uint32_t ora_m31_add(uint32_t a, uint32_t b) { uint32_t s = a + b; return s >= ORA_P ? s - ORA_P : s; }
uint32_t ora_m31_sub(uint32_t a, uint32_t b) { return a >= b ? a - b : ORA_P - (b - a); }
uint32_t ora_m31_neg(uint32_t a) { return a ? ORA_P - a : 0; }
uint32_t ora_m31_mul(uint32_t a, uint32_t b) { return m31_reduce((uint64_t)a * b); }
static uint32_t m31_pow(uint32_t b, uint32_t e)
{
    uint32_t r = 1;
    while (e) {
        if (e & 1) r = ora_m31_mul(r, b);
        b = ora_m31_mul(b, b);
        e >>= 1;
    }
    return r;
}
uint32_t ora_m31_inv(uint32_t a) { return m31_pow(a, ORA_P - 2); }

typedef struct { uint32_t r, i; } cm31;
static cm31 cm_add(cm31 a, cm31 b) { return (cm31){ora_m31_add(a.r, b.r), ora_m31_add(a.i, b.i)}; }
static cm31 cm_sub(cm31 a, cm31 b) { return (cm31){ora_m31_sub(a.r, b.r), ora_m31_sub(a.i, b.i)}; }
static cm31 cm_mul(cm31 a, cm31 b)
{
    return (cm31){ora_m31_sub(ora_m31_mul(a.r, b.r), ora_m31_mul(a.i, b.i)),
                  ora_m31_add(ora_m31_mul(a.r, b.i), ora_m31_mul(a.i, b.r))};
}
static int cm_inv(cm31 a, cm31 *out)
{
    if (!a.r && !a.i) return -1;
    uint32_t d = ora_m31_add(ora_m31_mul(a.r, a.r), ora_m31_mul(a.i, a.i));
    uint32_t di = ora_m31_inv(d);
    *out = (cm31){ora_m31_mul(a.r, di), ora_m31_mul(ora_m31_neg(a.i), di)};
    return 0;
}
static const cm31 QM_NONRES = {2, 1};

typedef struct { uint32_t v[4]; } qm31;
static qm31 qm_load(const uint32_t *p) { qm31 q; memcpy(q.v, p, 16); return q; }
static qm31 qm_add(qm31 a, qm31 b) { qm31 o; for (int k = 0; k < 4; ++k) o.v[k] = ora_m31_add(a.v[k], b.v[k]); return o; }
static qm31 qm_sub(qm31 a, qm31 b) { qm31 o; for (int k = 0; k < 4; ++k) o.v[k] = ora_m31_sub(a.v[k], b.v[k]); return o; }
static qm31 qm_scale(qm31 a, uint32_t s) { qm31 o; for (int k = 0; k < 4; ++k) o.v[k] = ora_m31_mul(a.v[k], s); return o; }
static qm31 qm_mul(qm31 x, qm31 y)
{
    cm31 a = {x.v[0], x.v[1]}, b = {x.v[2], x.v[3]}, c = {y.v[0], y.v[1]}, d = {y.v[2], y.v[3]};
    cm31 lo = cm_add(cm_mul(a, c), cm_mul(QM_NONRES, cm_mul(b, d)));
    cm31 hi = cm_add(cm_mul(a, d), cm_mul(b, c));
    return (qm31){{lo.r, lo.i, hi.r, hi.i}};
}
static int qm_is_zero(qm31 a) { return !(a.v[0] | a.v[1] | a.v[2] | a.v[3]); }
static int qm_canonical(const uint32_t *p, size_t n)
{
    for (size_t k = 0; k < n; ++k)
        if (p[k] >= ORA_P) return 0;
    return 1;
}

void ora_qm31_add(const uint32_t a[4], const uint32_t b[4], uint32_t out[4]) { qm31 r = qm_add(qm_load(a), qm_load(b)); memcpy(out, r.v, 16); }
void ora_qm31_sub(const uint32_t a[4], const uint32_t b[4], uint32_t out[4]) { qm31 r = qm_sub(qm_load(a), qm_load(b)); memcpy(out, r.v, 16); }
void ora_qm31_mul(const uint32_t a[4], const uint32_t b[4], uint32_t out[4]) { qm31 r = qm_mul(qm_load(a), qm_load(b)); memcpy(out, r.v, 16); }
/* QM31Field.swift:227-235 of its file: a^2 + (2+i) b^2, 2ab */
void ora_qm31_square(const uint32_t a[4], uint32_t out[4])
{
    cm31 x = {a[0], a[1]}, y = {a[2], a[3]};
    cm31 lo = cm_add(cm_mul(x, x), cm_mul(QM_NONRES, cm_mul(y, y)));
    cm31 ab = cm_mul(x, y);
    cm31 hi = cm_add(ab, ab);
    out[0] = lo.r; out[1] = lo.i; out[2] = hi.r; out[3] = hi.i;
}
int ora_qm31_inv(const uint32_t a4[4], uint32_t out[4])
{
    qm31 q = qm_load(a4);
    if (!qm_canonical(a4, 4) || qm_is_zero(q)) return -1;
    cm31 a = {a4[0], a4[1]}, b = {a4[2], a4[3]}, di;
    cm31 den = cm_sub(cm_mul(a, a), cm_mul(QM_NONRES, cm_mul(b, b)));
    if (cm_inv(den, &di)) return -1;
    cm31 lo = cm_mul(a, di);
    cm31 nb = {ora_m31_neg(b.r), ora_m31_neg(b.i)};
    cm31 hi = cm_mul(nb, di);
    out[0] = lo.r; out[1] = lo.i; out[2] = hi.r; out[3] = hi.i;
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Circle group over M31                                 CircleDomain.swift:3-102 */
/* ------------------------------------------------------------------------- */

typedef struct { uint32_t x, y; } cpt;
static const cpt CGEN = {2u, 1268011823u};

static cpt cpt_add(cpt a, cpt b)
{
    return (cpt){ora_m31_sub(ora_m31_mul(a.x, b.x), ora_m31_mul(a.y, b.y)),
                 ora_m31_add(ora_m31_mul(a.x, b.y), ora_m31_mul(a.y, b.x))};
}
/* generator * index by double-and-add.                           :42-56,99-101 */
static cpt cpt_from_index(uint32_t idx)
{
    cpt r = {1, 0}, p = CGEN;
    idx &= 0x7fffffffu;
    while (idx) {
        if (idx & 1) r = cpt_add(r, p);
        idx >>= 1;
        if (idx) p = cpt_add(p, p);
    }
    return r;
}
void ora_circle_point_from_index(uint32_t index, uint32_t out_xy[2])
{
    cpt p = cpt_from_index(index);
    out_xy[0] = p.x; out_xy[1] = p.y;
}

static size_t bitrev(size_t v, uint32_t bits)
{
    size_t r = 0;
    for (uint32_t k = 0; k < bits; ++k) { r = (r << 1) | (v & 1); v >>= 1; }
    return r;
}

/* Canonical domain: half coset G_{4n/2..} = halfOdds(logN-1) then its conjugate.
 * CircleDomain.swift:122-136,225-239,273-285 */
int ora_circle_point_index(uint32_t log_n, uint32_t natural, uint32_t *out_index)
{
    if (log_n < 1 || log_n > 30 || natural >= (1u << log_n)) return -1;
    uint32_t init = 1u << (31 - (log_n + 1));
    uint32_t step = 1u << (31 - (log_n - 1));
    uint32_t half = 1u << (log_n - 1);
    uint32_t idx;
    if (natural < half) idx = (uint32_t)((init + (uint64_t)step * natural) & 0x7fffffffu);
    else idx = (uint32_t)((0x80000000u - ((init + (uint64_t)step * (natural - half)) & 0x7fffffffu)) & 0x7fffffffu);
    *out_index = idx;
    return 0;
}

/* Points of the domain in natural order: first half by repeated addition of the
 * step, second half as conjugates (index negation = (x, -y)). */
static cpt *domain_natural_points(uint32_t log_n)
{
    size_t n = (size_t)1 << log_n, half = n / 2;
    cpt *pts = (cpt *)malloc(n * sizeof(cpt));
    if (!pts) return NULL;
    uint32_t i0;
    ora_circle_point_index(log_n, 0, &i0);
    cpt p = cpt_from_index(i0);
    cpt step = cpt_from_index(1u << (31 - (log_n - 1)));
    for (size_t k = 0; k < half; ++k) {
        pts[k] = p;
        pts[half + k] = (cpt){p.x, ora_m31_neg(p.y)};
        p = cpt_add(p, step);
    }
    return pts;
}

/* storage order = bit-reversed natural order.  CircleDomain.swift:303-331 */
int ora_circle_domain_points(uint32_t log_n, uint32_t *out)
{
    if (log_n < 1 || log_n > 30) return -1;
    cpt *pts = domain_natural_points(log_n);
    if (!pts) return -2;
    size_t n = (size_t)1 << log_n;
    for (size_t s = 0; s < n; ++s) {
        cpt p = pts[bitrev(s, log_n)];
        out[2 * s] = p.x; out[2 * s + 1] = p.y;
    }
    free(pts);
    return 0;
}

static uint32_t double_x(uint32_t x) { uint32_t sq = ora_m31_mul(x, x); return ora_m31_sub(ora_m31_add(sq, sq), 1); }

/* Montgomery batch inverse (all inputs non-zero). */
static int batch_inv(uint32_t *v, size_t n)
{
    if (!n) return 0;
    uint32_t *pre = (uint32_t *)malloc(n * sizeof(uint32_t));
    if (!pre) return -2;
    uint32_t acc = 1;
    for (size_t k = 0; k < n; ++k) {
        if (v[k] == 0) { free(pre); return -1; }
        pre[k] = acc;
        acc = ora_m31_mul(acc, v[k]);
    }
    uint32_t inv = ora_m31_inv(acc);
    for (size_t k = n; k-- > 0;) {
        uint32_t t = ora_m31_mul(inv, pre[k]);
        inv = ora_m31_mul(inv, v[k]);
        v[k] = t;
    }
    free(pre);
    return 0;
}

/* Stage-major FFT twiddles: stage 0 -> y, stage s>=1 -> x doubled s-1 times, of
 * the point at storage index h << (s+1).  CircleCodeword.swift:129-171 */
int ora_circle_fft_twiddles(uint32_t log_n, uint32_t *out)
{
    if (log_n < 1 || log_n > 30) return -1;
    cpt *pts = domain_natural_points(log_n);
    if (!pts) return -2;
    size_t n = (size_t)1 << log_n, o = 0;
    for (uint32_t s = 0; s < log_n; ++s) {
        size_t cnt = n >> (s + 1);
        for (size_t h = 0; h < cnt; ++h) {
            cpt p = pts[bitrev(h << (s + 1), log_n)];
            if (s == 0) out[o++] = p.y;
            else {
                uint32_t x = p.x;
                for (uint32_t k = 1; k < s; ++k) x = double_x(x);
                out[o++] = x;
            }
        }
    }
    free(pts);
    return 0;
}

/* Inverse-domain layers as M31 words: round 0 = 1/y of the pair's left point,
 * round r>=1 = 1/x of the line pair's left coordinate, x <- 2x^2-1 between rounds.
 * CircleDomain.swift:399-418,468-511,546-562 */
int ora_circle_fold_twiddles(uint32_t log_n, uint32_t rounds, uint32_t *out)
{
    if (log_n < 1 || log_n > 30 || rounds < 1 || rounds > log_n) return -1;
    cpt *pts = domain_natural_points(log_n);
    if (!pts) return -2;
    size_t half = (size_t)1 << (log_n - 1);
    uint32_t *xs = (uint32_t *)malloc(half * sizeof(uint32_t));
    if (!xs) { free(pts); return -2; }
    for (size_t p = 0; p < half; ++p) {
        cpt pt = pts[bitrev(p << 1, log_n)];
        out[p] = pt.y;
        xs[p] = pt.x;
    }
    free(pts);
    int rc = batch_inv(out, half);
    size_t o = half, cur = half;
    for (uint32_t r = 1; r < rounds && rc == 0; ++r) {
        size_t pairs = cur / 2;
        for (size_t q = 0; q < pairs; ++q) {
            uint32_t l = xs[2 * q], rr = xs[2 * q + 1];
            if (ora_m31_add(l, rr) != 0 || l == 0) { rc = -1; break; }
            out[o + q] = l;
            xs[q] = double_x(l);
        }
        if (rc) break;
        rc = batch_inv(out + o, pairs);
        o += pairs;
        cur = pairs;
    }
    free(xs);
    return rc;
}

static void put_u32(uint8_t *p, uint32_t v) { for (int k = 0; k < 4; ++k) p[k] = (uint8_t)(v >> (8 * k)); }
static void put_u64(uint8_t *p, uint64_t v) { for (int k = 0; k < 8; ++k) p[k] = (uint8_t)(v >> (8 * k)); }
static uint32_t get_u32(const uint8_t *p) { return (uint32_t)p[0] | (uint32_t)p[1] << 8 | (uint32_t)p[2] << 16 | (uint32_t)p[3] << 24; }
static uint64_t get_u64(const uint8_t *p) { return (uint64_t)get_u32(p) | (uint64_t)get_u32(p + 4) << 32; }

/* 44-byte descriptor.                              CircleProofFormat.swift:71-93 */
int ora_circle_domain_descriptor(uint32_t log_n, uint8_t out[44])
{
    if (log_n < 1 || log_n > 30) return -1;
    static const uint8_t magic[8] = {0x41, 0x5a, 0x4b, 0x43, 0x44, 0x56, 0x31, 0x00};
    memcpy(out, magic, 8);
    put_u32(out + 8, 1);
    put_u32(out + 12, ORA_P);
    put_u32(out + 16, 31);
    put_u32(out + 20, CGEN.x);
    put_u32(out + 24, CGEN.y);
    put_u32(out + 28, log_n);
    put_u32(out + 32, 1u << (31 - (log_n + 1)));
    put_u32(out + 36, log_n - 1);
    put_u32(out + 40, 1);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Codeword                                           CircleCodeword.swift:27-279 */
/* ------------------------------------------------------------------------- */

/* Radix-2 stages logN-1 .. 1 then 0 on interleaved line-basis coefficients.  :72-98 */
int ora_circle_fft(const uint32_t *coeffs, uint32_t log_n, uint32_t *out)
{
    if (log_n < 1 || log_n > 30) return -1;
    size_t n = (size_t)1 << log_n;
    uint32_t *tw = (uint32_t *)malloc((n - 1 + 1) * sizeof(uint32_t));
    if (!tw) return -2;
    int rc = ora_circle_fft_twiddles(log_n, tw);
    if (rc) { free(tw); return rc; }
    memcpy(out, coeffs, n * 16);
    /* stage order: logN-1 ... 1, then 0 */
    for (int pass = 0; pass < (int)log_n; ++pass) {
        uint32_t s = pass < (int)log_n - 1 ? (uint32_t)(log_n - 1 - pass) : 0;
        size_t cnt = n >> (s + 1);
        size_t off = s ? n - (n >> s) : 0;
        size_t span = (size_t)1 << s;
        for (size_t h = 0; h < cnt; ++h) {
            uint32_t t = tw[off + h];
            for (size_t lane = 0; lane < span; ++lane) {
                size_t li = (h << (s + 1)) + lane, ri = li + span;
                qm31 l = qm_load(out + 4 * li), r = qm_scale(qm_load(out + 4 * ri), t);
                qm31 a = qm_add(l, r), b = qm_sub(l, r);
                memcpy(out + 4 * li, a.v, 16);
                memcpy(out + 4 * ri, b.v, 16);
            }
        }
    }
    free(tw);
    return 0;
}

#define INV2 1073741824u

/* Chebyshev T_k * x = (T_{k-1} + T_{k+1}) / 2, T_0 * x = T_1.   :220-236 */
static void cheb_times_x(const uint32_t *in, size_t n, uint32_t *out /* n+1 */)
{
    memset(out, 0, (n + 1) * sizeof(uint32_t));
    for (size_t k = 0; k < n; ++k) {
        uint32_t c = in[k];
        if (!c) continue;
        if (k == 0) out[1] = ora_m31_add(out[1], c);
        else {
            uint32_t h = ora_m31_mul(c, INV2);
            out[k - 1] = ora_m31_add(out[k - 1], h);
            out[k + 1] = ora_m31_add(out[k + 1], h);
        }
    }
}

/* Recursive Chebyshev -> line-basis split.                             :238-263 */
static void cheb_to_line(qm31 *c, size_t n)
{
    if (n <= 1) return;
    size_t half = n / 2;
    qm31 *hi = (qm31 *)malloc(half * sizeof(qm31));
    hi[0] = c[half];
    for (size_t k = 1; k < half; ++k) hi[k] = qm_scale(c[half + k], 2);
    for (size_t lo = 1; lo < half; ++lo) c[lo] = qm_sub(c[lo], qm_scale(hi[half - lo], INV2));
    memcpy(c + half, hi, half * sizeof(qm31));
    free(hi);
    cheb_to_line(c, half);
    cheb_to_line(c + half, half);
}

/* Monomial -> line-FFT-basis (O(n^2), small n).                        :187-218 */
static int line_basis(const uint32_t *mono, uint32_t nm, size_t cap, qm31 *out)
{
    if (nm > cap) return -1;
    memset(out, 0, cap * sizeof(qm31));
    if (!nm) return 0;
    uint32_t *pw = (uint32_t *)calloc(nm + 1, sizeof(uint32_t));
    uint32_t *nx = (uint32_t *)calloc(nm + 2, sizeof(uint32_t));
    size_t plen = 1;
    pw[0] = 1;
    for (uint32_t d = 0; d < nm; ++d) {
        qm31 c = qm_load(mono + 4 * d);
        if (!qm_is_zero(c))
            for (size_t k = 0; k < plen; ++k)
                if (pw[k]) out[k] = qm_add(out[k], qm_scale(c, pw[k]));
        if (d + 1 < nm) {
            cheb_times_x(pw, plen, nx);
            ++plen;
            memcpy(pw, nx, plen * sizeof(uint32_t));
        }
    }
    free(pw);
    free(nx);
    cheb_to_line(out, cap);
    return 0;
}

/* Interleave x/y line coefficients: coeff[2i] = x_i, coeff[2i+1] = y_i.  :100-127 */
int ora_circle_basis_from_monomial(const uint32_t *x, uint32_t nx, const uint32_t *y, uint32_t ny,
                                   uint32_t log_n, uint32_t *out)
{
    if (log_n < 1 || log_n > 24) return -1;
    size_t half = (size_t)1 << (log_n - 1);
    if (nx > half || ny > half || !qm_canonical(x, 4 * (size_t)nx) || !qm_canonical(y, 4 * (size_t)ny)) return -1;
    qm31 *xl = (qm31 *)malloc(half * sizeof(qm31)), *yl = (qm31 *)malloc(half * sizeof(qm31));
    line_basis(x, nx, half, xl);
    line_basis(y, ny, half, yl);
    for (size_t k = 0; k < half; ++k) {
        memcpy(out + 8 * k, xl[k].v, 16);
        memcpy(out + 8 * k + 4, yl[k].v, 16);
    }
    free(xl);
    free(yl);
    return 0;
}

/* Direct Horner evaluation of P(x) + y Q(x) at each stored point.     :27-70,173-185 */
int ora_circle_evaluate_direct(const uint32_t *x, uint32_t nx, const uint32_t *y, uint32_t ny,
                               uint32_t log_n, uint32_t *out)
{
    if (log_n < 1 || log_n > 24) return -1;
    size_t n = (size_t)1 << log_n;
    uint32_t *pts = (uint32_t *)malloc(n * 2 * sizeof(uint32_t));
    ora_circle_domain_points(log_n, pts);
    for (size_t s = 0; s < n; ++s) {
        uint32_t px = pts[2 * s], py = pts[2 * s + 1];
        qm31 ax = {{0, 0, 0, 0}}, ay = {{0, 0, 0, 0}};
        for (uint32_t k = nx; k-- > 0;) ax = qm_add(qm_scale(ax, px), qm_load(x + 4 * k));
        for (uint32_t k = ny; k-- > 0;) ay = qm_add(qm_scale(ay, px), qm_load(y + 4 * k));
        qm31 v = qm_add(ax, qm_scale(ay, py));
        memcpy(out + 4 * s, v.v, 16);
    }
    free(pts);
    return 0;
}

/* The same Horner evaluation at `count` selected storage indices only (a
 * size-independent check of a large codeword).                    :27-70,173-185 */
int ora_circle_evaluate_at(const uint32_t *x, uint32_t nx, const uint32_t *y, uint32_t ny, uint32_t log_n,
                           const uint32_t *idx, size_t count, uint32_t *out)
{
    if (log_n < 1 || log_n > 24) return -1;
    size_t n = (size_t)1 << log_n;
    if (nx > n / 2 || ny > n / 2) return -1;
    for (size_t i = 0; i < count; ++i)
        if (idx[i] >= n) return -1;
    uint32_t *pts = (uint32_t *)malloc(n * 2 * sizeof(uint32_t));
    ora_circle_domain_points(log_n, pts);
    for (size_t i = 0; i < count; ++i) {
        uint32_t px = pts[2 * (size_t)idx[i]], py = pts[2 * (size_t)idx[i] + 1];
        qm31 ax = {{0, 0, 0, 0}}, ay = {{0, 0, 0, 0}};
        for (uint32_t k = nx; k-- > 0;) ax = qm_add(qm_scale(ax, px), qm_load(x + 4 * (size_t)k));
        for (uint32_t k = ny; k-- > 0;) ay = qm_add(qm_scale(ay, px), qm_load(y + 4 * (size_t)k));
        qm31 v = qm_add(ax, qm_scale(ay, py));
        memcpy(out + 4 * i, v.v, 16);
    }
    free(pts);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* FRI fold                                            QM31FRIFold.swift:11-44 */
/* ------------------------------------------------------------------------- */

/* out = (l + r + ch * ((l - r) * t)) * 2^-1 with t an M31 twiddle embedded as (t,0,0,0). */
int ora_fold_layer(const uint32_t *in, size_t n, const uint32_t *tw, const uint32_t ch4[4], uint32_t *out)
{
    if (n < 2 || n % 2) return -1;
    qm31 ch = qm_load(ch4);
    for (size_t k = 0; k < n / 2; ++k) {
        qm31 l = qm_load(in + 8 * k), r = qm_load(in + 8 * k + 4);
        if (tw[k] == 0) return -1;
        qm31 m = qm_add(qm_add(l, r), qm_mul(ch, qm_scale(qm_sub(l, r), tw[k])));
        qm31 o = qm_scale(m, INV2);
        memcpy(out + 4 * k, o.v, 16);
    }
    return 0;
}

/* Explicit-challenge chain.                           CircleDomain.swift:513-544 */
int ora_circle_fold_chain(const uint32_t *evals, uint32_t log_n, const uint32_t *challenges,
                          uint32_t rounds, uint32_t *out)
{
    if (rounds < 1 || rounds > log_n) return -1;
    size_t n = (size_t)1 << log_n;
    uint32_t *tw = (uint32_t *)malloc(n * sizeof(uint32_t));
    uint32_t *a = (uint32_t *)malloc(n * 16), *b = (uint32_t *)malloc(n * 16);
    int rc = ora_circle_fold_twiddles(log_n, rounds, tw);
    memcpy(a, evals, n * 16);
    size_t off = 0, cur = n;
    for (uint32_t r = 0; r < rounds && !rc; ++r) {
        rc = ora_fold_layer(a, cur, tw + off, challenges + 4 * r, b);
        off += cur / 2;
        cur /= 2;
        uint32_t *t = a; a = b; b = t;
    }
    if (!rc) memcpy(out, a, cur * 16);
    free(tw); free(a); free(b);
    return rc;
}

/* ------------------------------------------------------------------------- */
/* Merkle                                                MerkleOracle.swift:4-141 */
/* ------------------------------------------------------------------------- */

static int is_pow2(size_t v) { return v && !(v & (v - 1)); }
static uint32_t ilog2(size_t v) { uint32_t r = 0; while (v > 1) { v >>= 1; ++r; } return r; }

/* All levels bottom-up: leaf hashes (count), then count/2, ..., root. */
int ora_merkle_tree(const uint8_t *raw, size_t count, size_t stride, size_t len, uint8_t *nodes)
{
    if (!is_pow2(count) || stride < len) return -1;
    for (size_t k = 0; k < count; ++k) ora_sha3_256(raw + k * stride, len, nodes + 32 * k);
    uint8_t *lvl = nodes;
    for (size_t m = count; m > 1; m /= 2) {
        uint8_t *nxt = lvl + 32 * m;
        for (size_t k = 0; k < m / 2; ++k) ora_sha3_256(lvl + 64 * k, 64, nxt + 32 * k);
        lvl = nxt;
    }
    return 0;
}

int ora_merkle_root(const uint8_t *raw, size_t count, size_t stride, size_t len, uint8_t root[32])
{
    if (!is_pow2(count) || stride < len) return -1;
    uint8_t *nodes = (uint8_t *)malloc((2 * count - 1) * 32);
    if (!nodes) return -2;
    ora_merkle_tree(raw, count, stride, len, nodes);
    memcpy(root, nodes + (2 * count - 2) * 32, 32);
    free(nodes);
    return 0;
}

static void tree_siblings(const uint8_t *nodes, size_t count, size_t index, uint8_t *sib)
{
    const uint8_t *lvl = nodes;
    size_t k = 0;
    for (size_t m = count; m > 1; m /= 2, ++k) {
        memcpy(sib + 32 * k, lvl + 32 * (index ^ 1), 32);
        lvl += 32 * m;
        index >>= 1;
    }
}

/* Sibling path bottom-up.                                              :53-110 */
int ora_merkle_opening(const uint8_t *raw, size_t count, size_t stride, size_t len, size_t index,
                       uint8_t *siblings, uint8_t root[32])
{
    if (!is_pow2(count) || stride < len || index >= count) return -1;
    uint8_t *nodes = (uint8_t *)malloc((2 * count - 1) * 32);
    if (!nodes) return -2;
    ora_merkle_tree(raw, count, stride, len, nodes);
    tree_siblings(nodes, count, index, siblings);
    memcpy(root, nodes + (2 * count - 2) * 32, 32);
    free(nodes);
    return 0;
}

/* 1 = valid, 0 = mismatch.                                             :112-141 */
int ora_merkle_verify(const uint8_t *leaf, size_t leaf_len, size_t index, const uint8_t *sib,
                      uint32_t nsib, const uint8_t root[32])
{
    if (nsib >= 63 || index >= ((size_t)1 << nsib)) return -1;
    uint8_t d[32], buf[64];
    ora_sha3_256(leaf, leaf_len, d);
    for (uint32_t k = 0; k < nsib; ++k) {
        if (((index >> k) & 1) == 0) { memcpy(buf, d, 32); memcpy(buf + 32, sib + 32 * k, 32); }
        else { memcpy(buf, sib + 32 * k, 32); memcpy(buf + 32, d, 32); }
        ora_sha3_256(buf, 64, d);
    }
    return memcmp(d, root, 32) == 0;
}

/* ------------------------------------------------------------------------- */
/* Circle FRI transcript V1                        CircleProofFormat.swift:684-983 */
/* ------------------------------------------------------------------------- */

static const char CFRI_DOMAIN[] = "AppleZKProver.CircleFRI.PCS.V1";

static size_t base_frame(uint8_t *f, uint8_t type)          /* :966-973 */
{
    size_t dl = sizeof(CFRI_DOMAIN) - 1;
    put_u32(f, (uint32_t)dl);
    memcpy(f + 4, CFRI_DOMAIN, dl);
    put_u32(f + 4 + dl, 1);
    f[8 + dl] = type;
    return 9 + dl;
}

static void absorb_frame_header(ora_transcript *t)           /* :906-914 */
{
    uint8_t f[96];
    size_t o = base_frame(f, 0);
    put_u32(f + o, 1); put_u32(f + o + 4, 1); put_u32(f + o + 8, ORA_P);
    put_u32(f + o + 12, CGEN.x); put_u32(f + o + 16, CGEN.y);
    ora_transcript_absorb(t, f, o + 20, 0x06);
}
static void absorb_frame_security(ora_transcript *t, const ora_security *s)   /* :916-924 */
{
    uint8_t f[96];
    size_t o = base_frame(f, 1);
    put_u32(f + o, s->log_blowup); put_u32(f + o + 4, s->query_count);
    put_u32(f + o + 8, s->folding_step); put_u32(f + o + 12, s->grinding_bits);
    put_u32(f + o + 16, s->log_blowup * s->query_count + s->grinding_bits);
    ora_transcript_absorb(t, f, o + 20, 0x06);
}
static void absorb_frame_u32s(ora_transcript *t, uint8_t type, const uint32_t *w, int nw)
{
    uint8_t f[96];
    size_t o = base_frame(f, type);
    for (int k = 0; k < nw; ++k) put_u32(f + o + 4 * k, w[k]);
    ora_transcript_absorb(t, f, o + 4 * (size_t)nw, 0x06);
}

static void transcript_prefix(ora_transcript *t, uint32_t log_n, const ora_security *sec, const uint8_t digest[32])
{                                                              /* :796-818 */
    uint8_t desc[44];
    ora_transcript_init(t);
    absorb_frame_header(t);
    ora_circle_domain_descriptor(log_n, desc);
    ora_transcript_absorb(t, desc, 44, 0x06);
    absorb_frame_security(t, sec);
    uint32_t w = 32;
    absorb_frame_u32s(t, 2, &w, 1);
    ora_transcript_absorb(t, digest, 32, 0x06);
}

/* commitment frame, root, challenge frame, squeeze 4 limbs mod p.      :819-828 */
static void transcript_round(ora_transcript *t, uint32_t r, const uint8_t root[32], uint32_t ch[4])
{
    uint32_t w2[2] = {r, 32};
    absorb_frame_u32s(t, 3, w2, 2);
    ora_transcript_absorb(t, root, 32, 0x06);
    uint32_t w3[3] = {r, ORA_P, 4};
    absorb_frame_u32s(t, 4, w3, 3);
    ora_transcript_squeeze_u32(t, 4, ORA_P, ch);
}

static void transcript_final_layer(ora_transcript *t, const uint32_t *fin, size_t n)   /* :852-859 */
{
    uint8_t f[96];
    size_t o = base_frame(f, 5);
    put_u64(f + o, n);
    put_u64(f + o + 8, n * 16);
    ora_transcript_absorb(t, f, o + 16, 0x06);
    uint8_t *bytes = (uint8_t *)malloc(n * 16);
    for (size_t k = 0; k < 4 * n; ++k) put_u32(bytes + 4 * k, fin[k]);
    ora_transcript_absorb(t, bytes, n * 16, 0x06);
    free(bytes);
}

/* grinding frame + nonce, then a 32-byte squeeze.                      :885-896,956-963 */
static void grinding_digest(ora_transcript *t, uint32_t bits, uint64_t nonce, uint8_t d[32])
{
    uint32_t w[2] = {bits, 8};
    absorb_frame_u32s(t, 7, w, 2);
    uint8_t nb[8];
    put_u64(nb, nonce);
    ora_transcript_absorb(t, nb, 8, 0x06);
    ora_transcript_squeeze_bytes(t, 32, d);
}
/* digestMeetsTarget: the first `bits` bits of the digest are zero, most significant bit of
   byte 0 first; more bits than the digest holds never meet the target.       :663-681 */
int ora_digest_meets_target(const uint8_t *d, size_t len, uint32_t bits)
{
    if ((uint64_t)bits > 8 * (uint64_t)len) return 0;
    uint32_t k = 0;
    while (bits >= 8) { if (d[k]) return 0; ++k; bits -= 8; }
    if (!bits) return 1;
    return (d[k] & (uint8_t)(0xffu << (8 - bits))) == 0;
}
static int digest_meets(const uint8_t *d, uint32_t bits) { return ora_digest_meets_target(d, 32, bits); }
void ora_circle_grinding_digest(const uint64_t state[25], uint32_t bits, uint64_t nonce, uint8_t out[32])
{
    ora_transcript t;
    memcpy(t.s, state, sizeof t.s);
    grinding_digest(&t, bits, nonce, out);
}

static void transcript_queries(ora_transcript *t, uint32_t q, uint64_t pairs, uint32_t *out)   /* :708-717 */
{
    uint8_t f[96];
    size_t o = base_frame(f, 6);
    put_u32(f + o, q);
    put_u64(f + o + 4, pairs);
    ora_transcript_absorb(t, f, o + 12, 0x06);
    ora_transcript_squeeze_u32(t, q, (uint32_t)pairs, out);
}

int ora_circle_transcript_challenges(uint32_t log_n, const ora_security *sec, const uint8_t digest[32],
                                     const uint8_t *commitments, uint32_t rounds, uint32_t *out)
{
    ora_transcript t;
    transcript_prefix(&t, log_n, sec, digest);
    for (uint32_t r = 0; r < rounds; ++r) transcript_round(&t, r, commitments + 32 * r, out + 4 * r);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Proof builder + codec                   CircleProofFormat.swift:475-645,1172-1395 */
/* ------------------------------------------------------------------------- */

typedef struct { uint8_t *p; size_t n, cap; int overflow; } wbuf;
static void w_bytes(wbuf *w, const void *d, size_t n)
{
    if (w->n + n > w->cap) { w->overflow = 1; w->n += n; return; }
    memcpy(w->p + w->n, d, n);
    w->n += n;
}
static void w_u32(wbuf *w, uint32_t v) { uint8_t b[4]; put_u32(b, v); w_bytes(w, b, 4); }
static void w_u64(wbuf *w, uint64_t v) { uint8_t b[8]; put_u64(b, v); w_bytes(w, b, 8); }
static void w_qm(wbuf *w, const uint32_t *q) { for (int k = 0; k < 4; ++k) w_u32(w, q[k]); }

/* opening = u64 leaf index, 16-B value, u32 sibling count, siblings.    :617-628 */
static void w_opening(wbuf *w, uint64_t leaf, const uint32_t *val, const uint8_t *nodes, size_t count)
{
    w_u64(w, leaf);
    w_qm(w, val);
    uint32_t ns = ilog2(count);
    w_u32(w, ns);
    uint8_t sib[32 * 32];
    tree_siblings(nodes, count, (size_t)leaf, sib);
    w_bytes(w, sib, 32 * (size_t)ns);
}

int ora_circle_fri_prove(const uint32_t *evals, uint32_t log_n, const ora_security *sec,
                         const uint8_t digest[32], uint32_t rounds, const uint32_t *claimed, uint32_t n_claimed,
                         uint8_t *out, size_t cap, size_t *out_len, ora_prove_info *info)
{
    if (log_n < 1 || log_n > 28 || rounds < 1 || rounds > log_n || rounds > 32 || sec->folding_step != 1 ||
        sec->grinding_bits > 20 || sec->log_blowup == 0 || sec->query_count == 0 || sec->query_count > 4096)
        return -1;
    size_t n = (size_t)1 << log_n;
    if (!qm_canonical(evals, 4 * n)) return -1;
    for (uint32_t k = 0; k < n_claimed; ++k)
        if (claimed[k] >= n || (k && claimed[k] <= claimed[k - 1])) return -1;

    uint32_t *tw = (uint32_t *)malloc(n * sizeof(uint32_t));
    int rc = ora_circle_fold_twiddles(log_n, rounds, tw);
    if (rc) { free(tw); return rc; }

    /* committed layers and their full trees */
    uint32_t **layer = (uint32_t **)calloc(rounds + 1, sizeof(uint32_t *));
    uint8_t **tree = (uint8_t **)calloc(rounds, sizeof(uint8_t *));
    uint8_t roots[32 * 32];
    uint32_t ch[32 * 4];
    layer[0] = (uint32_t *)malloc(n * 16);
    memcpy(layer[0], evals, n * 16);
    ora_transcript t;
    transcript_prefix(&t, log_n, sec, digest);
    size_t cur = n, toff = 0;
    for (uint32_t r = 0; r < rounds; ++r) {
        uint8_t *bytes = (uint8_t *)malloc(cur * 16);
        for (size_t k = 0; k < 4 * cur; ++k) put_u32(bytes + 4 * k, layer[r][k]);
        tree[r] = (uint8_t *)malloc((2 * cur - 1) * 32);
        ora_merkle_tree(bytes, cur, 16, 16, tree[r]);
        free(bytes);
        memcpy(roots + 32 * r, tree[r] + (2 * cur - 2) * 32, 32);
        transcript_round(&t, r, roots + 32 * r, ch + 4 * r);
        layer[r + 1] = (uint32_t *)malloc(cur / 2 * 16);
        ora_fold_layer(layer[r], cur, tw + toff, ch + 4 * r, layer[r + 1]);
        toff += cur / 2;
        cur /= 2;
    }
    const uint32_t *fin = layer[rounds];
    transcript_final_layer(&t, fin, cur);

    if (info) memcpy(info->pre_grinding, t.s, sizeof t.s);
    int have_nonce = 0;
    uint64_t nonce = 0;
    if (sec->grinding_bits > 0) {                               /* :753-794 */
        for (;; ++nonce) {
            ora_transcript c = t;
            uint8_t d[32];
            grinding_digest(&c, sec->grinding_bits, nonce, d);
            if (digest_meets(d, sec->grinding_bits)) break;
        }
        uint8_t d[32];
        grinding_digest(&t, sec->grinding_bits, nonce, d);
        have_nonce = 1;
    }
    uint32_t *qp = (uint32_t *)malloc(sec->query_count * sizeof(uint32_t));
    transcript_queries(&t, sec->query_count, n / 2, qp);

    /* encode */
    wbuf w = {out, 0, cap, 0};
    static const uint8_t magic[8] = {0x41, 0x5a, 0x4b, 0x50, 0x46, 0x56, 0x31, 0x00};
    w_bytes(&w, magic, 8);
    w_u32(&w, 1);
    w_u32(&w, 1);
    uint8_t desc[44];
    ora_circle_domain_descriptor(log_n, desc);
    w_u32(&w, 44);
    w_bytes(&w, desc, 44);
    w_u32(&w, sec->log_blowup); w_u32(&w, sec->query_count);
    w_u32(&w, sec->folding_step); w_u32(&w, sec->grinding_bits);
    w_u32(&w, 32);
    w_bytes(&w, digest, 32);
    w_u32(&w, rounds);
    w_bytes(&w, roots, 32 * (size_t)rounds);
    w_u64(&w, cur);
    for (size_t k = 0; k < 4 * cur; ++k) w_u32(&w, fin[k]);
    w_u32(&w, sec->query_count);
    for (uint32_t q = 0; q < sec->query_count; ++q) {            /* makeQuery :1278-1331 */
        size_t pair = qp[q];
        w_u64(&w, pair);
        w_u32(&w, rounds);
        size_t lc = n;
        for (uint32_t r = 0; r < rounds; ++r) {
            w_u32(&w, r);
            w_u64(&w, pair);
            w_opening(&w, 2 * pair, layer[r] + 8 * pair, tree[r], lc);
            w_opening(&w, 2 * pair + 1, layer[r] + 8 * pair + 4, tree[r], lc);
            pair >>= 1;
            lc >>= 1;
        }
    }
    if (n_claimed || have_nonce) {
        w_u32(&w, n_claimed);
        for (uint32_t k = 0; k < n_claimed; ++k) w_opening(&w, claimed[k], layer[0] + 4 * (size_t)claimed[k], tree[0], n);
    }
    if (have_nonce) w_u64(&w, nonce);
    *out_len = w.n;

    if (info) {
        memcpy(info->commitments, roots, 32 * (size_t)rounds);
        memcpy(info->challenges, ch, 16 * (size_t)rounds);
        info->n_queries = sec->query_count;
        info->grinding_nonce = nonce;
        memcpy(info->query_pairs, qp, 4 * (size_t)sec->query_count);
    }
    for (uint32_t r = 0; r <= rounds; ++r) free(layer[r]);
    for (uint32_t r = 0; r < rounds; ++r) free(tree[r]);
    free(layer); free(tree); free(tw); free(qp);
    return w.overflow ? -3 : 0;
}

/* ------------------------------------------------------------------------- */
/* Decoder + verifier          CircleProofFormat.swift:340-473,536-607,1397-1546 */
/* ------------------------------------------------------------------------- */

typedef struct { const uint8_t *p; size_t n, o; int bad; } rbuf;
static const uint8_t *r_bytes(rbuf *r, size_t k)
{
    if (r->bad || r->o + k > r->n) { r->bad = 1; return NULL; }
    const uint8_t *s = r->p + r->o;
    r->o += k;
    return s;
}
static uint32_t r_u32(rbuf *r) { const uint8_t *s = r_bytes(r, 4); return s ? get_u32(s) : 0; }
static uint64_t r_u64(rbuf *r) { const uint8_t *s = r_bytes(r, 8); return s ? get_u64(s) : 0; }

/* verify one opening against root: leaf index, sibling count, path.   :1500-1520 */
static int check_opening(rbuf *r, uint64_t want_leaf, uint32_t want_nsib, const uint8_t *root, uint32_t val[4])
{
    uint64_t leaf = r_u64(r);
    const uint8_t *v = r_bytes(r, 16);
    uint32_t ns = r_u32(r);
    if (r->bad || ns > 62) return 0;
    const uint8_t *sib = r_bytes(r, 32 * (size_t)ns);
    if (r->bad) return 0;
    for (int k = 0; k < 4; ++k) { val[k] = get_u32(v + 4 * k); if (val[k] >= ORA_P) { r->bad = 1; return 0; } }
    if (leaf != want_leaf || ns != want_nsib) return 0;
    return ora_merkle_verify(v, 16, (size_t)leaf, sib, ns, root) == 1;
}

/* 1 = accept, 0 = reject, <0 = malformed encoding. */
int ora_circle_fri_verify(const uint8_t *proof, size_t len, const uint8_t digest[32])
{
    rbuf r = {proof, len, 0, 0};
    static const uint8_t magic[8] = {0x41, 0x5a, 0x4b, 0x50, 0x46, 0x56, 0x31, 0x00};
    const uint8_t *m = r_bytes(&r, 8);
    if (!m || memcmp(m, magic, 8)) return -1;
    if (r_u32(&r) != 1 || r_u32(&r) != 1) return -1;
    if (r_u32(&r) != 44) return -1;
    const uint8_t *desc = r_bytes(&r, 44);
    if (!desc) return -1;
    uint32_t log_n = get_u32(desc + 28);
    uint8_t want[44];
    if (log_n < 1 || log_n > 28 || ora_circle_domain_descriptor(log_n, want) || memcmp(desc, want, 44)) return -1;
    ora_security sec;
    sec.log_blowup = r_u32(&r); sec.query_count = r_u32(&r);
    sec.folding_step = r_u32(&r); sec.grinding_bits = r_u32(&r);
    if (r_u32(&r) != 32) return -1;
    const uint8_t *dg = r_bytes(&r, 32);
    uint32_t rounds = r_u32(&r);
    if (r.bad || !dg || rounds < 1 || rounds > log_n || rounds > 32) return -1;
    if (sec.folding_step != 1 || sec.log_blowup == 0 || sec.query_count == 0 || sec.query_count > 4096 || sec.grinding_bits > 64)
        return -1;
    const uint8_t *roots = r_bytes(&r, 32 * (size_t)rounds);
    uint64_t nfin = r_u64(&r);
    size_t n = (size_t)1 << log_n;
    if (r.bad || nfin != (n >> rounds)) return -1;
    const uint8_t *finb = r_bytes(&r, 16 * (size_t)nfin);
    if (!finb) return -1;
    uint32_t *fin = (uint32_t *)malloc(16 * (size_t)nfin);
    for (size_t k = 0; k < 4 * nfin; ++k) {
        fin[k] = get_u32(finb + 4 * k);
        if (fin[k] >= ORA_P) { free(fin); return -1; }
    }
    int ok = memcmp(dg, digest, 32) == 0;

    /* transcript */
    ora_transcript t;
    transcript_prefix(&t, log_n, &sec, dg);
    uint32_t ch[32 * 4];
    for (uint32_t i = 0; i < rounds; ++i) transcript_round(&t, i, roots + 32 * i, ch + 4 * i);
    transcript_final_layer(&t, fin, nfin);

    uint32_t nq = r_u32(&r);
    if (r.bad || nq != sec.query_count) { free(fin); return -1; }
    size_t qstart = r.o;
    /* skip queries to reach the optional tail (claimed openings + nonce) */
    for (uint32_t q = 0; q < nq && !r.bad; ++q) {
        r_u64(&r);
        uint32_t nl = r_u32(&r);
        if (nl != rounds) { r.bad = 1; break; }
        for (uint32_t l = 0; l < nl && !r.bad; ++l) {
            r_u32(&r); r_u64(&r);
            for (int side = 0; side < 2 && !r.bad; ++side) {
                r_u64(&r); r_bytes(&r, 16);
                uint32_t ns = r_u32(&r);
                if (ns > 62) { r.bad = 1; break; }
                r_bytes(&r, 32 * (size_t)ns);
            }
        }
    }
    if (r.bad) { free(fin); return -1; }
    int have_nonce = 0;
    uint64_t nonce = 0;
    if (r.o < r.n) {
        uint32_t nc = r_u32(&r);
        for (uint32_t k = 0; k < nc && !r.bad; ++k) {
            uint64_t leaf = r_u64(&r);
            r_bytes(&r, 16);
            uint32_t ns = r_u32(&r);
            if (ns != log_n || leaf >= n) { r.bad = 1; break; }
            r_bytes(&r, 32 * (size_t)ns);
        }
        if (r.o < r.n) { nonce = r_u64(&r); have_nonce = 1; }
    }
    if (r.bad || r.o != r.n) { free(fin); return -1; }
    if ((sec.grinding_bits > 0) != have_nonce) { free(fin); return -1; }
    if (have_nonce) {
        uint8_t d[32];
        grinding_digest(&t, sec.grinding_bits, nonce, d);
        if (!digest_meets(d, sec.grinding_bits)) ok = 0;
    }
    uint32_t *qp = (uint32_t *)malloc(4 * (size_t)nq);
    transcript_queries(&t, nq, n / 2, qp);

    uint32_t *tw = (uint32_t *)malloc(n * sizeof(uint32_t));
    ora_circle_fold_twiddles(log_n, rounds, tw);
    size_t *toff = (size_t *)malloc(sizeof(size_t) * rounds);
    for (uint32_t i = 0, o = 0; i < rounds; ++i) { toff[i] = o; o += (uint32_t)(n >> (i + 1)); }

    rbuf rq = {proof, len, qstart, 0};
    for (uint32_t q = 0; q < nq && ok; ++q) {                   /* :1420-1480 */
        uint64_t ip = r_u64(&rq);
        r_u32(&rq);
        if (ip != qp[q]) { ok = 0; break; }
        uint64_t pair = ip;
        int have_exp = 0;
        uint64_t exp_idx = 0;
        uint32_t exp_val[4];
        for (uint32_t i = 0; i < rounds && ok; ++i) {
            size_t lc = n >> i;
            uint32_t li = r_u32(&rq);
            uint64_t pi = r_u64(&rq);
            if (li != i || pi != pair || pair >= lc / 2) { ok = 0; break; }
            uint32_t lv[4], rv[4];
            int okl = check_opening(&rq, 2 * pair, ilog2(lc), roots + 32 * i, lv);
            int okr = check_opening(&rq, 2 * pair + 1, ilog2(lc), roots + 32 * i, rv);
            if (!okl || !okr) { ok = 0; break; }
            if (have_exp) {
                const uint32_t *cmp = exp_idx == 2 * pair ? lv : exp_idx == 2 * pair + 1 ? rv : NULL;
                if (!cmp || memcmp(cmp, exp_val, 16)) { ok = 0; break; }
            }
            uint32_t pairv[8];
            memcpy(pairv, lv, 16);
            memcpy(pairv + 4, rv, 16);
            ora_fold_layer(pairv, 2, tw + toff[i] + pair, ch + 4 * i, exp_val);
            exp_idx = pair;
            have_exp = 1;
            pair >>= 1;
        }
        if (ok && (exp_idx >= nfin || memcmp(fin + 4 * exp_idx, exp_val, 16))) ok = 0;
    }
    free(fin); free(qp); free(tw); free(toff);
    return ok;
}

/* ------------------------------------------------------------------------- */
/* Sum-check                                       SumcheckOracle.swift:212-276, */
/*                                          817-939,1066-1106,1288-1299,1311-1386 */
/* ------------------------------------------------------------------------- */

static size_t frame_dom(uint8_t *f, const char *dom, uint8_t type)
{
    size_t dl = strlen(dom);
    put_u32(f, (uint32_t)dl);
    memcpy(f + 4, dom, dl);
    put_u32(f + 4 + dl, 1);
    f[8 + dl] = type;
    return 9 + dl;
}

int ora_sumcheck_chunk(const uint32_t *ev, size_t n, uint32_t rounds, uint32_t *final_out,
                       uint32_t *coeffs_out, uint32_t *challenges_out)
{
    if (n < 2 || !is_pow2(n) || rounds < 1 || rounds > ilog2(n) || !qm_canonical(ev, n)) return -1;
    static const char dom[] = "AppleZKProver.SumcheckChunk";
    ora_transcript t;
    ora_transcript_init(&t);
    uint8_t f[128];
    size_t o = frame_dom(f, dom, 0);
    put_u64(f + o, n); put_u32(f + o + 8, rounds); put_u32(f + o + 12, ORA_P);
    ora_transcript_absorb(&t, f, o + 16, 0x06);
    uint32_t *cur = (uint32_t *)malloc(n * 4), *nxt = (uint32_t *)malloc(n * 4);
    uint8_t *bytes = (uint8_t *)malloc(n * 4);
    memcpy(cur, ev, n * 4);
    size_t m = n, co = 0;
    for (uint32_t r = 0; r < rounds; ++r) {
        o = frame_dom(f, dom, 1);
        put_u32(f + o, r); put_u64(f + o + 4, m); put_u64(f + o + 12, m);
        ora_transcript_absorb(&t, f, o + 20, 0x06);
        for (size_t k = 0; k < m; ++k) put_u32(bytes + 4 * k, cur[k]);
        ora_transcript_absorb(&t, bytes, 4 * m, 0x06);
        o = frame_dom(f, dom, 2);
        put_u32(f + o, r); put_u32(f + o + 4, ORA_P);
        ora_transcript_absorb(&t, f, o + 8, 0x06);
        uint32_t c;
        ora_transcript_squeeze_u32(&t, 1, ORA_P, &c);
        challenges_out[r] = c;
        memcpy(coeffs_out + co, cur, 4 * m);
        co += m;
        for (size_t k = 0; k < m / 2; ++k) nxt[k] = ora_m31_add(cur[2 * k], ora_m31_mul(cur[2 * k + 1], c));
        m /= 2;
        uint32_t *tmp = cur; cur = nxt; nxt = tmp;
    }
    memcpy(final_out, cur, 4 * m);
    free(cur); free(nxt); free(bytes);
    return 0;
}

void ora_sumcheck_words_digest(const uint32_t *w, size_t n, uint8_t out[32])
{
    static const char dom[] = "AppleZKProver.M31Sumcheck.Words.V1";
    size_t dl = sizeof(dom) - 1, tot = 4 + dl + 4 + 8 + 4 * n;
    uint8_t *b = (uint8_t *)malloc(tot);
    put_u32(b, (uint32_t)dl);
    memcpy(b + 4, dom, dl);
    put_u32(b + 4 + dl, 1);
    put_u64(b + 8 + dl, n);
    for (size_t k = 0; k < n; ++k) put_u32(b + 16 + dl + 4 * k, w[k]);
    ora_sha3_256(b, tot, out);
    free(b);
}

int ora_sumcheck_multilinear(const uint32_t *ev, size_t n, uint32_t *round_polys, uint32_t *challenges,
                             uint32_t *final_eval, uint32_t *claimed_sum, uint8_t digest_out[32])
{
    if (n < 2 || !is_pow2(n) || !qm_canonical(ev, n)) return -1;
    static const char dom[] = "AppleZKProver.M31MultilinearSumcheck";
    uint32_t rounds = ilog2(n), sum = 0;
    for (size_t k = 0; k < n; ++k) sum = ora_m31_add(sum, ev[k]);
    ora_sumcheck_words_digest(ev, n, digest_out);
    *claimed_sum = sum;
    ora_transcript t;
    ora_transcript_init(&t);
    uint8_t f[160];
    size_t o = frame_dom(f, dom, 0);
    put_u64(f + o, n); put_u32(f + o + 8, rounds); put_u32(f + o + 12, sum);
    memcpy(f + o + 16, digest_out, 32);
    ora_transcript_absorb(&t, f, o + 48, 0x06);
    uint32_t *cur = (uint32_t *)malloc(n * 4);
    memcpy(cur, ev, n * 4);
    size_t m = n;
    for (uint32_t r = 0; r < rounds; ++r) {
        uint32_t e0 = 0, e1 = 0;
        for (size_t k = 0; k < m / 2; ++k) { e0 = ora_m31_add(e0, cur[2 * k]); e1 = ora_m31_add(e1, cur[2 * k + 1]); }
        round_polys[2 * r] = e0;
        round_polys[2 * r + 1] = e1;
        o = frame_dom(f, dom, 1);
        put_u32(f + o, r); put_u32(f + o + 4, e0); put_u32(f + o + 8, e1);
        ora_transcript_absorb(&t, f, o + 12, 0x06);
        o = frame_dom(f, dom, 2);
        put_u32(f + o, r); put_u32(f + o + 4, ORA_P);
        ora_transcript_absorb(&t, f, o + 8, 0x06);
        uint32_t c;
        ora_transcript_squeeze_u32(&t, 1, ORA_P, &c);
        challenges[r] = c;
        uint32_t omc = ora_m31_sub(1, c);
        for (size_t k = 0; k < m / 2; ++k)
            cur[k] = ora_m31_add(ora_m31_mul(omc, cur[2 * k]), ora_m31_mul(c, cur[2 * k + 1]));
        m /= 2;
    }
    *final_eval = cur[0];
    free(cur);
    return 0;
}

/* M31MultilinearSumcheckProofCodecV1.encode, SumcheckOracle.swift:1176-1194, of the
 * proof built above: magic "AZKMSV1\0", u32 proof version, u32 statement version,
 * u64 lanes, u32 rounds, u32 claimed sum, 32-byte digest, u64-counted initial words,
 * u64-counted (e0, e1) pairs, u64-counted challenges, u32 final evaluation. */
size_t ora_sumcheck_multilinear_proof(const uint32_t *ev, size_t n, uint8_t *out, size_t cap)
{
    if (n < 2 || !is_pow2(n)) return 0;
    uint32_t rounds = ilog2(n);
    size_t need = 8 + 4 + 4 + 8 + 4 + 4 + 32 + (8 + 4 * n) + (8 + 8 * (size_t)rounds) + (8 + 4 * (size_t)rounds) + 4;
    if (cap < need) return 0;
    uint32_t *rp = (uint32_t *)malloc(8 * rounds), *ch = (uint32_t *)malloc(4 * rounds), fin, cs;
    uint8_t dg[32];
    if (ora_sumcheck_multilinear(ev, n, rp, ch, &fin, &cs, dg)) { free(rp); free(ch); return 0; }
    static const uint8_t magic[8] = {0x41, 0x5a, 0x4b, 0x4d, 0x53, 0x56, 0x31, 0x00};
    uint8_t *p = out;
    memcpy(p, magic, 8); p += 8;
    put_u32(p, 1); p += 4;
    put_u32(p, 1); p += 4;
    put_u64(p, n); p += 8;
    put_u32(p, rounds); p += 4;
    put_u32(p, cs); p += 4;
    memcpy(p, dg, 32); p += 32;
    put_u64(p, n); p += 8;
    for (size_t k = 0; k < n; ++k, p += 4) put_u32(p, ev[k]);
    put_u64(p, rounds); p += 8;
    for (uint32_t r = 0; r < 2 * rounds; ++r, p += 4) put_u32(p, rp[r]);
    put_u64(p, rounds); p += 8;
    for (uint32_t r = 0; r < rounds; ++r, p += 4) put_u32(p, ch[r]);
    put_u32(p, fin); p += 4;
    free(rp);
    free(ch);
    return (size_t)(p - out);
}

/* ------------------------------------------------------------------------- */
/* SplitMix64 inputs                         Planner/MetalProofPlanner.swift:758-789 */
/* ------------------------------------------------------------------------- */

static uint64_t smix_next(uint64_t *s)
{
    *s += 0x9e3779b97f4a7c15ull;
    uint64_t z = *s;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

void ora_splitmix_bytes(uint64_t seed, size_t n, uint8_t *out)
{
    uint64_t s = seed ? seed : 0x243f6a8885a308d3ull, cur = 0;
    int rem = 0;
    for (size_t k = 0; k < n; ++k) {
        if (!rem) { cur = smix_next(&s); rem = 8; }
        out[k] = (uint8_t)cur;
        cur >>= 8;
        --rem;
    }
}

/* u32 words from the same LE byte stream, masked to 31 bits, value p rejected. */
void ora_splitmix_m31(uint64_t seed, size_t n, uint32_t *out)
{
    uint64_t s = seed ? seed : 0x243f6a8885a308d3ull;
    size_t k = 0;
    while (k < n) {
        uint64_t z = smix_next(&s);
        uint32_t w0 = (uint32_t)z & 0x7fffffffu, w1 = (uint32_t)(z >> 32) & 0x7fffffffu;
        if (w0 != ORA_P) out[k++] = w0;
        if (k < n && w1 != ORA_P) out[k++] = w1;
    }
}
