/*
 * CPU oracle for the Circle-PCS commit path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference's CPU oracles (Swift sources of
 * AppleZKProver, cited per function in azk_oracle.c).  Only tests/, the smoke
 * check and the bench's CPU-baseline leg may load this library, and only as the
 * checker; the product (apple_amd/libazk.so) never links or calls it.
 *
 * Every function that returns int returns 0 on success and a negative value
 * on a rejected layout (the reference throws invalidInputLayout there).
 */
#ifndef AZK_ORACLE_H
#define AZK_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORA_P 2147483647u

/* ---- Keccak / SHA3 ---- */
void ora_keccak_f1600(uint64_t st[25]);
int ora_sponge(const uint8_t *data, size_t len, int rate, uint8_t suffix, uint8_t *out, size_t outlen);
void ora_sha3_256(const uint8_t *data, size_t len, uint8_t out[32]);
void ora_keccak_256(const uint8_t *data, size_t len, uint8_t out[32]);
int ora_keccakf_batch(const uint8_t *in, size_t in_stride, uint8_t *out, size_t out_stride, size_t count);
int ora_hash_oneblock_batch(const uint8_t *in, size_t stride, size_t len, uint8_t *out,
                            size_t out_stride, size_t count, uint8_t suffix);

/* ---- transcript sponge ---- */
typedef struct { uint64_t s[25]; } ora_transcript;
void ora_transcript_init(ora_transcript *t);
void ora_transcript_absorb(ora_transcript *t, const uint8_t *data, size_t len, uint8_t suffix);
int ora_transcript_squeeze_u32(const ora_transcript *t, uint32_t count, uint32_t modulus, uint32_t *out);
void ora_transcript_squeeze_bytes(const ora_transcript *t, size_t count, uint8_t *out);

/* ---- fields ---- */
uint32_t ora_m31_add(uint32_t a, uint32_t b);
uint32_t ora_m31_sub(uint32_t a, uint32_t b);
uint32_t ora_m31_neg(uint32_t a);
uint32_t ora_m31_mul(uint32_t a, uint32_t b);
uint32_t ora_m31_inv(uint32_t a);
void ora_qm31_add(const uint32_t a[4], const uint32_t b[4], uint32_t out[4]);
void ora_qm31_sub(const uint32_t a[4], const uint32_t b[4], uint32_t out[4]);
void ora_qm31_mul(const uint32_t a[4], const uint32_t b[4], uint32_t out[4]);
void ora_qm31_square(const uint32_t a[4], uint32_t out[4]);
int ora_qm31_inv(const uint32_t a[4], uint32_t out[4]);

/* ---- Circle domain ---- */
void ora_circle_point_from_index(uint32_t index, uint32_t out_xy[2]);
int ora_circle_point_index(uint32_t log_n, uint32_t natural, uint32_t *out_index);
int ora_circle_domain_points(uint32_t log_n, uint32_t *out_xy);            /* storage order, 2N words */
int ora_circle_fft_twiddles(uint32_t log_n, uint32_t *out);                 /* N-1 words */
int ora_circle_fold_twiddles(uint32_t log_n, uint32_t rounds, uint32_t *out); /* sum N>>(r+1) words */
int ora_circle_domain_descriptor(uint32_t log_n, uint8_t out[44]);

/* ---- codeword ---- */
int ora_circle_fft(const uint32_t *coeffs, uint32_t log_n, uint32_t *out);
int ora_circle_basis_from_monomial(const uint32_t *x, uint32_t nx, const uint32_t *y, uint32_t ny,
                                   uint32_t log_n, uint32_t *out);
int ora_circle_evaluate_direct(const uint32_t *x, uint32_t nx, const uint32_t *y, uint32_t ny,
                               uint32_t log_n, uint32_t *out);
int ora_circle_evaluate_at(const uint32_t *x, uint32_t nx, const uint32_t *y, uint32_t ny, uint32_t log_n,
                           const uint32_t *idx, size_t count, uint32_t *out);

/* ---- fold ---- */
int ora_fold_layer(const uint32_t *in, size_t n, const uint32_t *tw, const uint32_t ch[4], uint32_t *out);
int ora_circle_fold_chain(const uint32_t *evals, uint32_t log_n, const uint32_t *challenges,
                          uint32_t rounds, uint32_t *out);

/* ---- Merkle ---- */
int ora_merkle_root(const uint8_t *raw, size_t count, size_t stride, size_t len, uint8_t root[32]);
int ora_merkle_tree(const uint8_t *raw, size_t count, size_t stride, size_t len, uint8_t *nodes);
int ora_merkle_opening(const uint8_t *raw, size_t count, size_t stride, size_t len, size_t index,
                       uint8_t *siblings, uint8_t root[32]);
int ora_merkle_verify(const uint8_t *leaf, size_t leaf_len, size_t index, const uint8_t *siblings,
                      uint32_t nsib, const uint8_t root[32]);

/* ---- Circle PCS / FRI proof ---- */
typedef struct {
    uint32_t log_blowup, query_count, folding_step, grinding_bits;
} ora_security;

typedef struct {
    uint8_t commitments[32 * 32];   /* up to 32 rounds */
    uint32_t challenges[32 * 4];
    uint32_t query_pairs[4096];
    uint32_t n_queries;
    uint64_t grinding_nonce;        /* 0 without grinding */
    uint64_t pre_grinding[25];      /* transcript state after the final layer, before the grinding frame */
} ora_prove_info;

int ora_circle_fri_prove(const uint32_t *evals, uint32_t log_n, const ora_security *sec,
                         const uint8_t digest[32], uint32_t rounds, const uint32_t *claimed, uint32_t n_claimed,
                         uint8_t *out, size_t cap, size_t *out_len, ora_prove_info *info);
int ora_circle_fri_verify(const uint8_t *proof, size_t len, const uint8_t digest[32]);
/* digestMeetsTarget over a digest of any length (CircleProofFormat.swift:663-681). */
int ora_digest_meets_target(const uint8_t *digest, size_t len, uint32_t bits);
/* Grinding digest of `nonce` from a pre-grinding transcript state (:885-896,956-963). */
void ora_circle_grinding_digest(const uint64_t state[25], uint32_t bits, uint64_t nonce, uint8_t out[32]);
int ora_circle_transcript_challenges(uint32_t log_n, const ora_security *sec, const uint8_t digest[32],
                                     const uint8_t *commitments, uint32_t rounds, uint32_t *out);

/* ---- sum-check ---- */
int ora_sumcheck_chunk(const uint32_t *evals, size_t n, uint32_t rounds, uint32_t *final_out,
                       uint32_t *coeffs_out, uint32_t *challenges_out);
int ora_sumcheck_multilinear(const uint32_t *evals, size_t n, uint32_t *round_polys, uint32_t *challenges,
                             uint32_t *final_eval, uint32_t *claimed_sum, uint8_t digest_out[32]);
void ora_sumcheck_words_digest(const uint32_t *words, size_t n, uint8_t out[32]);
/* codec bytes of the multilinear proof; returns the length, 0 on a bad layout */
size_t ora_sumcheck_multilinear_proof(const uint32_t *evals, size_t n, uint8_t *out, size_t cap);

/* ---- seeded inputs ---- */
void ora_splitmix_bytes(uint64_t seed, size_t n, uint8_t *out);
void ora_splitmix_m31(uint64_t seed, size_t n, uint32_t *out);

#ifdef __cplusplus
}
#endif
#endif
