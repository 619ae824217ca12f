// Host-side Keccak for the parts of the path that the reference also runs on the CPU
// after readback: the final-layer absorb and query squeeze of the Circle FRI transcript
// (CirclePCSFRIResidentProverV1.prove re-derives them on the CPU,
// CircleFRIFoldPlan.swift:875-897), the constant transcript prefix, and the *_verified
// CPU gates (SHA3BatchHasher.swift:211-263).
#pragma once
#include <stdint.h>
#include <string.h>

namespace azk::host {

inline uint64_t rol(uint64_t v, int n) { return (v << n) | (v >> (64 - n)); }

constexpr uint64_t kRoundConstants[24] = {
    0x0000000000000001ull, 0x0000000000008082ull, 0x800000000000808Aull, 0x8000000080008000ull,
    0x000000000000808Bull, 0x0000000080000001ull, 0x8000000080008081ull, 0x8000000000008009ull,
    0x000000000000008Aull, 0x0000000000000088ull, 0x0000000080008009ull, 0x000000008000000Aull,
    0x000000008000808Bull, 0x800000000000008Bull, 0x8000000000008089ull, 0x8000000000008003ull,
    0x8000000000008002ull, 0x8000000000000080ull, 0x000000000000800Aull, 0x800000008000000Aull,
    0x8000000080008081ull, 0x8000000000008080ull, 0x0000000080000001ull, 0x8000000080008008ull};

// One round from the 25 lanes S<xy> (x + 5y numbering: S00..S24) into D<xy>.  The five
// chi rows are formed from their pi/rho sources directly: row y of the output takes lanes
// (x + 3y, x) of the input for x = 0..4, each with its theta column term and rotation.
#define AZK_KECCAK_ROUND(S, D, rc)                                                                              \
    do {                                                                                                        \
        const uint64_t c0 = S##00 ^ S##05 ^ S##10 ^ S##15 ^ S##20, c1 = S##01 ^ S##06 ^ S##11 ^ S##16 ^ S##21;    \
        const uint64_t c2 = S##02 ^ S##07 ^ S##12 ^ S##17 ^ S##22, c3 = S##03 ^ S##08 ^ S##13 ^ S##18 ^ S##23;    \
        const uint64_t c4 = S##04 ^ S##09 ^ S##14 ^ S##19 ^ S##24;                                               \
        const uint64_t d0 = c4 ^ rol(c1, 1), d1 = c0 ^ rol(c2, 1), d2 = c1 ^ rol(c3, 1), d3 = c2 ^ rol(c4, 1);    \
        const uint64_t d4 = c3 ^ rol(c0, 1);                                                                    \
        uint64_t b0 = S##00 ^ d0, b1 = rol(S##06 ^ d1, 44), b2 = rol(S##12 ^ d2, 43), b3 = rol(S##18 ^ d3, 21);  \
        uint64_t b4 = rol(S##24 ^ d4, 14);                                                                      \
        D##00 = b0 ^ (~b1 & b2) ^ (rc); D##01 = b1 ^ (~b2 & b3); D##02 = b2 ^ (~b3 & b4);                       \
        D##03 = b3 ^ (~b4 & b0); D##04 = b4 ^ (~b0 & b1);                                                       \
        b0 = rol(S##03 ^ d3, 28); b1 = rol(S##09 ^ d4, 20); b2 = rol(S##10 ^ d0, 3); b3 = rol(S##16 ^ d1, 45);    \
        b4 = rol(S##22 ^ d2, 61);                                                                               \
        D##05 = b0 ^ (~b1 & b2); D##06 = b1 ^ (~b2 & b3); D##07 = b2 ^ (~b3 & b4);                              \
        D##08 = b3 ^ (~b4 & b0); D##09 = b4 ^ (~b0 & b1);                                                       \
        b0 = rol(S##01 ^ d1, 1); b1 = rol(S##07 ^ d2, 6); b2 = rol(S##13 ^ d3, 25); b3 = rol(S##19 ^ d4, 8);      \
        b4 = rol(S##20 ^ d0, 18);                                                                               \
        D##10 = b0 ^ (~b1 & b2); D##11 = b1 ^ (~b2 & b3); D##12 = b2 ^ (~b3 & b4);                              \
        D##13 = b3 ^ (~b4 & b0); D##14 = b4 ^ (~b0 & b1);                                                       \
        b0 = rol(S##04 ^ d4, 27); b1 = rol(S##05 ^ d0, 36); b2 = rol(S##11 ^ d1, 10); b3 = rol(S##17 ^ d2, 15);  \
        b4 = rol(S##23 ^ d3, 56);                                                                               \
        D##15 = b0 ^ (~b1 & b2); D##16 = b1 ^ (~b2 & b3); D##17 = b2 ^ (~b3 & b4);                              \
        D##18 = b3 ^ (~b4 & b0); D##19 = b4 ^ (~b0 & b1);                                                       \
        b0 = rol(S##02 ^ d2, 62); b1 = rol(S##08 ^ d3, 55); b2 = rol(S##14 ^ d4, 39); b3 = rol(S##15 ^ d0, 41);  \
        b4 = rol(S##21 ^ d1, 2);                                                                                \
        D##20 = b0 ^ (~b1 & b2); D##21 = b1 ^ (~b2 & b3); D##22 = b2 ^ (~b3 & b4);                              \
        D##23 = b3 ^ (~b4 & b0); D##24 = b4 ^ (~b0 & b1);                                                       \
    } while (0)

// Keccak-f[1600] (FIPS 202), lanes a[x + 5y].  The state lives in two sets of 25 locals
// and alternates between them round by round, so the compiler keeps it in registers
// (about twice the throughput of an array-indexed loop under clang).
inline void keccak_f(uint64_t a[25])
{
    uint64_t s00 = a[0], s01 = a[1], s02 = a[2], s03 = a[3], s04 = a[4], s05 = a[5], s06 = a[6], s07 = a[7];
    uint64_t s08 = a[8], s09 = a[9], s10 = a[10], s11 = a[11], s12 = a[12], s13 = a[13], s14 = a[14], s15 = a[15];
    uint64_t s16 = a[16], s17 = a[17], s18 = a[18], s19 = a[19], s20 = a[20], s21 = a[21], s22 = a[22], s23 = a[23];
    uint64_t s24 = a[24];
    uint64_t t00, t01, t02, t03, t04, t05, t06, t07, t08, t09, t10, t11, t12, t13, t14, t15, t16, t17, t18, t19, t20,
        t21, t22, t23, t24;
    for (int r = 0; r < 24; r += 2) {
        AZK_KECCAK_ROUND(s, t, kRoundConstants[r]);
        AZK_KECCAK_ROUND(t, s, kRoundConstants[r + 1]);
    }
    a[0] = s00; a[1] = s01; a[2] = s02; a[3] = s03; a[4] = s04; a[5] = s05; a[6] = s06; a[7] = s07; a[8] = s08;
    a[9] = s09; a[10] = s10; a[11] = s11; a[12] = s12; a[13] = s13; a[14] = s14; a[15] = s15; a[16] = s16;
    a[17] = s17; a[18] = s18; a[19] = s19; a[20] = s20; a[21] = s21; a[22] = s22; a[23] = s23; a[24] = s24;
}
#undef AZK_KECCAK_ROUND

inline void xor_block(uint64_t a[25], const uint8_t *p, size_t n)
{
    for (size_t i = 0; i < n / 8; ++i) {
        uint64_t v;
        memcpy(&v, p + 8 * i, 8);
        a[i] ^= v;
    }
}

// Transcript sponge: each absorb is one SHA3-padded message into the running state;
// squeezes read a copy (SHA3Oracle.TranscriptState, SHA3Oracle.swift:8-98).
struct Transcript {
    uint64_t s[25] = {0};
    void absorb(const uint8_t *p, size_t len, uint8_t suffix = 0x06)
    {
        size_t off = 0;
        for (; off + 136 <= len; off += 136) {
            xor_block(s, p + off, 136);
            keccak_f(s);
        }
        uint8_t tail[136] = {0};
        memcpy(tail, p + off, len - off);
        tail[len - off] ^= suffix;
        tail[135] ^= 0x80;
        xor_block(s, tail, 136);
        keccak_f(s);
    }
    // rejection sampling against 2^32 - (2^32 mod m), 34 words per block
    void squeeze_u32(uint32_t count, uint32_t modulus, uint32_t *out) const
    {
        uint64_t b[25];
        memcpy(b, s, sizeof b);
        const uint64_t limit = (1ull << 32) - ((1ull << 32) % modulus);
        uint32_t got = 0;
        for (uint64_t cand = 0; got < count; ++cand) {
            if (cand > 0 && cand % 34 == 0) keccak_f(b);
            uint32_t wi = (uint32_t)(cand % 34);
            uint64_t w = (uint32_t)(b[wi >> 1] >> ((wi & 1) * 32));
            if (w < limit) out[got++] = (uint32_t)(w % modulus);
        }
    }
};

// SHA3-256 (suffix 0x06) or Keccak-256 (suffix 0x01) of an arbitrary message.
inline void sha3_256(const uint8_t *p, size_t len, uint8_t out[32], uint8_t suffix = 0x06)
{
    Transcript t;
    t.absorb(p, len, suffix);
    memcpy(out, t.s, 32);
}

// M31SumcheckEncodingV1.digestWords (SumcheckOracle.swift:1288-1299): SHA3-256 of u32 len ||
// "AppleZKProver.M31Sumcheck.Words.V1" || u32 1 || u64 n || the words LE (sumcheck.hip).
void sumcheck_words_digest(const uint32_t *words, uint64_t n, uint8_t out[32]);

} // namespace azk::host
