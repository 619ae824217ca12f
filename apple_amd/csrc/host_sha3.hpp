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

// Unrolled Keccak-f[1600] (FIPS 202), lanes a[x + 5y].
inline void keccak_f(uint64_t a[25])
{
    static const uint64_t rc[24] = {
        0x0000000000000001ull, 0x0000000000008082ull, 0x800000000000808Aull, 0x8000000080008000ull,
        0x000000000000808Bull, 0x0000000080000001ull, 0x8000000080008081ull, 0x8000000000008009ull,
        0x000000000000008Aull, 0x0000000000000088ull, 0x0000000080008009ull, 0x000000008000000Aull,
        0x000000008000808Bull, 0x800000000000008Bull, 0x8000000000008089ull, 0x8000000000008003ull,
        0x8000000000008002ull, 0x8000000000000080ull, 0x000000000000800Aull, 0x800000008000000Aull,
        0x8000000080008081ull, 0x8000000000008080ull, 0x0000000080000001ull, 0x8000000080008008ull};
    for (int round = 0; round < 24; ++round) {
        uint64_t c0 = a[0] ^ a[5] ^ a[10] ^ a[15] ^ a[20], c1 = a[1] ^ a[6] ^ a[11] ^ a[16] ^ a[21];
        uint64_t c2 = a[2] ^ a[7] ^ a[12] ^ a[17] ^ a[22], c3 = a[3] ^ a[8] ^ a[13] ^ a[18] ^ a[23];
        uint64_t c4 = a[4] ^ a[9] ^ a[14] ^ a[19] ^ a[24];
        uint64_t d0 = c4 ^ rol(c1, 1), d1 = c0 ^ rol(c2, 1), d2 = c1 ^ rol(c3, 1), d3 = c2 ^ rol(c4, 1),
                 d4 = c3 ^ rol(c0, 1);
        uint64_t b0 = a[0] ^ d0, b10 = rol(a[1] ^ d1, 1), b20 = rol(a[2] ^ d2, 62), b5 = rol(a[3] ^ d3, 28),
                 b15 = rol(a[4] ^ d4, 27), b16 = rol(a[5] ^ d0, 36), b1 = rol(a[6] ^ d1, 44), b11 = rol(a[7] ^ d2, 6),
                 b21 = rol(a[8] ^ d3, 55), b6 = rol(a[9] ^ d4, 20), b7 = rol(a[10] ^ d0, 3), b17 = rol(a[11] ^ d1, 10),
                 b2 = rol(a[12] ^ d2, 43), b12 = rol(a[13] ^ d3, 25), b22 = rol(a[14] ^ d4, 39),
                 b23 = rol(a[15] ^ d0, 41), b8 = rol(a[16] ^ d1, 45), b18 = rol(a[17] ^ d2, 15),
                 b3 = rol(a[18] ^ d3, 21), b13 = rol(a[19] ^ d4, 8), b14 = rol(a[20] ^ d0, 18),
                 b24 = rol(a[21] ^ d1, 2), b9 = rol(a[22] ^ d2, 61), b19 = rol(a[23] ^ d3, 56),
                 b4 = rol(a[24] ^ d4, 14);
        a[0] = b0 ^ (~b1 & b2) ^ rc[round]; a[1] = b1 ^ (~b2 & b3); a[2] = b2 ^ (~b3 & b4); a[3] = b3 ^ (~b4 & b0); a[4] = b4 ^ (~b0 & b1);
        a[5] = b5 ^ (~b6 & b7); a[6] = b6 ^ (~b7 & b8); a[7] = b7 ^ (~b8 & b9); a[8] = b8 ^ (~b9 & b5); a[9] = b9 ^ (~b5 & b6);
        a[10] = b10 ^ (~b11 & b12); a[11] = b11 ^ (~b12 & b13); a[12] = b12 ^ (~b13 & b14); a[13] = b13 ^ (~b14 & b10); a[14] = b14 ^ (~b10 & b11);
        a[15] = b15 ^ (~b16 & b17); a[16] = b16 ^ (~b17 & b18); a[17] = b17 ^ (~b18 & b19); a[18] = b18 ^ (~b19 & b15); a[19] = b19 ^ (~b15 & b16);
        a[20] = b20 ^ (~b21 & b22); a[21] = b21 ^ (~b22 & b23); a[22] = b22 ^ (~b23 & b24); a[23] = b23 ^ (~b24 & b20); a[24] = b24 ^ (~b20 & b21);
    }
}

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

} // namespace azk::host
