// Host-side SHA3 for the library's *_verified CPU gates (the reference's verified APIs
// recompute on the CPU and throw correctnessValidationFailed on mismatch, e.g.
// SHA3BatchHasher.swift:211-263).  Not used on any compute path.
#pragma once
#include <stdint.h>
#include <string.h>

namespace azk::host {

inline uint64_t rol(uint64_t v, int n) { return n ? (v << n) | (v >> (64 - n)) : v; }

inline void keccak_f(uint64_t a[25])
{
    static const uint64_t rc[24] = {
        0x0000000000000001ull, 0x0000000000008082ull, 0x800000000000808Aull, 0x8000000080008000ull,
        0x000000000000808Bull, 0x0000000080000001ull, 0x8000000080008081ull, 0x8000000000008009ull,
        0x000000000000008Aull, 0x0000000000000088ull, 0x0000000080008009ull, 0x000000008000000Aull,
        0x000000008000808Bull, 0x800000000000008Bull, 0x8000000000008089ull, 0x8000000000008003ull,
        0x8000000000008002ull, 0x8000000000000080ull, 0x000000000000800Aull, 0x800000008000000Aull,
        0x8000000080008081ull, 0x8000000000008080ull, 0x0000000080000001ull, 0x8000000080008008ull};
    static const int r[25] = {0, 1, 62, 28, 27, 36, 44, 6, 55, 20, 3, 10, 43, 25, 39, 41, 45, 15, 21, 8, 18, 2, 61, 56, 14};
    for (int round = 0; round < 24; ++round) {
        uint64_t c[5], b[25];
        for (int x = 0; x < 5; ++x) c[x] = a[x] ^ a[x + 5] ^ a[x + 10] ^ a[x + 15] ^ a[x + 20];
        for (int x = 0; x < 5; ++x) {
            uint64_t d = c[(x + 4) % 5] ^ rol(c[(x + 1) % 5], 1);
            for (int y = 0; y < 25; y += 5) a[x + y] ^= d;
        }
        for (int x = 0; x < 5; ++x)
            for (int y = 0; y < 5; ++y) b[y + 5 * ((2 * x + 3 * y) % 5)] = rol(a[x + 5 * y], r[x + 5 * y]);
        for (int y = 0; y < 25; y += 5)
            for (int x = 0; x < 5; ++x) a[x + y] = b[x + y] ^ (~b[(x + 1) % 5 + y] & b[(x + 2) % 5 + y]);
        a[0] ^= rc[round];
    }
}

// SHA3-256 (suffix 0x06) or Keccak-256 (suffix 0x01) of an arbitrary message.
inline void sha3_256(const uint8_t *p, size_t len, uint8_t out[32], uint8_t suffix = 0x06)
{
    uint64_t a[25] = {0};
    uint8_t blk[136];
    size_t off = 0;
    for (;;) {
        size_t take = len - off < 136 ? len - off : 136;
        memset(blk, 0, sizeof blk);
        memcpy(blk, p + off, take);
        bool last = take < 136;
        if (last) {
            blk[take] ^= suffix;
            blk[135] ^= 0x80;
        }
        for (int i = 0; i < 17; ++i) {
            uint64_t v;
            memcpy(&v, blk + 8 * i, 8);
            a[i] ^= v;
        }
        keccak_f(a);
        off += take;
        if (last) break;
    }
    memcpy(out, a, 32);
}

} // namespace azk::host
