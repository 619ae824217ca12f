// Keccak-f[1600] and fixed-shape SHA3-256 sponges for gfx950.
//
// Two forms of the same permutation:
//  * kf_v: one state per lane, each 64-bit lane as a VGPR pair.  Rotations are
//    v_alignbit_b32 pairs, the 5-way column parities and chi are v_bitop3_b32
//    (truth tables 0x96 = xor3, chi via the compiler's own bitop3 formation):
//    ~190 VALU ops per round, 4555 per permutation.
//  * kf_s: wave-uniform state; the compiler keeps it in SGPRs and runs 64-bit
//    SALU ops.  Used where one serial sponge sits on the critical path (the
//    Fiat-Shamir transcript): its latency is several times lower than one VALU wave.
// Semantics follow the reference's SHA3 oracle (Sources/AppleZKProver/SHA3Oracle.swift:
// 181-280) and its Metal keccak_f1600 (Resources/HashMerkleKernels.metal:308-348);
// round constants and rho offsets are the FIPS 202 tables.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace azk {

// ---------------------------------------------------------------- constants
__constant__ static const uint32_t kRcLo[24] = {
    0x00000001u, 0x00008082u, 0x0000808Au, 0x80008000u, 0x0000808Bu, 0x80000001u,
    0x80008081u, 0x00008009u, 0x0000008Au, 0x00000088u, 0x80008009u, 0x8000000Au,
    0x8000808Bu, 0x0000008Bu, 0x00008089u, 0x00008003u, 0x00008002u, 0x00000080u,
    0x0000800Au, 0x8000000Au, 0x80008081u, 0x00008080u, 0x80000001u, 0x80008008u};
__constant__ static const uint32_t kRcHi[24] = {
    0x00000000u, 0x00000000u, 0x80000000u, 0x80000000u, 0x00000000u, 0x00000000u,
    0x80000000u, 0x80000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u,
    0x00000000u, 0x80000000u, 0x80000000u, 0x80000000u, 0x80000000u, 0x80000000u,
    0x00000000u, 0x80000000u, 0x80000000u, 0x80000000u, 0x00000000u, 0x80000000u};

// ---------------------------------------------------------------- VALU form
struct u2 { uint32_t lo, hi; };

__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c)
{
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}
__device__ __forceinline__ u2 x2(u2 a, u2 b) { return {a.lo ^ b.lo, a.hi ^ b.hi}; }
__device__ __forceinline__ u2 x5(u2 a, u2 b, u2 c, u2 d, u2 e)
{
    return {xor3(xor3(a.lo, b.lo, c.lo), d.lo, e.lo), xor3(xor3(a.hi, b.hi, c.hi), d.hi, e.hi)};
}
// 64-bit rotate-left by a compile-time n, as alignbit pairs (n = 32 is a swap).
template <int N>
__device__ __forceinline__ u2 rot(u2 v)
{
    if constexpr (N == 0) return v;
    else if constexpr (N == 32) return {v.hi, v.lo};
    else if constexpr (N < 32)
        return {__builtin_amdgcn_alignbit(v.lo, v.hi, 32 - N), __builtin_amdgcn_alignbit(v.hi, v.lo, 32 - N)};
    else
        return {__builtin_amdgcn_alignbit(v.hi, v.lo, 64 - N), __builtin_amdgcn_alignbit(v.lo, v.hi, 64 - N)};
}
__device__ __forceinline__ u2 chi(u2 a, u2 b, u2 c) { return {a.lo ^ (~b.lo & c.lo), a.hi ^ (~b.hi & c.hi)}; }

__device__ __forceinline__ void kround_v(u2 (&a)[25], uint32_t rclo, uint32_t rchi)
{
    u2 c0 = x5(a[0], a[5], a[10], a[15], a[20]);
    u2 c1 = x5(a[1], a[6], a[11], a[16], a[21]);
    u2 c2 = x5(a[2], a[7], a[12], a[17], a[22]);
    u2 c3 = x5(a[3], a[8], a[13], a[18], a[23]);
    u2 c4 = x5(a[4], a[9], a[14], a[19], a[24]);
    u2 d0 = x2(c4, rot<1>(c1)), d1 = x2(c0, rot<1>(c2)), d2 = x2(c1, rot<1>(c3));
    u2 d3 = x2(c2, rot<1>(c4)), d4 = x2(c3, rot<1>(c0));
    // theta + rho + pi: B[y, 2x+3y] = rot(A[x, y] ^ D[x], r[x, y])
    u2 b[25];
    b[0] = x2(a[0], d0);
    b[10] = rot<1>(x2(a[1], d1));
    b[20] = rot<62>(x2(a[2], d2));
    b[5] = rot<28>(x2(a[3], d3));
    b[15] = rot<27>(x2(a[4], d4));
    b[16] = rot<36>(x2(a[5], d0));
    b[1] = rot<44>(x2(a[6], d1));
    b[11] = rot<6>(x2(a[7], d2));
    b[21] = rot<55>(x2(a[8], d3));
    b[6] = rot<20>(x2(a[9], d4));
    b[7] = rot<3>(x2(a[10], d0));
    b[17] = rot<10>(x2(a[11], d1));
    b[2] = rot<43>(x2(a[12], d2));
    b[12] = rot<25>(x2(a[13], d3));
    b[22] = rot<39>(x2(a[14], d4));
    b[23] = rot<41>(x2(a[15], d0));
    b[8] = rot<45>(x2(a[16], d1));
    b[18] = rot<15>(x2(a[17], d2));
    b[3] = rot<21>(x2(a[18], d3));
    b[13] = rot<8>(x2(a[19], d4));
    b[14] = rot<18>(x2(a[20], d0));
    b[24] = rot<2>(x2(a[21], d1));
    b[9] = rot<61>(x2(a[22], d2));
    b[19] = rot<56>(x2(a[23], d3));
    b[4] = rot<14>(x2(a[24], d4));
#pragma unroll
    for (int y = 0; y < 25; y += 5) {
        u2 t0 = b[y], t1 = b[y + 1], t2 = b[y + 2], t3 = b[y + 3], t4 = b[y + 4];
        a[y] = chi(t0, t1, t2);
        a[y + 1] = chi(t1, t2, t3);
        a[y + 2] = chi(t2, t3, t4);
        a[y + 3] = chi(t3, t4, t0);
        a[y + 4] = chi(t4, t0, t1);
    }
    a[0].lo ^= rclo;
    a[0].hi ^= rchi;
}

// First and last rounds are peeled so sparse inputs constant-fold and unused
// output lanes are dropped; the middle 22 rounds stay rolled (two per trip) so a
// kernel can hold several call sites without blowing the instruction cache.
__device__ __forceinline__ void kf_v(u2 (&a)[25])
{
    kround_v(a, 0x00000001u, 0u);
#pragma unroll 2
    for (int r = 1; r < 23; ++r) kround_v(a, kRcLo[r], kRcHi[r]);
    kround_v(a, 0x80008008u, 0x80000000u);
}

// ---------------------------------------------------------------- wavefront-cooperative form
// One state spread over 25 lanes, two states per wave (lanes 0..24 and 32..56); lane
// l = x + 5y holds A[x, y] as a (lo, hi) VGPR pair, which is also the state's word l
// in sponge order.  Cross-lane steps pull with ds_bpermute; rho is a per-lane swap +
// alignbit.  About 45 instructions per round instead of the one-lane form's 190, so one
// permutation finishes sooner when only a few states are in flight (a transcript step,
// the top levels of a Merkle tree).  Lanes 25..31 and
// 57..63 carry a copy of nothing useful and are never read.  Each round takes two
// exchanges: theta pulls columns x-1 and x+1 over all rows (20 pulls), and pi + chi
// pull a lane's B[x], B[x+1], B[x+2] straight from their pi sources (6 pulls).
struct CoopLane {
    int cm[5], cp[5];    // byte addresses (lane * 4) of A[x-1, k] and A[x+1, k], k = 0..4
    int pi0, pi1, pi2;   // pi sources of B[x, y], B[x+1, y], B[x+2, y]: A[(X + 3y) % 5, X]
    uint32_t sh;         // rho: rotl by r = optional half swap, then alignbit by sh
    bool sw, lead;       // lead: lane 0 of its state (iota)
    uint32_t word;       // l = x + 5y (25.. for idle lanes)
};

__device__ __forceinline__ CoopLane coop_lane()
{
    // rho offsets r[x + 5y] as (swap, shift) pairs: r >= 32 swaps halves, r' = r mod 32
    // is alignbit by 32 - r' (r' = 0 flips the swap and shifts by 0).
    constexpr uint32_t kSwap = 0x00c1d165u;  // bit l: swap for lane l
    constexpr uint8_t kShift[25] = {0, 31, 2, 4, 5, 28, 20, 26, 9, 12, 29, 22, 21,
                                    7, 25, 23, 19, 17, 11, 24, 14, 30, 3, 8, 18};
    const uint32_t lane = __lane_id(), base = lane & 32u, l = lane & 31u;
    CoopLane c;
    c.word = l;
    if (l >= 25) {
        const int self = (int)(lane * 4);
        for (int k = 0; k < 5; ++k) c.cm[k] = c.cp[k] = self;
        c.pi0 = c.pi1 = c.pi2 = self;
        c.sh = 0;
        c.sw = false;
        c.lead = false;
        return c;
    }
    const uint32_t x = l % 5, y = l / 5;
    auto at = [&](uint32_t xx, uint32_t yy) { return (int)((base + xx % 5 + 5 * (yy % 5)) * 4); };
    for (uint32_t k = 0; k < 5; ++k) {
        c.cm[k] = at(x + 4, k);
        c.cp[k] = at(x + 1, k);
    }
    c.pi0 = at(x + 3 * y, x);
    c.pi1 = at(x + 1 + 3 * y, x + 1);
    c.pi2 = at(x + 2 + 3 * y, x + 2);
    c.sh = kShift[l];
    c.sw = (kSwap >> l) & 1u;
    c.lead = l == 0;
    return c;
}

__device__ __forceinline__ uint32_t bperm(int addr, uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_ds_bpermute(addr, (int)v);
}

__device__ __forceinline__ void kf_coop(uint32_t &lo, uint32_t &hi, const CoopLane &c)
{
#pragma unroll 2
    for (int r = 0; r < 24; ++r) {
        // theta: columns x-1 and x+1 over all five rows in one exchange
        uint32_t ml[5], mh[5], pl[5], ph[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            ml[k] = bperm(c.cm[k], lo);
            mh[k] = bperm(c.cm[k], hi);
            pl[k] = bperm(c.cp[k], lo);
            ph[k] = bperm(c.cp[k], hi);
        }
        const uint32_t cml = xor3(xor3(ml[0], ml[1], ml[2]), ml[3], ml[4]);
        const uint32_t cmh = xor3(xor3(mh[0], mh[1], mh[2]), mh[3], mh[4]);
        const uint32_t cpl = xor3(xor3(pl[0], pl[1], pl[2]), pl[3], pl[4]);
        const uint32_t cph = xor3(xor3(ph[0], ph[1], ph[2]), ph[3], ph[4]);
        lo = xor3(lo, cml, __builtin_amdgcn_alignbit(cpl, cph, 31));
        hi = xor3(hi, cmh, __builtin_amdgcn_alignbit(cph, cpl, 31));
        // rho at the source lane; pi and chi's two row neighbours in one exchange
        const uint32_t h2 = c.sw ? lo : hi, l2 = c.sw ? hi : lo;
        const uint32_t th = __builtin_amdgcn_alignbit(h2, l2, c.sh), tl = __builtin_amdgcn_alignbit(l2, h2, c.sh);
        const uint32_t b0l = bperm(c.pi0, tl), b1l = bperm(c.pi1, tl), b2l = bperm(c.pi2, tl);
        const uint32_t b0h = bperm(c.pi0, th), b1h = bperm(c.pi1, th), b2h = bperm(c.pi2, th);
        lo = b0l ^ (~b1l & b2l);
        hi = b0h ^ (~b1h & b2h);
        if (c.lead) {  // iota
            lo ^= kRcLo[r];
            hi ^= kRcHi[r];
        }
    }
}

// ---------------------------------------------------------------- SALU form
__device__ __forceinline__ uint64_t rotl64(uint64_t x, int n) { return n == 0 ? x : ((x << n) | (x >> (64 - n))); }

__device__ __forceinline__ void kround_s(uint64_t (&a)[25], uint64_t rc)
{
    uint64_t c[5], d[5], b[25];
#pragma unroll
    for (int x = 0; x < 5; ++x) c[x] = a[x] ^ a[x + 5] ^ a[x + 10] ^ a[x + 15] ^ a[x + 20];
#pragma unroll
    for (int x = 0; x < 5; ++x) d[x] = c[(x + 4) % 5] ^ rotl64(c[(x + 1) % 5], 1);
    constexpr int R[25] = {0, 1, 62, 28, 27, 36, 44, 6, 55, 20, 3, 10, 43, 25, 39, 41, 45, 15, 21, 8, 18, 2, 61, 56, 14};
#pragma unroll
    for (int x = 0; x < 5; ++x)
#pragma unroll
        for (int y = 0; y < 5; ++y) b[y + 5 * ((2 * x + 3 * y) % 5)] = rotl64(a[x + 5 * y] ^ d[x], R[x + 5 * y]);
#pragma unroll
    for (int y = 0; y < 25; y += 5)
#pragma unroll
        for (int x = 0; x < 5; ++x) a[y + x] = b[y + x] ^ (~b[y + (x + 1) % 5] & b[y + (x + 2) % 5]);
    a[0] ^= rc;
}

__device__ __forceinline__ void kf_s(uint64_t (&a)[25])
{
#pragma unroll 1
    for (int r = 0; r < 24; ++r) kround_s(a, (uint64_t)kRcLo[r] | ((uint64_t)kRcHi[r] << 32));
}

// Wave-uniform 64-bit load: the value is moved to SGPRs so the sponge runs on the
// scalar unit (measured 8.8 us per permutation on one wave vs 12.3 us on VALU).
__device__ __forceinline__ uint64_t ld_uniform(const uint64_t *p)
{
    uint64_t v = *p;
    uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v), hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return (uint64_t)lo | ((uint64_t)hi << 32);
}

// ---------------------------------------------------------------- sponges
__device__ __forceinline__ void zero_state(u2 (&a)[25])
{
#pragma unroll
    for (int i = 0; i < 25; ++i) a[i] = {0u, 0u};
}

// SHA3-256 of a 16-byte QM31 leaf (LE u32 a, b, c, d): one permutation.
// Reference: sha3_256_absorb_qm31_leaf, Resources/HashMerkleKernels.metal:733-740.
__device__ __forceinline__ void sha3_leaf16(const uint32_t v[4], uint32_t out[8])
{
    u2 a[25];
    zero_state(a);
    a[0] = {v[0], v[1]};
    a[1] = {v[2], v[3]};
    a[2] = {0x06u, 0u};
    a[16] = {0u, 0x80000000u};
    kf_v(a);
#pragma unroll
    for (int i = 0; i < 4; ++i) { out[2 * i] = a[i].lo; out[2 * i + 1] = a[i].hi; }
}

// SHA3-256(left || right), two 32-byte digests: one permutation (Merkle parent).
// Reference: MerkleOracle.swift:16-26, HashMerkleKernels.metal:1082-1098.
__device__ __forceinline__ void sha3_parent(const uint32_t l[8], const uint32_t r[8], uint32_t out[8])
{
    u2 a[25];
    zero_state(a);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        a[i] = {l[2 * i], l[2 * i + 1]};
        a[4 + i] = {r[2 * i], r[2 * i + 1]};
    }
    a[8] = {0x06u, 0u};
    a[16] = {0u, 0x80000000u};
    kf_v(a);
#pragma unroll
    for (int i = 0; i < 4; ++i) { out[2 * i] = a[i].lo; out[2 * i + 1] = a[i].hi; }
}

// One-block sponge over 34 message words (already zeroed beyond `len` bytes),
// len <= 136; a 136-byte message takes a second, padding-only block.
// Reference: SHA3Oracle.swift:181-219 with the one-block bound of
// SHA3BatchHasher.swift:161-166; Metal absorb_fixed_136 HashMerkleKernels.metal:633-657.
__device__ __forceinline__ void sponge_oneblock(const uint32_t (&w)[34], uint32_t len, uint32_t suffix, uint32_t out[8])
{
    u2 a[25];
    zero_state(a);
#pragma unroll
    for (int i = 0; i < 17; ++i) a[i] = {w[2 * i], w[2 * i + 1]};
    if (len == 136) {
        kf_v(a);
        a[0].lo ^= suffix;
    } else {
        // byte len ^= suffix
        uint32_t word = len >> 2, sh = (len & 3u) * 8u;
#pragma unroll
        for (int i = 0; i < 17; ++i) {
            if (word == (uint32_t)(2 * i)) a[i].lo ^= suffix << sh;
            if (word == (uint32_t)(2 * i + 1)) a[i].hi ^= suffix << sh;
        }
    }
    a[16].hi ^= 0x80000000u;
    kf_v(a);
#pragma unroll
    for (int i = 0; i < 4; ++i) { out[2 * i] = a[i].lo; out[2 * i + 1] = a[i].hi; }
}

} // namespace azk
