// Keccak-f[1600] and fixed-shape SHA3-256 sponges for gfx950.
//
// Two forms of the same permutation:
//  * kf_v: one state per lane, each 64-bit lane as two VGPRs.  Rotations are
//    v_alignbit_b32 pairs, the 5-way column parities and chi are v_bitop3_b32
//    (truth tables 0x96 = xor3, 0xD2 = chi): 180 VALU ops per round, the middle
//    22 rounds on registers placed for double-rate issue.  The throughput form.
//  * kf_w: one state per wave, one state word per lane (plus copies); DPP and
//    permlane swaps for theta, one ds_bpermute exchange per round for pi and chi.
//    The latency form (2.54 us per dependent permutation, against 8.8 us for a
//    wave-uniform state on the scalar unit): Merkle tree tops and the Fiat-Shamir
//    transcripts of the commit and of the sum-check.
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

// kround_pinned: the same round as inline assembly over hand-placed VGPRs (state in
// v8..v57, rho/pi output and theta temporaries in v58..v107), written by
// gen_keccak_round.py so that no three-source op reads three VGPRs of one parity.
#include "keccak_round_asm.hpp"

// First and last rounds are compiled from kround_v so sparse inputs constant-fold and
// unused output lanes are dropped; the middle 22 rounds are kround_pinned, rolled two per
// trip so a kernel can hold several call sites without blowing the instruction cache.
__device__ __forceinline__ void kf_v(u2 (&a)[25])
{
    kround_v(a, 0x00000001u, 0u);
#pragma unroll 2
    for (int r = 1; r < 23; ++r) kround_pinned(a, kRcLo[r], kRcHi[r]);
    kround_v(a, 0x80008008u, 0x80000000u);
}

// ---------------------------------------------------------------- two lanes per state
// Lanes 2k and 2k + 1 hold one state: the even lane the low, the odd lane the high 32-bit
// halves of all 25 words.  A 64-bit rotation by r is one alignbit of this lane's half with
// the partner's (DPP quad_perm [1,0,3,2]), the same formula on both lanes: alignbit(mine,
// partner, 32 - r) for r < 32, alignbit(partner, mine, 64 - r) above.  Theta's parities,
// chi and iota stay per half: about 125 instructions per round per lane against 190 for
// one lane per state, for levels of a Merkle tree with 16..128 parents per block.  Both
// lanes of a pair must be active.
__device__ __forceinline__ uint32_t pswap(uint32_t v) { return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xf, 0xf, false); }

template <int N>
__device__ __forceinline__ uint32_t hrot(uint32_t v)
{
    if constexpr (N == 0) return v;
    else if constexpr (N < 32) return __builtin_amdgcn_alignbit(v, pswap(v), 32 - N);
    else return __builtin_amdgcn_alignbit(pswap(v), v, 64 - N);
}

__device__ __forceinline__ void kround_h(uint32_t (&a)[25], uint32_t rc)
{
    const uint32_t c0 = xor3(xor3(a[0], a[5], a[10]), a[15], a[20]);
    const uint32_t c1 = xor3(xor3(a[1], a[6], a[11]), a[16], a[21]);
    const uint32_t c2 = xor3(xor3(a[2], a[7], a[12]), a[17], a[22]);
    const uint32_t c3 = xor3(xor3(a[3], a[8], a[13]), a[18], a[23]);
    const uint32_t c4 = xor3(xor3(a[4], a[9], a[14]), a[19], a[24]);
    const uint32_t d0 = c4 ^ hrot<1>(c1), d1 = c0 ^ hrot<1>(c2), d2 = c1 ^ hrot<1>(c3);
    const uint32_t d3 = c2 ^ hrot<1>(c4), d4 = c3 ^ hrot<1>(c0);
    uint32_t b[25];  // B[y, 2x+3y] = rot(A[x, y] ^ D[x], r[x, y]), as in kround_v
    b[0] = a[0] ^ d0;
    b[10] = hrot<1>(a[1] ^ d1);
    b[20] = hrot<62>(a[2] ^ d2);
    b[5] = hrot<28>(a[3] ^ d3);
    b[15] = hrot<27>(a[4] ^ d4);
    b[16] = hrot<36>(a[5] ^ d0);
    b[1] = hrot<44>(a[6] ^ d1);
    b[11] = hrot<6>(a[7] ^ d2);
    b[21] = hrot<55>(a[8] ^ d3);
    b[6] = hrot<20>(a[9] ^ d4);
    b[7] = hrot<3>(a[10] ^ d0);
    b[17] = hrot<10>(a[11] ^ d1);
    b[2] = hrot<43>(a[12] ^ d2);
    b[12] = hrot<25>(a[13] ^ d3);
    b[22] = hrot<39>(a[14] ^ d4);
    b[23] = hrot<41>(a[15] ^ d0);
    b[8] = hrot<45>(a[16] ^ d1);
    b[18] = hrot<15>(a[17] ^ d2);
    b[3] = hrot<21>(a[18] ^ d3);
    b[13] = hrot<8>(a[19] ^ d4);
    b[14] = hrot<18>(a[20] ^ d0);
    b[24] = hrot<2>(a[21] ^ d1);
    b[9] = hrot<61>(a[22] ^ d2);
    b[19] = hrot<56>(a[23] ^ d3);
    b[4] = hrot<14>(a[24] ^ d4);
#pragma unroll
    for (int y = 0; y < 25; y += 5) {
        const uint32_t t0 = b[y], t1 = b[y + 1], t2 = b[y + 2], t3 = b[y + 3], t4 = b[y + 4];
        a[y] = t0 ^ (~t1 & t2);
        a[y + 1] = t1 ^ (~t2 & t3);
        a[y + 2] = t2 ^ (~t3 & t4);
        a[y + 3] = t3 ^ (~t4 & t0);
        a[y + 4] = t4 ^ (~t0 & t1);
    }
    a[0] ^= rc;
}

// `hi`: this lane holds the high halves (the odd lane of its pair).  Every round is
// unrolled, the round constants scalar operands: 5.36 us per dependent permutation on one
// wave against 5.83 us for a loop unrolled twice.
__device__ __forceinline__ void kf_h(uint32_t (&a)[25], bool hi)
{
#pragma unroll
    for (int r = 0; r < 24; ++r) {
        uint32_t rl = kRcLo[r], rh = kRcHi[r];
        asm volatile("" : "+s"(rl), "+s"(rh));
        kround_h(a, hi ? rh : rl);
    }
}

__device__ __forceinline__ uint32_t bperm(int addr, uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_ds_bpermute(addr, (int)v);
}

// ---------------------------------------------------------------- one state per wave
// Lane l = 8y + p holds A[x(p), y] for y < 5, where x(p) = 4, 0, 1, 2, 3, 4, 0, 1 for
// p = 0..7: positions 1..5 are the state, 0, 6 and 7 hold copies of columns 4, 0 and 1 so
// that every column neighbour a real lane needs sits one lane away inside its 8-lane
// group.  Lanes 40..63 hold zero.  Per round:
//  * theta: column parities by an xor all-reduce over the eight groups (DPP row_ror:8,
//    then v_permlane16_swap and v_permlane32_swap), C[x-1] and C[x+1] by DPP row_shr:1
//    and row_shl:1 -- no LDS traffic;
//  * rho at the source lane (per-lane half swap + alignbit);
//  * pi and chi's two row neighbours in one ds_bpermute exchange: every lane (copies
//    included) pulls B[x], B[x+1], B[x+2] of its row straight from their pi sources;
//  * the zero lanes are re-zeroed, iota goes to both lanes that hold A[0, 0].
// One LDS round trip per round: 2.54 us per permutation on one wave, against 5.2 us for a
// 25-lane layout that also pulls theta's columns with ds_bpermute (two round trips) and
// ~9 us for the one-lane form.
struct WaveLane {
    int s0, s1, s2;   // byte addresses of the pi sources of B[x, y], B[x+1, y], B[x+2, y]
    uint32_t sh;      // rho: optional half swap, then alignbit by sh
    uint32_t keep;    // ~0 on lanes 0..39, 0 on the zero lanes
    uint32_t lead;    // ~0 on the two lanes holding A[0, 0]
    bool sw;
    bool own;         // the lane is its word's position 1..5 (not a copy)
    uint32_t word;    // state word x + 5y held by this lane (25 on the zero lanes)
};

// Lane of the state's own copy of word w (position 1 + x of group y).
__device__ __forceinline__ uint32_t wave_lane_of(uint32_t w) { return 8u * (w / 5u) + 1u + w % 5u; }

__device__ __forceinline__ WaveLane wave_lane()
{
    constexpr uint32_t kSwap = 0x00c1d165u;  // bit w: rho swaps the halves of word w
    constexpr uint8_t kShift[25] = {0, 31, 2, 4, 5, 28, 20, 26, 9, 12, 29, 22, 21,
                                    7, 25, 23, 19, 17, 11, 24, 14, 30, 3, 8, 18};
    constexpr uint32_t kX = 0x10432104u;     // x(p) in nibble p
    const uint32_t lane = __lane_id(), y = lane >> 3, x = (kX >> (4 * (lane & 7u))) & 15u;
    WaveLane c;
    if (y >= 5) {
        c.s0 = c.s1 = c.s2 = (int)(lane * 4);
        c.sh = 0;
        c.keep = c.lead = 0;
        c.sw = false;
        c.own = false;
        c.word = 25;
        return c;
    }
    // B[X, y] = rho(A)[(X + 3y) mod 5, X]
    auto src = [&](uint32_t k) {
        const uint32_t X = (x + k) % 5u;
        return (int)(4 * (8 * X + 1 + (X + 3 * y) % 5u));
    };
    c.s0 = src(0);
    c.s1 = src(1);
    c.s2 = src(2);
    c.word = x + 5 * y;
    c.sh = kShift[c.word];
    c.sw = (kSwap >> c.word) & 1u;
    c.keep = ~0u;
    c.lead = c.word == 0 ? ~0u : 0u;
    c.own = (lane & 7u) - 1u < 5u;
    return c;
}

// Column parities of both halves over the eight groups.  Rows of 16 lanes hold groups
// (0, 1), (2, 3), (4, 5), (6, 7); row_ror:8 adds the two groups of a row, the first
// permlane16 swap pairs rows (0, 1) and (2, 3) of the low halves with those of the high
// halves, leaving [lo 0..3, hi 0..3, lo 4..7, hi 4..7] by row; the permlane32 swap adds
// rows 0 + 2 and 1 + 3, and the last permlane16 swap spreads lo and hi to every row.
__device__ __forceinline__ void wave_parity(uint32_t lo, uint32_t hi, uint32_t &cl, uint32_t &ch)
{
    const uint32_t ul = lo ^ __builtin_amdgcn_mov_dpp(lo, 0x128, 0xf, 0xf, false);  // row_ror:8
    const uint32_t uh = hi ^ __builtin_amdgcn_mov_dpp(hi, 0x128, 0xf, 0xf, false);
    const auto a = __builtin_amdgcn_permlane16_swap(ul, uh, false, false);
    const uint32_t x = a[0] ^ a[1];
    const auto b = __builtin_amdgcn_permlane32_swap(x, x, false, false);
    const uint32_t z = b[0] ^ b[1];
    const auto d = __builtin_amdgcn_permlane16_swap(z, z, false, false);
    cl = d[0];
    ch = d[1];
}

// Every round is unrolled: the round constants become scalar operands (no readlane on the
// chain), theta's rotation by one is taken before the neighbour shifts so that both DPP
// reads fold into the xors, and the permlane swaps get their two operands from two xors
// rather than an xor and a copy.  2.54 us per dependent permutation on one wave, against
// 2.89 us for the same round in a loop unrolled twice.
__device__ __forceinline__ uint32_t xor_ror8(uint32_t v)
{
    return v ^ (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x128, 0xf, 0xf, false);  // row_ror:8
}

__device__ __forceinline__ void kf_w(uint32_t &lo, uint32_t &hi, const WaveLane &c)
{
#pragma unroll
    for (int r = 0; r < 24; ++r) {
        // theta: column parities in every lane, then C[x - 1] and rot1(C[x + 1])
        const auto a = __builtin_amdgcn_permlane16_swap(xor_ror8(lo), xor_ror8(hi), false, false);
        const auto b = __builtin_amdgcn_permlane32_swap(a[0] ^ a[1], a[1] ^ a[0], false, false);
        const auto d = __builtin_amdgcn_permlane16_swap(b[0] ^ b[1], b[1] ^ b[0], false, false);
        const uint32_t cl = d[0], ch = d[1];
        const uint32_t rl = __builtin_amdgcn_alignbit(cl, ch, 31), rh = __builtin_amdgcn_alignbit(ch, cl, 31);
        lo ^= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)cl, 0x111, 0xf, 0xf, true);  // row_shr:1
        hi ^= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)ch, 0x111, 0xf, 0xf, true);
        lo ^= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)rl, 0x101, 0xf, 0xf, true);  // row_shl:1
        hi ^= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)rh, 0x101, 0xf, 0xf, true);
        // rho at the source lane, then pi and chi's row neighbours in one exchange
        const uint32_t h2 = c.sw ? lo : hi, l2 = c.sw ? hi : lo;
        const uint32_t th = __builtin_amdgcn_alignbit(h2, l2, c.sh), tl = __builtin_amdgcn_alignbit(l2, h2, c.sh);
        const uint32_t b0l = bperm(c.s0, tl), b1l = bperm(c.s1, tl), b2l = bperm(c.s2, tl);
        const uint32_t b0h = bperm(c.s0, th), b1h = bperm(c.s1, th), b2h = bperm(c.s2, th);
        // iota: the constants stay scalar operands (not per-round masks held in VGPRs)
        uint32_t rcl = kRcLo[r], rch = kRcHi[r];
        asm volatile("" : "+s"(rcl), "+s"(rch));
        lo = ((b0l ^ (~b1l & b2l)) & c.keep) ^ (rcl & c.lead);
        hi = ((b0h ^ (~b1h & b2h)) & c.keep) ^ (rch & c.lead);
    }
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
