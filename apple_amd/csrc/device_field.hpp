// Exact Mersenne-31 arithmetic and its extensions for gfx950.
//
// Canonical representatives in [0, p) everywhere, p = 2^31 - 1.  The values agree
// with the reference's M31Field / CM31Field / QM31Field oracles
// (Sources/AppleZKProver/SumcheckOracle.swift:3-147, CM31Field.swift:41-96,
// QM31Field.swift:38-89); QM31 = CM31[u]/(u^2 - (2 + i)).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace azk {

constexpr uint32_t M31_P = 0x7fffffffu;

__device__ __forceinline__ uint32_t m31_add(uint32_t a, uint32_t b)
{
    uint32_t s = a + b;
    return s >= M31_P ? s - M31_P : s;
}
__device__ __forceinline__ uint32_t m31_sub(uint32_t a, uint32_t b)
{
    return a >= b ? a - b : a + (M31_P - b);
}
__device__ __forceinline__ uint32_t m31_neg(uint32_t a) { return a ? M31_P - a : 0; }

// For a, b <= p the product t <= p^2 < 2^62, so one fold at bit 31 gives
// r = (t mod 2^31) + (t >> 31) <= (2^31 - 1) + (2^31 - 2) = 2p - 1, and one conditional
// subtract (v_add + v_min_u32) makes it canonical: 6 VALU ops per product.  The 64-bit
// product is one v_mad_u64_u32 (measured 15% more products per second on gfx950 than
// the v_mul_lo_u32 / v_mul_hi_u32 pair).
__device__ __forceinline__ uint32_t m31_mul(uint32_t a, uint32_t b)
{
    const uint64_t t = (uint64_t)a * b;
    const uint32_t r = ((uint32_t)t & M31_P) + (uint32_t)(t >> 31); // <= 2p - 1
    return r >= M31_P ? r - M31_P : r;
}

// t < 2^63 (a sum of two products of canonical values) -> t mod p.
__device__ __forceinline__ uint32_t m31_reduce63(uint64_t t)
{
    const uint64_t s = (t & M31_P) + (t >> 31);              // < 2^31 + 2^32
    uint32_t r = ((uint32_t)s & M31_P) + (uint32_t)(s >> 31); // < 2^31 + 3
    return r >= M31_P ? r - M31_P : r;
}

// (x0 + x1 i)(w0 + w1 i) with w1n = p - w1: real x0 w0 + x1 w1n, imaginary x0 w1 + x1 w0,
// each one 64-bit sum of two products and one reduction.
__device__ __forceinline__ void cm_mul_lazy(uint32_t x0, uint32_t x1, uint32_t w0, uint32_t w1, uint32_t w1n,
                                            uint32_t &r0, uint32_t &r1)
{
    r0 = m31_reduce63((uint64_t)x0 * w0 + (uint64_t)x1 * w1n);
    r1 = m31_reduce63((uint64_t)x0 * w1 + (uint64_t)x1 * w0);
}

__device__ __forceinline__ uint32_t m31_pow(uint32_t b, uint32_t e)
{
    uint32_t r = 1;
    while (e) {
        if (e & 1) r = m31_mul(r, b);
        b = m31_mul(b, b);
        e >>= 1;
    }
    return r;
}
__device__ __forceinline__ uint32_t m31_inv(uint32_t a) { return m31_pow(a, M31_P - 2); }

struct qm31 { uint32_t a, b, c, d; };

__device__ __forceinline__ qm31 qm_add(qm31 x, qm31 y)
{
    return {m31_add(x.a, y.a), m31_add(x.b, y.b), m31_add(x.c, y.c), m31_add(x.d, y.d)};
}
__device__ __forceinline__ qm31 qm_sub(qm31 x, qm31 y)
{
    return {m31_sub(x.a, y.a), m31_sub(x.b, y.b), m31_sub(x.c, y.c), m31_sub(x.d, y.d)};
}
__device__ __forceinline__ qm31 qm_scale(qm31 x, uint32_t s)
{
    return {m31_mul(x.a, s), m31_mul(x.b, s), m31_mul(x.c, s), m31_mul(x.d, s)};
}
// (x0 + x1 i) (y0 + y1 i)
__device__ __forceinline__ void cm_mul(uint32_t x0, uint32_t x1, uint32_t y0, uint32_t y1, uint32_t &r0, uint32_t &r1)
{
    r0 = m31_sub(m31_mul(x0, y0), m31_mul(x1, y1));
    r1 = m31_add(m31_mul(x0, y1), m31_mul(x1, y0));
}
// (A + B u)(C + D u) = AC + (2 + i) BD + (AD + BC) u
__device__ __forceinline__ qm31 qm_mul(qm31 x, qm31 y)
{
    uint32_t ac0, ac1, bd0, bd1, ad0, ad1, bc0, bc1;
    cm_mul(x.a, x.b, y.a, y.b, ac0, ac1);
    cm_mul(x.c, x.d, y.c, y.d, bd0, bd1);
    cm_mul(x.a, x.b, y.c, y.d, ad0, ad1);
    cm_mul(x.c, x.d, y.a, y.b, bc0, bc1);
    // (2 + i)(bd0 + bd1 i) = (2 bd0 - bd1) + (bd0 + 2 bd1) i
    uint32_t n0 = m31_sub(m31_add(bd0, bd0), bd1);
    uint32_t n1 = m31_add(bd0, m31_add(bd1, bd1));
    return {m31_add(ac0, n0), m31_add(ac1, n1), m31_add(ad0, bc0), m31_add(ad1, bc1)};
}

constexpr uint32_t M31_INV2 = 1u << 30;

// x * 2^-1 for canonical x: x / 2 when x is even, (x + p) / 2 = (x >> 1) + 2^30 when odd;
// the same canonical value as m31_mul(x, M31_INV2) in three ops instead of ten.
__device__ __forceinline__ uint32_t m31_half(uint32_t x) { return (x >> 1) | ((x & 1u) << 30); }

// FRI fold of one adjacent pair with an M31 twiddle t embedded as (t, 0, 0, 0):
// (l + r + ch * ((l - r) * t)) * 2^-1.   QM31FRIFold.swift:11-44
__device__ __forceinline__ qm31 fri_fold_pair(qm31 l, qm31 r, uint32_t t, qm31 ch)
{
    qm31 m = qm_add(qm_add(l, r), qm_mul(ch, qm_scale(qm_sub(l, r), t)));
    return {m31_half(m.a), m31_half(m.b), m31_half(m.c), m31_half(m.d)};
}

// Circle group: (x1, y1) * (x2, y2) = (x1 x2 - y1 y2, x1 y2 + y1 x2).  CircleDomain.swift:31-40
struct cpoint { uint32_t x, y; };
__device__ __forceinline__ cpoint cp_add(cpoint p, cpoint q)
{
    return {m31_sub(m31_mul(p.x, q.x), m31_mul(p.y, q.y)), m31_add(m31_mul(p.x, q.y), m31_mul(p.y, q.x))};
}
// generator^index by double-and-add, generator (2, 1268011823) of order 2^31.
__device__ __forceinline__ cpoint cp_from_index(uint32_t idx)
{
    cpoint r = {1u, 0u}, p = {2u, 1268011823u};
    idx &= 0x7fffffffu;
    for (int k = 0; k < 31; ++k) {
        if (idx & 1u) r = cp_add(r, p);
        p = cp_add(p, p);
        idx >>= 1;
    }
    return r;
}
__device__ __forceinline__ uint32_t double_x(uint32_t x)
{
    uint32_t sq = m31_mul(x, x);
    return m31_sub(m31_add(sq, sq), 1u);
}

} // namespace azk
