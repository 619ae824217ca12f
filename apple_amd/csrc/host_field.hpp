// Host M31 / CM31 / QM31 arithmetic and the canonical Circle domain in storage order:
// shared by the proof verifier and the CPU gates of the *_verified entry points.
// Canonical representatives in [0, p) throughout (M31Field, QM31Field of the reference,
// SumcheckOracle.swift:3-147, QM31Field.swift:38-89; CircleDomain.swift:31-60,273-331).
#pragma once
#include <stdint.h>

#include <vector>

namespace azk::host {

constexpr uint32_t kM31 = 0x7fffffffu;

inline uint32_t fadd(uint32_t a, uint32_t b) { uint32_t s = a + b; return s >= kM31 ? s - kM31 : s; }
inline uint32_t fsub(uint32_t a, uint32_t b) { return a >= b ? a - b : a + (kM31 - b); }
inline uint32_t fmul(uint32_t a, uint32_t b)
{
    const uint64_t t = (uint64_t)a * b;
    uint32_t r = ((uint32_t)t & kM31) + (uint32_t)(t >> 31);
    r = (r & kM31) + (r >> 31);
    return r >= kM31 ? r - kM31 : r;
}
inline uint32_t fpow(uint32_t b, uint32_t e)
{
    uint32_t r = 1;
    for (; e; e >>= 1, b = fmul(b, b))
        if (e & 1) r = fmul(r, b);
    return r;
}
inline uint32_t finv(uint32_t a) { return fpow(a, kM31 - 2); }

struct Q { uint32_t a, b, c, d; };
inline bool operator==(const Q &x, const Q &y) { return x.a == y.a && x.b == y.b && x.c == y.c && x.d == y.d; }
inline Q qadd(Q x, Q y) { return {fadd(x.a, y.a), fadd(x.b, y.b), fadd(x.c, y.c), fadd(x.d, y.d)}; }
inline Q qsub(Q x, Q y) { return {fsub(x.a, y.a), fsub(x.b, y.b), fsub(x.c, y.c), fsub(x.d, y.d)}; }
inline Q qscale(Q x, uint32_t s) { return {fmul(x.a, s), fmul(x.b, s), fmul(x.c, s), fmul(x.d, s)}; }
// QM31 = CM31[u]/(u^2 - (2 + i)): (A + Bu)(C + Du) = AC + (2 + i)BD + (AD + BC)u
inline Q qmul(Q x, Q y)
{
    auto cm = [](uint32_t x0, uint32_t x1, uint32_t y0, uint32_t y1, uint32_t &r0, uint32_t &r1) {
        r0 = fsub(fmul(x0, y0), fmul(x1, y1));
        r1 = fadd(fmul(x0, y1), fmul(x1, y0));
    };
    uint32_t ac0, ac1, bd0, bd1, ad0, ad1, bc0, bc1;
    cm(x.a, x.b, y.a, y.b, ac0, ac1);
    cm(x.c, x.d, y.c, y.d, bd0, bd1);
    cm(x.a, x.b, y.c, y.d, ad0, ad1);
    cm(x.c, x.d, y.a, y.b, bc0, bc1);
    const uint32_t n0 = fsub(fadd(bd0, bd0), bd1), n1 = fadd(bd0, fadd(bd1, bd1));
    return {fadd(ac0, n0), fadd(ac1, n1), fadd(ad0, bc0), fadd(ad1, bc1)};
}
inline bool canonical(const Q &x) { return x.a < kM31 && x.b < kM31 && x.c < kM31 && x.d < kM31; }

// ---------------------------------------------------------------- circle domain
struct Pt { uint32_t x, y; };
inline Pt padd(Pt p, Pt q) { return {fsub(fmul(p.x, q.x), fmul(p.y, q.y)), fadd(fmul(p.x, q.y), fmul(p.y, q.x))}; }
// generator^idx, generator (2, 1268011823) of order 2^31
inline Pt point_at(uint32_t idx)
{
    Pt r{1, 0}, g{2, 1268011823u};
    for (idx &= 0x7fffffffu; idx; idx >>= 1, g = padd(g, g))
        if (idx & 1) r = padd(r, g);
    return r;
}
inline uint32_t double_x(uint32_t x) { return fsub(fadd(fmul(x, x), fmul(x, x)), 1); }

struct Domain {
    uint32_t log_n = 0, init = 0, half_log = 0;
    // Canonical domain of 2^log_n points (log_n <= 30): half coset from index 2^(30 - n)
    // with step 2^(32 - n) (isCanonical: 4 * init = 2^(31 - halfLog)).
    static Domain canonical_of(uint32_t log_n)
    {
        Domain d;
        d.log_n = log_n;
        d.half_log = log_n - 1;
        d.init = 1u << (30 - log_n);
        return d;
    }
    // Circle point of storage index s (bit-reversed order): natural i < N/2 is
    // init + step * i, the upper half the conjugates.
    Pt storage_point(uint64_t s) const
    {
        uint64_t nat = 0, v = s;
        for (uint32_t b = 0; b < log_n; ++b, v >>= 1) nat = (nat << 1) | (v & 1);
        const uint64_t half = 1ull << half_log, step = 1ull << (31 - half_log);
        uint32_t idx;
        if (nat < half) idx = (uint32_t)((init + step * nat) & 0x7fffffffu);
        else idx = (uint32_t)((0x80000000ull - ((init + step * (nat - half)) & 0x7fffffffu)) & 0x7fffffffu);
        return point_at(idx);
    }
    // Every point in storage order, one group addition per point.
    std::vector<Pt> storage_points() const
    {
        const uint64_t n = 1ull << log_n, half = n >> 1;
        std::vector<Pt> lower(half), out(n);
        const Pt step = point_at((uint32_t)(1ull << (31 - half_log)));
        Pt cur = point_at(init);
        for (uint64_t i = 0; i < half; ++i, cur = padd(cur, step)) lower[i] = cur;
        for (uint64_t s = 0; s < n; ++s) {
            uint64_t nat = 0, v = s;
            for (uint32_t b = 0; b < log_n; ++b, v >>= 1) nat = (nat << 1) | (v & 1);
            out[s] = nat < half ? lower[nat] : Pt{lower[nat - half].x, lower[nat - half].y ? kM31 - lower[nat - half].y : 0};
        }
        return out;
    }
    // Round r's domain coordinate of pair p: y of the pair's left point for r = 0, x of
    // the left point doubled r - 1 times afterwards (CircleFRILayerOracleV1.inverseDomainLayers,
    // CircleDomain.swift:469-508); the fold multiplies by its inverse.
    uint32_t fold_coordinate(uint32_t r, uint64_t p) const
    {
        if (r == 0) return storage_point(2 * p).y;
        uint32_t x = storage_point(p << (r + 1)).x;
        for (uint32_t k = 1; k < r; ++k) x = double_x(x);
        return x;
    }
};

// In-place inverses of nonzero values (one inversion, three products each).
inline void batch_inverse(std::vector<uint32_t> &v)
{
    if (v.empty()) return;
    std::vector<uint32_t> pre(v.size());
    uint32_t acc = 1;
    for (size_t i = 0; i < v.size(); ++i) {
        pre[i] = acc;
        acc = fmul(acc, v[i]);
    }
    uint32_t inv = finv(acc);
    for (size_t i = v.size(); i-- > 0;) {
        const uint32_t x = v[i];
        v[i] = fmul(inv, pre[i]);
        inv = fmul(inv, x);
    }
}

} // namespace azk::host
