// CPU gates: host computations of the path's results for the *_verified entry points,
// and their C-ABI exports (the reference exposes its CPU oracles as public API too:
// CircleCodewordOracle, CircleFRILayerOracleV1).  Host code only; the device path never
// calls into this file except to compare after its own result is complete.
#include <algorithm>
#include <cstring>
#include <thread>
#include <vector>

#include "circle_frames.hpp"
#include "host_field.hpp"
#include "host_gates.hpp"
#include "host_sha3.hpp"
#include "prover_internal.hpp"
#include "runtime_common.hpp"

using namespace azk;
using namespace azk::host;

namespace azk::host {

namespace {

std::vector<Q> load_q(const uint32_t *w, size_t n)
{
    std::vector<Q> v(n);
    for (size_t i = 0; i < n; ++i) v[i] = {w[4 * i], w[4 * i + 1], w[4 * i + 2], w[4 * i + 3]};
    return v;
}
void store_q(const std::vector<Q> &v, uint32_t *w)
{
    for (size_t i = 0; i < v.size(); ++i) {
        w[4 * i] = v[i].a;
        w[4 * i + 1] = v[i].b;
        w[4 * i + 2] = v[i].c;
        w[4 * i + 3] = v[i].d;
    }
}
bool all_canonical(const uint32_t *w, size_t words)
{
    for (size_t i = 0; i < words; ++i)
        if (w[i] >= kM31) return false;
    return true;
}

// Stage s twiddle h: y (s = 0) or x doubled s - 1 times (s >= 1) of the point at storage
// index h << (s + 1) (CircleCodewordOracle.circleFFTTwiddles, CircleCodeword.swift:129-164).
uint32_t fft_twiddle(const std::vector<Pt> &pts, uint32_t s, uint64_t h)
{
    const Pt p = pts[h << (s + 1)];
    if (s == 0) return p.y;
    uint32_t x = p.x;
    for (uint32_t k = 1; k < s; ++k) x = double_x(x);
    return x;
}

// pi(t) = 2 t^2 - 1 = T_2(t)
Q q_pi(Q t)
{
    const Q sq = qmul(t, t);
    return qsub(qadd(sq, sq), Q{1, 0, 0, 0});
}

} // namespace

void circle_fft(const uint32_t *coeffs, uint32_t log_n, uint32_t *out)
{
    const uint64_t n = 1ull << log_n;
    std::vector<Q> v = load_q(coeffs, n);
    const std::vector<Pt> pts = Domain::canonical_of(log_n).storage_points();
    for (int sk = (int)log_n - 1; sk >= 0; --sk) {  // stages log n - 1 .. 1, then 0
        const uint32_t s = (uint32_t)sk;
        const uint64_t cnt = n >> (s + 1), lanes = 1ull << s;
        for (uint64_t h = 0; h < cnt; ++h) {
            const uint32_t t = fft_twiddle(pts, s, h);
            for (uint64_t lane = 0; lane < lanes; ++lane) {
                const uint64_t l = (h << (s + 1)) + lane, r = l + lanes;
                const Q sr = qscale(v[r], t), left = v[l];
                v[l] = qadd(left, sr);
                v[r] = qsub(left, sr);
            }
        }
    }
    store_q(v, out);
}

bool fold_round(const std::vector<Pt> &pts, uint32_t r, const uint32_t *in, uint64_t n_out, const uint32_t ch4[4],
                uint32_t *out)
{
    std::vector<uint32_t> t(n_out);
    for (uint64_t p = 0; p < n_out; ++p) {  // the pair's domain coordinate, then its inverse
        if (r == 0) {
            t[p] = pts[2 * p].y;
        } else {
            uint32_t x = pts[p << (r + 1)].x;
            for (uint32_t k = 1; k < r; ++k) x = double_x(x);
            t[p] = x;
        }
    }
    for (uint32_t c : t)
        if (!c) return false;  // inverseDomainLayers refuses a zero coordinate
    batch_inverse(t);
    const Q ch{ch4[0], ch4[1], ch4[2], ch4[3]}, inv2{1u << 30, 0, 0, 0};
    for (uint64_t p = 0; p < n_out; ++p) {  // (l + r + ch ((l - r) t)) / 2, QM31FRIFold.swift:11-44
        const Q l{in[8 * p], in[8 * p + 1], in[8 * p + 2], in[8 * p + 3]};
        const Q rr{in[8 * p + 4], in[8 * p + 5], in[8 * p + 6], in[8 * p + 7]};
        const Q o = qmul(qadd(qadd(l, rr), qmul(ch, qscale(qsub(l, rr), t[p]))), inv2);
        out[4 * p] = o.a;
        out[4 * p + 1] = o.b;
        out[4 * p + 2] = o.c;
        out[4 * p + 3] = o.d;
    }
    return true;
}

bool fold_chain(const uint32_t *evals, uint32_t log_n, const uint32_t *challenges, uint32_t rounds, uint32_t *out)
{
    const std::vector<Pt> pts = Domain::canonical_of(log_n).storage_points();
    std::vector<uint32_t> cur(evals, evals + ((size_t)4 << log_n)), next;
    for (uint32_t r = 0; r < rounds; ++r) {
        const uint64_t n_out = cur.size() / 8;
        next.assign(4 * n_out, 0);
        if (!fold_round(pts, r, cur.data(), n_out, challenges + 4 * r, next.data())) return false;
        cur.swap(next);
    }
    memcpy(out, cur.data(), 4 * cur.size());
    return true;
}

namespace {

// body(lo, hi) over [0, n) on up to 16 host threads, ranges of at least 4096 items.
template <class F>
void parallel_for(uint64_t n, F &&body)
{
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const unsigned nt = (unsigned)std::min<uint64_t>(hw, std::max<uint64_t>(1, n / 4096));
    std::vector<std::thread> pool;
    for (unsigned t = 1; t < nt; ++t) pool.emplace_back([&, t] { body(n * t / nt, n * (t + 1) / nt); });
    body(0, n / nt);
    for (auto &th : pool) th.join();
}

// Every node of the SHA3-256 Merkle tree over `count` 16-byte QM31 leaves, level by level
// from the leaf digests up (level l starts at node 2 count - 2 count >> l).
std::vector<uint8_t> qm31_merkle_tree(const uint32_t *leaves, uint64_t count)
{
    std::vector<uint8_t> nodes(32 * (2 * count - 1));
    parallel_for(count, [&](uint64_t lo, uint64_t hi) {
        for (uint64_t i = lo; i < hi; ++i) sha3_256((const uint8_t *)(leaves + 4 * i), 16, nodes.data() + 32 * i);
    });
    uint64_t at = 0;
    for (uint64_t m = count; m > 1; m >>= 1) {
        const uint8_t *lv = nodes.data() + 32 * at;
        uint8_t *up = nodes.data() + 32 * (at + m);
        parallel_for(m / 2, [&](uint64_t lo, uint64_t hi) {
            for (uint64_t i = lo; i < hi; ++i) sha3_256(lv + 64 * i, 64, up + 32 * i);
        });
        at += m;
    }
    return nodes;
}

void put32(std::vector<uint8_t> &v, uint32_t x) { for (int k = 0; k < 4; ++k) v.push_back((uint8_t)(x >> (8 * k))); }
void put64(std::vector<uint8_t> &v, uint64_t x) { for (int k = 0; k < 8; ++k) v.push_back((uint8_t)(x >> (8 * k))); }

} // namespace

bool circle_prove(const uint32_t *coeffs, const azk_circle_prover_desc &d, std::vector<uint8_t> &proof)
{
    const uint32_t log_n = d.log_n, R = d.rounds;
    const uint64_t n = 1ull << log_n;
    std::vector<std::vector<uint32_t>> layer(R + 1);
    layer[0].resize(4 * n);
    circle_fft(coeffs, log_n, layer[0].data());
    const std::vector<Pt> pts = Domain::canonical_of(log_n).storage_points();
    std::vector<std::vector<uint8_t>> tree(R);
    std::vector<uint32_t> ch(4 * R);
    Transcript t;
    for (const auto &m : frames::prefix_messages(log_n, d.log_blowup, d.query_count, d.folding_step, d.grinding_bits,
                                                 d.public_input_digest))
        t.absorb(m.data(), m.size());
    for (uint32_t r = 0; r < R; ++r) {  // commit, absorb the root, squeeze, fold (:1189-1260)
        const uint64_t m = n >> r;
        tree[r] = qm31_merkle_tree(layer[r].data(), m);
        const std::vector<uint8_t> cf = frames::commitment_frame(r), xf = frames::challenge_frame(r);
        t.absorb(cf.data(), cf.size());
        t.absorb(tree[r].data() + 32 * (2 * m - 2), 32);
        t.absorb(xf.data(), xf.size());
        t.squeeze_u32(4, kM31, ch.data() + 4 * r);
        layer[r + 1].resize(2 * m);
        if (!fold_round(pts, r, layer[r].data(), m / 2, ch.data() + 4 * r, layer[r + 1].data())) return false;
    }
    const uint64_t nf = n >> R;
    std::vector<uint8_t> f = frames::base_frame(5);  // final layer frame and bytes (:852-859)
    put64(f, nf);
    put64(f, 16 * nf);
    t.absorb(f.data(), f.size());
    t.absorb(reinterpret_cast<const uint8_t *>(layer[R].data()), 16 * nf);
    uint64_t nonce = 0;
    if (d.grinding_bits) {  // findGrindingNonce, counting up from 0 (:753-794)
        f = frames::base_frame(7);
        put32(f, d.grinding_bits);
        put32(f, 8);
        t.absorb(f.data(), f.size());
        for (;; ++nonce) {
            Transcript c = t;
            c.absorb(reinterpret_cast<const uint8_t *>(&nonce), 8);
            if (azk_circle_grinding_digest_meets_target(reinterpret_cast<const uint8_t *>(c.s), 32, d.grinding_bits)) break;
        }
        t.absorb(reinterpret_cast<const uint8_t *>(&nonce), 8);
    }
    f = frames::base_frame(6);  // query frame, then the pair indices (:708-717)
    put32(f, d.query_count);
    put64(f, n / 2);
    t.absorb(f.data(), f.size());
    std::vector<uint32_t> pairs(d.query_count);
    t.squeeze_u32(d.query_count, (uint32_t)(n / 2), pairs.data());
    // CirclePCSFRIProofCodecV1.encode (:475-535)
    proof.clear();
    const uint8_t magic[8] = {0x41, 0x5a, 0x4b, 0x50, 0x46, 0x56, 0x31, 0x00};
    proof.insert(proof.end(), magic, magic + 8);
    put32(proof, 1);
    put32(proof, 1);
    const std::vector<uint8_t> desc = frames::descriptor(log_n);
    put32(proof, (uint32_t)desc.size());
    proof.insert(proof.end(), desc.begin(), desc.end());
    for (uint32_t w : {d.log_blowup, d.query_count, d.folding_step, d.grinding_bits}) put32(proof, w);
    put32(proof, 32);
    proof.insert(proof.end(), d.public_input_digest, d.public_input_digest + 32);
    put32(proof, R);
    for (uint32_t r = 0; r < R; ++r) {
        const uint64_t m = n >> r;
        proof.insert(proof.end(), tree[r].begin() + 32 * (2 * m - 2), tree[r].begin() + 32 * (2 * m - 1));
    }
    put64(proof, nf);
    for (uint32_t w : layer[R]) put32(proof, w);
    put32(proof, d.query_count);
    auto opening = [&](uint32_t r, uint64_t leaf) {  // leaf index, value, siblings bottom-up
        const uint64_t m = n >> r;
        const uint32_t depth = log_n - r;
        put64(proof, leaf);
        for (int k = 0; k < 4; ++k) put32(proof, layer[r][4 * leaf + k]);
        put32(proof, depth);
        uint64_t at = 0, idx = leaf;
        for (uint64_t lv = m; lv > 1; lv >>= 1, idx >>= 1) {
            const uint8_t *sib = tree[r].data() + 32 * (at + (idx ^ 1));
            proof.insert(proof.end(), sib, sib + 32);
            at += lv;
        }
    };
    for (uint32_t q = 0; q < d.query_count; ++q) {  // makeQuery (:1278-1331)
        uint64_t pair = pairs[q];
        put64(proof, pair);
        put32(proof, R);
        for (uint32_t r = 0; r < R; ++r, pair >>= 1) {
            put32(proof, r);
            put64(proof, pair);
            opening(r, 2 * pair);
            opening(r, 2 * pair + 1);
        }
    }
    if (d.grinding_bits) {  // no claimed openings, then the nonce (:516-525)
        put32(proof, 0);
        put64(proof, nonce);
    }
    return true;
}

void qm31_merkle_root(const uint32_t *leaves, uint64_t count, uint8_t root[32])
{
    auto parallel = [](uint64_t n, auto &&body) { parallel_for(n, body); };
    std::vector<uint8_t> level(32 * count), up;
    parallel(count, [&](uint64_t lo, uint64_t hi) {
        for (uint64_t i = lo; i < hi; ++i) sha3_256((const uint8_t *)(leaves + 4 * i), 16, level.data() + 32 * i);
    });
    for (uint64_t m = count; m > 1; m >>= 1) {
        up.assign(16 * m, 0);
        parallel(m / 2, [&](uint64_t lo, uint64_t hi) {
            for (uint64_t i = lo; i < hi; ++i) sha3_256(level.data() + 64 * i, 64, up.data() + 32 * i);
        });
        level.swap(up);
    }
    memcpy(root, level.data(), 32);
}

bool basis_consistent(const uint32_t *x, uint32_t nx, const uint32_t *y, uint32_t ny, uint32_t log_n,
                      const uint32_t *basis, uint32_t points)
{
    const uint64_t n = 1ull << log_n, half = n >> 1;
    // points from a sponge over the shape and samples of the inputs and the basis
    Transcript t;
    uint32_t hdr[3] = {log_n, nx, ny};
    t.absorb((const uint8_t *)hdr, sizeof hdr);
    const size_t sample = std::min<uint64_t>(n, 64);
    t.absorb((const uint8_t *)basis, 16 * sample);
    t.absorb((const uint8_t *)(basis + 4 * (n - sample)), 16 * sample);
    if (nx) t.absorb((const uint8_t *)(x + 4 * (size_t)(nx - 1)), 16);
    if (ny) t.absorb((const uint8_t *)(y + 4 * (size_t)(ny - 1)), 16);
    std::vector<uint32_t> w(8 * points);
    t.squeeze_u32((uint32_t)w.size(), kM31, w.data());
    std::vector<Q> v(half);
    for (uint32_t f = 0; f < 2; ++f) {
        const uint32_t *mono = f ? y : x;
        const uint32_t cnt = f ? ny : nx;
        for (uint32_t k = 0; k < points; ++k) {
            const uint32_t *pw = w.data() + 8 * k + 4 * f;
            const Q z{pw[0], pw[1], pw[2], pw[3]};
            Q lhs{0, 0, 0, 0};  // Horner over the monomial column
            for (uint32_t i = cnt; i-- > 0;)
                lhs = qadd(qmul(lhs, z), Q{mono[4 * i], mono[4 * i + 1], mono[4 * i + 2], mono[4 * i + 3]});
            // sum_i c_i prod_k T_{2^k}(z)^(bit k of i): fold the index bits from the lowest
            for (uint64_t i = 0; i < half; ++i) {
                const uint32_t *c = basis + 4 * (2 * i + f);
                v[i] = {c[0], c[1], c[2], c[3]};
            }
            Q tk = z;
            for (uint64_t m = half; m > 1; m >>= 1, tk = q_pi(tk))
                for (uint64_t j = 0; j < m / 2; ++j) v[j] = qadd(v[2 * j], qmul(tk, v[2 * j + 1]));
            if (!(lhs == v[0])) return false;
        }
    }
    return true;
}

} // namespace azk::host

extern "C" {

azk_status azk_host_circle_prove(const uint32_t *coeffs, const azk_circle_prover_desc *desc, uint8_t *proof,
                                 size_t capacity, size_t *len)
{
    CHECK_ARG(coeffs && desc && proof && len);
    azk_status s = azk::prv::validate_desc(*desc);
    if (s) return s;
    if (desc->log_n > 24) return rt::fail(AZK_ERR_INVALID_LAYOUT, "the host proof builder stops at 2^24");
    if (!all_canonical(coeffs, (size_t)4 << desc->log_n)) return rt::fail(AZK_ERR_INVALID_LAYOUT, "non-canonical coefficient");
    std::vector<uint8_t> out;
    try {
        if (!host::circle_prove(coeffs, *desc, out)) return rt::fail(AZK_ERR_INVALID_LAYOUT, "zero fold coordinate");
    } catch (const std::bad_alloc &) {
        return rt::fail(AZK_ERR_ALLOCATION, "host proof builder: out of host memory");
    }
    if (capacity < out.size()) return rt::fail(AZK_ERR_INVALID_LAYOUT, "proof capacity %zu < %zu", capacity, out.size());
    memcpy(proof, out.data(), out.size());
    *len = out.size();
    return AZK_OK;
}

azk_status azk_host_circle_fft(const uint32_t *coeffs, uint32_t log_n, uint32_t *out)
{
    CHECK_ARG(coeffs && out);
    if (log_n < 1 || log_n > 26) return rt::fail(AZK_ERR_INVALID_LAYOUT, "domain log size %u outside 1..26", log_n);
    if (!all_canonical(coeffs, (size_t)4 << log_n)) return rt::fail(AZK_ERR_INVALID_LAYOUT, "non-canonical coefficient");
    circle_fft(coeffs, log_n, out);
    return AZK_OK;
}

azk_status azk_host_fold_chain(const uint32_t *evals, uint32_t log_n, const uint32_t *challenges, uint32_t rounds,
                               uint32_t *out)
{
    CHECK_ARG(evals && challenges && out);
    if (log_n < 1 || log_n > 26 || rounds < 1 || rounds > log_n)
        return rt::fail(AZK_ERR_INVALID_LAYOUT, "fold chain of %u rounds on 2^%u", rounds, log_n);
    if (!all_canonical(evals, (size_t)4 << log_n) || !all_canonical(challenges, (size_t)4 * rounds))
        return rt::fail(AZK_ERR_INVALID_LAYOUT, "non-canonical evaluation or challenge");
    if (!fold_chain(evals, log_n, challenges, rounds, out)) return rt::fail(AZK_ERR_INVALID_LAYOUT, "zero domain coordinate");
    return AZK_OK;
}

azk_status azk_witness_basis_check(const uint32_t *x, uint32_t nx, const uint32_t *y, uint32_t ny, uint32_t log_n,
                                   const uint32_t *basis, int *consistent)
{
    CHECK_ARG(basis && consistent && (x || !nx) && (y || !ny));
    *consistent = 0;
    if (log_n < 1 || log_n > 26) return rt::fail(AZK_ERR_INVALID_LAYOUT, "domain log size %u outside 1..26", log_n);
    const uint64_t half = 1ull << (log_n - 1);
    if (nx > half || ny > half) return rt::fail(AZK_ERR_INVALID_LAYOUT, "more coefficients than the capacity");
    *consistent = basis_consistent(x, nx, y, ny, log_n, basis) ? 1 : 0;
    return AZK_OK;
}

} // extern "C"
