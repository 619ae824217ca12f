// Host verifier of Circle PCS/FRI proofs in CirclePCSFRIProofCodecV1 bytes: the service
// a downstream consumer runs on what the prover emits.
//
// Reference: CirclePCSFRIProofCodecV1.decode (Sources/AppleZKProver/CircleProofFormat.swift:
// 537-645) with the proof's construction rules (CirclePCSFRIProofV1.init :355-476, the
// domain descriptor codec :71-127, CircleFRISecurityParametersV1 :129-158), the
// transcript CircleFRITranscriptV1.derive (:684-983, grinding :753-794) and
// CirclePCSFRIProofVerifierV1.verify (:1397-1546).  Malformed bytes are a layout error
// (the reference's decode throws); a well-formed proof that does not verify is
// reported as not accepted (the reference returns false).
//
// Beside it, the statement-level checks a consumer of a polynomial commitment runs:
// CirclePCSFRIPolynomialClaimDigestV1, CirclePCSFRIPolynomialVerifierV1 and
// CirclePCSFRIContractVerifierV1 (:1095-1151,1548-1752).
#include <new>
#include <vector>

#include "host_field.hpp"
#include "host_gates.hpp"
#include "host_sha3.hpp"
#include "runtime_common.hpp"

using namespace azk;
using namespace azk::rt;
using namespace azk::host;

namespace {

// ---------------------------------------------------------------- decode
struct Reader {
    const uint8_t *p;
    size_t n, at = 0;
    bool ok = true;
    bool take(size_t k) { if (!ok || k > n - at) { ok = false; return false; } return true; }
    uint32_t u32() { if (!take(4)) return 0; uint32_t v; memcpy(&v, p + at, 4); at += 4; return v; }
    uint64_t u64() { if (!take(8)) return 0; uint64_t v; memcpy(&v, p + at, 8); at += 8; return v; }
    const uint8_t *bytes(size_t k) { if (!take(k)) return nullptr; const uint8_t *q = p + at; at += k; return q; }
    Q qm() { Q q{u32(), u32(), u32(), u32()}; return q; }
    bool end() const { return at == n; }
};

struct Opening { uint64_t leaf; Q value; const uint8_t *siblings; uint32_t count; };
struct Layer { uint32_t index; uint64_t pair; Opening left, right; };
struct Query { uint64_t pair; std::vector<Layer> layers; };

struct Proof {
    const uint8_t *descriptor = nullptr;  // the 44 bytes as read
    Domain dom;
    uint32_t log_blowup = 0, queries = 0, folding = 0, grinding = 0;
    const uint8_t *digest = nullptr;
    std::vector<const uint8_t *> roots;
    std::vector<Q> final_layer;
    std::vector<Query> qs;
    std::vector<Opening> claimed;
    bool has_nonce = false;
    uint64_t nonce = 0;
};

bool read_opening(Reader &r, Opening &o)
{
    o.leaf = r.u64();
    o.value = r.qm();
    o.count = r.u32();
    if (!r.ok || o.count > 64) return false;
    o.siblings = r.bytes((size_t)32 * o.count);
    return r.ok && canonical(o.value);
}

const uint8_t kProofMagic[8] = {0x41, 0x5a, 0x4b, 0x50, 0x46, 0x56, 0x31, 0x00};
const uint8_t kDescMagic[8] = {0x41, 0x5a, 0x4b, 0x43, 0x44, 0x56, 0x31, 0x00};

azk_status decode(const uint8_t *data, size_t len, Proof &pf)
{
    auto bad = [](const char *what) { return fail(AZK_ERR_INVALID_LAYOUT, "Circle PCS proof: %s", what); };
    Reader r{data, len};
    const uint8_t *magic = r.bytes(8);
    if (!magic || memcmp(magic, kProofMagic, 8)) return bad("magic");
    if (r.u32() != 1 || r.u32() != 1) return bad("proof or transcript version");
    // domain descriptor (44 bytes, length-prefixed)
    if (r.u32() != 44) return bad("descriptor length");
    pf.descriptor = r.bytes(44);
    if (!pf.descriptor) return bad("descriptor");
    {
        Reader d{pf.descriptor, 44};
        const uint8_t *m = d.bytes(8);
        if (memcmp(m, kDescMagic, 8) || d.u32() != 1 || d.u32() != kP || d.u32() != 31 || d.u32() != 2 ||
            d.u32() != 1268011823u)
            return bad("descriptor header");
        pf.dom.log_n = d.u32();
        pf.dom.init = d.u32() & 0x7fffffffu;
        pf.dom.half_log = d.u32();
        const uint32_t storage = d.u32();
        if (pf.dom.log_n < 1 || pf.dom.log_n > 30 || pf.dom.half_log != pf.dom.log_n - 1 || storage != 1)
            return bad("descriptor shape");
        // isCanonical: 4 * halfCosetInitialIndex = subgroup generator of the half coset
        if (((uint64_t)pf.dom.init * 4 & 0x7fffffffu) != ((1ull << (31 - pf.dom.half_log)) & 0x7fffffffu))
            return bad("descriptor coset");
    }
    pf.log_blowup = r.u32();
    pf.queries = r.u32();
    pf.folding = r.u32();
    pf.grinding = r.u32();
    if (!r.ok || !pf.log_blowup || !pf.queries || !pf.folding || pf.grinding > 64 ||
        (uint64_t)pf.log_blowup * pf.queries + pf.grinding > 0xffffffffull || pf.folding != 1)
        return bad("security parameters");
    if (r.u32() != 32) return bad("public input digest length");
    pf.digest = r.bytes(32);
    const uint32_t R = r.u32();
    if (!r.ok || R == 0 || R > pf.dom.log_n) return bad("commitment count");
    for (uint32_t k = 0; k < R; ++k) pf.roots.push_back(r.bytes(32));
    const uint64_t nfin = r.u64();
    if (!r.ok || nfin != (1ull << pf.dom.log_n) >> R) return bad("final layer size");
    // sizes come from untrusted counts: the bytes must be there before anything is allocated
    // (CanonicalByteReader.readBytes throws on short input, CircleProofFormat.swift:550-561)
    if (16 * nfin > r.n - r.at) return bad("final layer truncated");
    pf.final_layer.resize(nfin);
    for (auto &q : pf.final_layer) {
        q = r.qm();
        if (!canonical(q)) return bad("non-canonical final layer");
    }
    const uint32_t nq = r.u32();
    if (!r.ok || nq != pf.queries) return bad("query count");
    const uint64_t n = 1ull << pf.dom.log_n;
    // smallest encoded query: pair + layer count, then per layer index + pair + two openings
    // of at least one sibling each (leaf index, value, count, 32 bytes)
    const uint64_t min_query = 12 + (uint64_t)R * (12 + 2 * (28 + 32));
    if ((uint64_t)nq * min_query > r.n - r.at) return bad("queries truncated");
    pf.qs.resize(nq);
    for (Query &q : pf.qs) {
        q.pair = r.u64();
        const uint32_t nl = r.u32();
        if (!r.ok || nl != R || q.pair >= n / 2) return bad("query shape");
        q.layers.resize(nl);
        for (uint32_t k = 0; k < nl; ++k) {  // validateQueryShape (:420-446)
            Layer &l = q.layers[k];
            l.index = r.u32();
            l.pair = r.u64();
            if (!read_opening(r, l.left) || !read_opening(r, l.right)) return bad("opening");
            const uint64_t m = n >> k;
            const uint32_t depth = pf.dom.log_n - k;
            if (l.index != k || l.pair != q.pair >> k || m <= 1 || l.pair >= m / 2 || l.left.leaf != 2 * l.pair ||
                l.right.leaf != 2 * l.pair + 1 || l.left.count != depth || l.right.count != depth)
                return bad("query layer shape");
        }
    }
    if (!r.end()) {  // claimed evaluation openings, then the grinding nonce
        const uint32_t nc = r.u32();
        if (!r.ok || nc > n) return bad("claimed openings");
        uint64_t prev = 0;
        for (uint32_t k = 0; k < nc; ++k) {
            Opening o;
            if (!read_opening(r, o) || o.leaf >= n || o.count != pf.dom.log_n || (k && o.leaf <= prev))
                return bad("claimed opening shape");
            prev = o.leaf;
            pf.claimed.push_back(o);
        }
        if (!r.end()) {
            pf.nonce = r.u64();
            pf.has_nonce = true;
        }
    }
    if (!r.ok || !r.end()) return bad("trailing or truncated bytes");
    if ((pf.grinding > 0) != pf.has_nonce) return bad("grinding nonce shape");
    return AZK_OK;
}

// ---------------------------------------------------------------- transcript
void put32(std::vector<uint8_t> &f, uint32_t v) { for (int i = 0; i < 4; ++i) f.push_back((uint8_t)(v >> (8 * i))); }
void put64(std::vector<uint8_t> &f, uint64_t v) { for (int i = 0; i < 8; ++i) f.push_back((uint8_t)(v >> (8 * i))); }
std::vector<uint8_t> frame(uint8_t type)
{
    static const char dom[] = "AppleZKProver.CircleFRI.PCS.V1";
    std::vector<uint8_t> f;
    put32(f, (uint32_t)(sizeof dom - 1));
    f.insert(f.end(), dom, dom + sizeof dom - 1);
    put32(f, 1);
    f.push_back(type);
    return f;
}

// CircleFRITranscriptV1.derive; false when the grinding nonce misses its target.
bool derive(const Proof &pf, std::vector<Q> &challenges, std::vector<uint32_t> &pairs)
{
    host::Transcript t;
    auto absorb = [&](const std::vector<uint8_t> &f) { t.absorb(f.data(), f.size()); };
    std::vector<uint8_t> f = frame(0);
    for (uint32_t w : {1u, 1u, kP, 2u, 1268011823u}) put32(f, w);
    absorb(f);
    // the descriptor as CircleDomainDescriptorCodecV1.encode writes it (:75-93)
    f.assign(kDescMagic, kDescMagic + 8);
    for (uint32_t w : {1u, kP, 31u, 2u, 1268011823u, pf.dom.log_n, pf.dom.init, pf.dom.half_log, 1u}) put32(f, w);
    absorb(f);
    f = frame(1);
    for (uint32_t w : {pf.log_blowup, pf.queries, pf.folding, pf.grinding, pf.log_blowup * pf.queries + pf.grinding})
        put32(f, w);
    absorb(f);
    f = frame(2);
    put32(f, 32);
    absorb(f);
    t.absorb(pf.digest, 32);
    for (uint32_t k = 0; k < pf.roots.size(); ++k) {
        f = frame(3);
        put32(f, k);
        put32(f, 32);
        absorb(f);
        t.absorb(pf.roots[k], 32);
        f = frame(4);
        put32(f, k);
        put32(f, kP);
        put32(f, 4);
        absorb(f);
        uint32_t l[4];
        t.squeeze_u32(4, kP, l);
        challenges.push_back({l[0], l[1], l[2], l[3]});
    }
    f = frame(5);
    put64(f, pf.final_layer.size());
    put64(f, pf.final_layer.size() * 16);
    absorb(f);
    std::vector<uint8_t> fl(pf.final_layer.size() * 16);
    for (size_t k = 0; k < pf.final_layer.size(); ++k) memcpy(fl.data() + 16 * k, &pf.final_layer[k], 16);
    t.absorb(fl.data(), fl.size());
    if (pf.grinding) {  // applyGrindingIfNeeded: frame, nonce, then 32 squeezed bytes
        f = frame(7);
        put32(f, pf.grinding);
        put32(f, 8);
        absorb(f);
        uint8_t nb[8];
        memcpy(nb, &pf.nonce, 8);
        t.absorb(nb, 8);
        uint8_t dg[32];
        memcpy(dg, t.s, 32);
        if (!azk_circle_grinding_digest_meets_target(dg, 32, pf.grinding)) return false;
    }
    f = frame(6);
    put32(f, pf.queries);
    put64(f, (1ull << pf.dom.log_n) / 2);
    absorb(f);
    pairs.resize(pf.queries);
    t.squeeze_u32(pf.queries, (uint32_t)((1ull << pf.dom.log_n) / 2), pairs.data());
    return true;
}

// MerkleOracle.verifySHA3_256 of a 16-byte QM31 leaf.
bool verify_opening(const Opening &o, const uint8_t *root)
{
    uint8_t node[32], msg[64];
    host::sha3_256(reinterpret_cast<const uint8_t *>(&o.value), 16, node);
    uint64_t idx = o.leaf;
    for (uint32_t k = 0; k < o.count; ++k, idx >>= 1) {
        const uint8_t *sib = o.siblings + 32 * k;
        if (idx & 1) { memcpy(msg, sib, 32); memcpy(msg + 32, node, 32); }
        else { memcpy(msg, node, 32); memcpy(msg + 32, sib, 32); }
        host::sha3_256(msg, 64, node);
    }
    return memcmp(node, root, 32) == 0;
}

bool verify(const Proof &pf, const uint8_t *expected_digest)
{
    if (memcmp(pf.digest, expected_digest, 32)) return false;
    std::vector<Q> ch;
    std::vector<uint32_t> pairs;
    if (!derive(pf, ch, pairs)) return false;
    const Q inv2{1u << 30, 0, 0, 0};
    for (size_t qi = 0; qi < pf.qs.size(); ++qi) {
        const Query &q = pf.qs[qi];
        if (q.pair != pairs[qi]) return false;
        uint64_t pair = pairs[qi];
        bool have = false;
        uint64_t exp_index = 0;
        Q exp_value{};
        for (uint32_t r = 0; r < q.layers.size(); ++r) {
            const Layer &l = q.layers[r];
            if (!verify_opening(l.left, pf.roots[r]) || !verify_opening(l.right, pf.roots[r])) return false;
            if (have) {
                if (exp_index == 2 * pair) { if (!(exp_value == l.left.value)) return false; }
                else if (exp_index == 2 * pair + 1) { if (!(exp_value == l.right.value)) return false; }
                else return false;
            }
            const uint32_t coord = pf.dom.fold_coordinate(r, pair);
            if (coord == 0) return false;
            const Q t{finv(coord), 0, 0, 0};
            const Q even = qadd(l.left.value, l.right.value), odd = qsub(l.left.value, l.right.value);
            exp_value = qmul(qadd(even, qmul(ch[r], qmul(odd, t))), inv2);
            exp_index = pair;
            have = true;
            pair >>= 1;
        }
        if (!have || exp_index >= pf.final_layer.size() || !(pf.final_layer[exp_index] == exp_value)) return false;
    }
    return true;
}

} // namespace

// The proof verifies, and every opened first-layer value equals the Circle FFT of the
// line-basis coefficients at that leaf: the CPU gate of the prover's *_verified calls.
extern "C" azk_status azk_circle_verify_with_coefficients(const uint8_t *proof, size_t len,
                                                          const uint8_t *public_input_digest, const uint32_t *coeffs,
                                                          size_t coeff_words, int *accepted)
{
    CHECK_ARG(proof && public_input_digest && coeffs && accepted);
    *accepted = 0;
    try {
        Proof pf;
        azk_status s = decode(proof, len, pf);
        if (s) return s;
        const size_t n = (size_t)1 << pf.dom.log_n;
        if (pf.dom.log_n > 26 || coeff_words != 4 * n)
            return fail(AZK_ERR_INVALID_LAYOUT, "need %zu coefficient words for the proof's 2^%u domain", 4 * n, pf.dom.log_n);
        for (size_t i = 0; i < coeff_words; ++i)
            if (coeffs[i] >= kP) return fail(AZK_ERR_INVALID_LAYOUT, "non-canonical coefficient");
        if (!verify(pf, public_input_digest)) return AZK_OK;
        std::vector<uint32_t> cw(4 * n);
        host::circle_fft(coeffs, pf.dom.log_n, cw.data());
        auto same = [&](const Opening &o) { return memcmp(&o.value, cw.data() + 4 * o.leaf, 16) == 0; };
        for (const Query &q : pf.qs)
            if (!same(q.layers[0].left) || !same(q.layers[0].right)) return AZK_OK;
        for (const Opening &o : pf.claimed)
            if (!same(o)) return AZK_OK;
        *accepted = 1;
    } catch (const std::bad_alloc &) {
        return fail(AZK_ERR_ALLOCATION, "Circle PCS proof: out of host memory");
    }
    return AZK_OK;
}

extern "C" azk_status azk_circle_verify(const uint8_t *proof, size_t len, const uint8_t *public_input_digest,
                                        int *accepted)
{
    CHECK_ARG(proof && public_input_digest && accepted);
    *accepted = 0;
    try {
        Proof pf;
        azk_status s = decode(proof, len, pf);
        if (s) return s;
        *accepted = verify(pf, public_input_digest) ? 1 : 0;
    } catch (const std::bad_alloc &) {
        return fail(AZK_ERR_ALLOCATION, "Circle PCS proof: out of host memory");
    }
    return AZK_OK;
}

// ====================================================================== grinding target

extern "C" int azk_circle_grinding_digest_meets_target(const uint8_t *digest, size_t len, uint32_t bits)
{
    if ((uint64_t)bits > 8 * (uint64_t)len || (!digest && bits)) return 0;
    size_t i = 0;
    for (; bits >= 8; bits -= 8, ++i)
        if (digest[i]) return 0;
    return !bits || (digest[i] & (uint8_t)(0xffu << (8 - bits))) == 0;
}

// ====================================================================== polynomial claims

namespace {

// CirclePCSFRIPolynomialClaimV1.init, CircleCodewordPolynomial.init and
// CirclePCSFRIEvaluationClaimV1.init (CircleProofFormat.swift:1000-1046, CircleCodeword.swift:7-20):
// with `with_values`, the claimed points must lie on the circle and the values be canonical.
azk_status check_claim(const azk_circle_polynomial_claim *c, bool with_values)
{
    auto bad = [](const char *what) { return fail(AZK_ERR_INVALID_LAYOUT, "polynomial claim: %s", what); };
    if (!c) return fail(AZK_ERR_INVALID_ARGUMENT, "null claim");
    if (c->log_n < 1 || c->log_n > 30) return bad("domain log size outside 1..30");
    if (!c->x_count && !c->y_count) return bad("no coefficients");
    if ((c->x_count && !c->x_coeffs) || (c->y_count && !c->y_coeffs)) return fail(AZK_ERR_INVALID_ARGUMENT, "null coefficients");
    for (size_t i = 0; i < 4ull * c->x_count; ++i)
        if (c->x_coeffs[i] >= kP) return bad("non-canonical x coefficient");
    for (size_t i = 0; i < 4ull * c->y_count; ++i)
        if (c->y_coeffs[i] >= kP) return bad("non-canonical y coefficient");
    if (!c->claim_count) return bad("no evaluation claims");
    if (!c->storage_indices) return fail(AZK_ERR_INVALID_ARGUMENT, "null storage indices");
    const uint64_t n = 1ull << c->log_n;
    for (uint32_t k = 0; k < c->claim_count; ++k)
        if (c->storage_indices[k] >= n || (k && c->storage_indices[k] <= c->storage_indices[k - 1]))
            return bad("storage indices not strictly increasing inside the domain");
    if (with_values) {
        if (!c->points || !c->values) return fail(AZK_ERR_INVALID_ARGUMENT, "null claim points or values");
        for (uint32_t k = 0; k < c->claim_count; ++k) {
            const uint32_t x = c->points[2 * k], y = c->points[2 * k + 1];
            if (x >= kP || y >= kP || fadd(fmul(x, x), fmul(y, y)) != 1) return bad("claimed point off the circle");
            for (int w = 0; w < 4; ++w)
                if (c->values[4 * k + w] >= kP) return bad("non-canonical claimed value");
        }
    }
    return AZK_OK;
}

// CircleCodewordOracle.evaluate(polynomial:at:) (CircleCodeword.swift:53-70,173-185):
// X(x) + y Y(x), Horner in x over QM31 coefficients.
Q evaluate_at(const azk_circle_polynomial_claim *c, Pt p)
{
    auto horner = [&](const uint32_t *co, uint32_t count) {
        Q acc{0, 0, 0, 0};
        for (uint32_t i = count; i-- > 0;) acc = qadd(qscale(acc, p.x), Q{co[4 * i], co[4 * i + 1], co[4 * i + 2], co[4 * i + 3]});
        return acc;
    };
    return qadd(horner(c->x_coeffs, c->x_count), qscale(horner(c->y_coeffs, c->y_count), p.y));
}

// CirclePCSFRIPolynomialClaimDigestV1 frames (:1118-1150): u32 len || domain string || u8 type.
std::vector<uint8_t> claim_frame(uint8_t type)
{
    static const char dom[] = "AppleZKProver.CirclePCS.PolynomialClaim.V1";
    std::vector<uint8_t> f;
    put32(f, (uint32_t)(sizeof dom - 1));
    f.insert(f.end(), dom, dom + sizeof dom - 1);
    f.push_back(type);
    return f;
}

void claim_digest(const azk_circle_polynomial_claim *c, uint8_t out[32])
{
    static const char dom[] = "AppleZKProver.CirclePCS.PolynomialClaim.V1";
    std::vector<uint8_t> m = claim_frame(0);  // domain frame: the string again, proof and transcript versions
    put32(m, (uint32_t)(sizeof dom - 1));
    m.insert(m.end(), dom, dom + sizeof dom - 1);
    put32(m, 1);
    put32(m, 1);
    m.insert(m.end(), kDescMagic, kDescMagic + 8);  // CircleDomainDescriptorCodecV1.encode
    for (uint32_t w : {1u, kP, 31u, 2u, 1268011823u, c->log_n, 1u << (30 - c->log_n), c->log_n - 1, 1u}) put32(m, w);
    std::vector<uint8_t> f = claim_frame(1);  // polynomial frame: counts and the 16-byte element size
    put32(f, c->x_count);
    put32(f, c->y_count);
    put32(f, 16);
    m.insert(m.end(), f.begin(), f.end());
    for (size_t i = 0; i < 4ull * c->x_count; ++i) put32(m, c->x_coeffs[i]);
    for (size_t i = 0; i < 4ull * c->y_count; ++i) put32(m, c->y_coeffs[i]);
    f = claim_frame(2);  // evaluation-claim frame
    put32(f, c->claim_count);
    m.insert(m.end(), f.begin(), f.end());
    for (uint32_t k = 0; k < c->claim_count; ++k) {
        put64(m, c->storage_indices[k]);
        put32(m, c->points[2 * k]);
        put32(m, c->points[2 * k + 1]);
        for (int w = 0; w < 4; ++w) put32(m, c->values[4 * k + w]);
    }
    host::sha3_256(m.data(), m.size(), out);
}

// CirclePCSFRIPolynomialVerifierV1.verify (:1548-1659) on a decoded proof.
bool verify_polynomial(const Proof &pf, const azk_circle_polynomial_claim *c)
{
    uint8_t digest[32];
    claim_digest(c, digest);
    if (memcmp(pf.digest, digest, 32) || pf.dom.log_n != c->log_n || pf.claimed.size() != c->claim_count) return false;
    // coefficientCountsFitDomainAndSecurity (:1585-1599)
    if (pf.log_blowup > pf.dom.log_n) return false;
    const uint64_t cap = std::max<uint64_t>(1, (1ull << pf.dom.log_n) >> pf.log_blowup);
    if (c->x_count > cap || c->y_count > cap) return false;
    if (!verify(pf, digest)) return false;
    const Domain dom = Domain::canonical_of(c->log_n);
    for (uint32_t k = 0; k < c->claim_count; ++k) {
        const Opening &o = pf.claimed[k];
        const Pt want = dom.storage_point(c->storage_indices[k]);  // claimMatchesDomainAndPolynomial (:1601-1625)
        const Q value{c->values[4 * k], c->values[4 * k + 1], c->values[4 * k + 2], c->values[4 * k + 3]};
        if (want.x != c->points[2 * k] || want.y != c->points[2 * k + 1] || !(evaluate_at(c, want) == value)) return false;
        if (o.leaf != c->storage_indices[k] || !(o.value == value)) return false;
        if (o.count != c->log_n || !verify_opening(o, pf.roots[0])) return false;  // verifyClaimOpening (:1627-1647)
    }
    return true;
}

} // namespace

extern "C" azk_status azk_circle_polynomial_claim_make(const azk_circle_polynomial_claim *claim, uint32_t *points,
                                                       uint32_t *values)
{
    azk_status s = check_claim(claim, false);
    if (s) return s;
    CHECK_ARG(points && values);
    const Domain dom = Domain::canonical_of(claim->log_n);
    for (uint32_t k = 0; k < claim->claim_count; ++k) {
        const Pt p = dom.storage_point(claim->storage_indices[k]);
        const Q v = evaluate_at(claim, p);
        points[2 * k] = p.x;
        points[2 * k + 1] = p.y;
        memcpy(values + 4 * k, &v, 16);
    }
    return AZK_OK;
}

extern "C" azk_status azk_circle_polynomial_claim_digest(const azk_circle_polynomial_claim *claim, uint8_t digest[32])
{
    azk_status s = check_claim(claim, true);
    if (s) return s;
    CHECK_ARG(digest);
    try {
        claim_digest(claim, digest);
    } catch (const std::bad_alloc &) {
        return fail(AZK_ERR_ALLOCATION, "polynomial claim digest: out of host memory");
    }
    return AZK_OK;
}

extern "C" azk_status azk_circle_verify_polynomial(const uint8_t *proof, size_t len, const azk_circle_polynomial_claim *claim,
                                                   int *accepted)
{
    CHECK_ARG(proof && accepted);
    *accepted = 0;
    azk_status s = check_claim(claim, true);
    if (s) return s;
    try {
        Proof pf;
        s = decode(proof, len, pf);
        if (s) return s;
        *accepted = verify_polynomial(pf, claim) ? 1 : 0;
    } catch (const std::bad_alloc &) {
        return fail(AZK_ERR_ALLOCATION, "Circle PCS proof: out of host memory");
    }
    return AZK_OK;
}

extern "C" azk_status azk_circle_verify_contract(const uint8_t *proof, size_t len, const azk_circle_parameter_set *params,
                                                 const azk_circle_polynomial_claim *claim, int *accepted)
{
    CHECK_ARG(proof && params && accepted);
    *accepted = 0;
    azk_status s = check_claim(claim, true);
    if (s) return s;
    // CirclePCSFRIParameterSetV1.init / CircleFRISecurityParametersV1.init (:135-153,216-236)
    const uint64_t nominal = (uint64_t)params->log_blowup * params->queries + params->grinding_bits;
    if (!params->log_blowup || !params->queries || params->grinding_bits > 64 || nominal > 0xffffffffull ||
        !params->target_soundness_bits || nominal < params->target_soundness_bits)
        return fail(AZK_ERR_INVALID_LAYOUT, "Circle PCS parameter set: invalid security parameters");
    // CirclePCSFRIStatementV1.init: validateDomain and the coefficient budget (:1664-1700)
    if (claim->log_n <= params->log_blowup)
        return fail(AZK_ERR_INVALID_LAYOUT, "Circle PCS statement: domain 2^%u not above the blowup 2^%u", claim->log_n,
                    params->log_blowup);
    const uint64_t used = (uint64_t)claim->x_count + claim->y_count;
    if (used > ((1ull << claim->log_n) >> params->log_blowup))
        return fail(AZK_ERR_INVALID_LAYOUT, "Circle PCS statement: %llu coefficients over the committed capacity",
                    (unsigned long long)used);
    try {
        Proof pf;
        s = decode(proof, len, pf);
        if (s) return s;
        // acceptsProofShape (:270-285)
        const uint32_t rounds = claim->log_n - params->log_blowup;
        bool ok = pf.dom.log_n == claim->log_n && pf.log_blowup == params->log_blowup && pf.queries == params->queries &&
                  pf.folding == 1 && pf.grinding == params->grinding_bits && pf.roots.size() == rounds &&
                  pf.final_layer.size() == (1ull << params->log_blowup);
        for (size_t k = 1; ok && k < pf.final_layer.size(); ++k) ok = pf.final_layer[k] == pf.final_layer[0];
        *accepted = ok && verify_polynomial(pf, claim) ? 1 : 0;
    } catch (const std::bad_alloc &) {
        return fail(AZK_ERR_ALLOCATION, "Circle PCS proof: out of host memory");
    }
    return AZK_OK;
}
