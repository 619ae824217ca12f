// M31 sum-check chunk plan: the reference's GPU sum-check.  Each round absorbs the whole active
// vector into the transcript (a round frame, the words LE, a challenge frame), squeezes one
// challenge r mod p and folds next[i] = v[2i] + r v[2i + 1].  The proof (M31SumcheckProofCodecV1)
// carries every round's vector, so every vector comes back to the host regardless; the device
// keeps the vector resident and folds it (an HBM-bound pass), the host runs the one serial sponge
// over the bytes it receives, and the initial-vector digest runs on a second host thread.
//
// Reference: SumcheckOracle.m31Chunk (Sources/AppleZKProver/SumcheckOracle.swift:212-276),
// MetalSumcheckChunkPlan (Planner/MetalSumcheckChunkPlan.swift:42-606, executeLocked :260-399),
// kernels sumcheck_round_eval_u32 / sumcheck_fold_halve_u32 (Resources/HashMerkleKernels.metal:
// 2185-2223), framing SumcheckTranscriptFraming (SumcheckOracle.swift:1311-1360), the proof
// M31SumcheckProofBuilderV1 / M31SumcheckProofCodecV1 / M31SumcheckVerifierV1 (:540-700,1108-1175).
#include <new>
#include <thread>

#include "device_field.hpp"
#include "host_field.hpp"
#include "host_sha3.hpp"
#include "runtime_common.hpp"

using namespace azk;
using namespace azk::rt;

namespace {

__global__ __launch_bounds__(256) void k_chunk_fold(const uint2 *__restrict__ cur, uint32_t *__restrict__ next,
                                                    uint64_t pairs, uint32_t challenge)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= pairs) return;
    const uint2 v = cur[i];
    next[i] = m31_add(v.x, m31_mul(v.y, challenge));
}

void put32(std::vector<uint8_t> &f, uint32_t v) { for (int i = 0; i < 4; ++i) f.push_back((uint8_t)(v >> (8 * i))); }
void put64(std::vector<uint8_t> &f, uint64_t v) { for (int i = 0; i < 8; ++i) f.push_back((uint8_t)(v >> (8 * i))); }

// SumcheckTranscriptFraming.baseFrame: u32 len || "AppleZKProver.SumcheckChunk" || u32 1 || u8 type.
std::vector<uint8_t> frame(uint8_t type)
{
    static const char dom[] = "AppleZKProver.SumcheckChunk";
    std::vector<uint8_t> f;
    put32(f, (uint32_t)(sizeof dom - 1));
    f.insert(f.end(), dom, dom + sizeof dom - 1);
    put32(f, 1);
    f.push_back(type);
    return f;
}
std::vector<uint8_t> header_frame(uint64_t n, uint32_t rounds)
{
    std::vector<uint8_t> f = frame(0);
    put64(f, n);
    put32(f, rounds);
    put32(f, kP);
    return f;
}
std::vector<uint8_t> round_frame(uint32_t r, uint64_t active)
{
    std::vector<uint8_t> f = frame(1);
    put32(f, r);
    put64(f, active);
    put64(f, active);  // coefficient words: the whole active vector
    return f;
}
std::vector<uint8_t> challenge_frame(uint32_t r)
{
    std::vector<uint8_t> f = frame(2);
    put32(f, r);
    put32(f, kP);
    return f;
}

// Absorbs one round's vector and squeezes its challenge.
uint32_t round_challenge(host::Transcript &t, uint32_t r, const uint32_t *v, uint64_t active)
{
    const std::vector<uint8_t> rf = round_frame(r, active), cf = challenge_frame(r);
    t.absorb(rf.data(), rf.size());
    t.absorb(reinterpret_cast<const uint8_t *>(v), 4 * active);  // words LE on a little-endian host
    t.absorb(cf.data(), cf.size());
    uint32_t c = 0;
    t.squeeze_u32(1, kP, &c);
    return c;
}

uint64_t coefficient_words(uint64_t n, uint32_t rounds)
{
    uint64_t total = 0;
    for (uint32_t r = 0; r < rounds; ++r) total += n >> r;
    return total;
}

constexpr uint8_t kChunkMagic[8] = {0x41, 0x5a, 0x4b, 0x53, 0x43, 0x56, 0x31, 0x00};  // "AZKSCV1\0"

} // namespace

struct azk_sumcheck_chunk_plan {
    azk_context *ctx = nullptr;
    uint64_t n = 0;
    uint32_t rounds = 0;
    DeviceMem vec;      // two halves: the ping-pong buffers of the folded vectors
    DeviceMem staging;  // host-array inputs, uploaded on first use
    std::mutex mu;
};

extern "C" azk_status azk_sumcheck_chunk_plan_create(azk_context *ctx, uint64_t lane_count, uint32_t rounds,
                                                     azk_sumcheck_chunk_plan **out)
{
    CHECK_ARG(ctx && out);
    // SumcheckChunkDescriptor validation (MetalSumcheckChunkPlan.swift:490-498)
    if (!is_pow2(lane_count) || lane_count < 2 || lane_count > (1ull << 30) || rounds < 1 || rounds > ilog2(lane_count))
        return fail(AZK_ERR_INVALID_LAYOUT, "sum-check chunk: %llu lanes (power of two, 2..2^30), %u rounds (1..log2)",
                    (unsigned long long)lane_count, rounds);
    HIP_TRY(hipSetDevice(ctx->device));
    auto *p = new azk_sumcheck_chunk_plan();
    p->ctx = ctx;
    p->n = lane_count;
    p->rounds = rounds;
    azk_status s = p->vec.ensure(4 * lane_count);  // n/2 + n/4 <= n words
    if (s) {
        delete p;
        return s;
    }
    ctx_retain(ctx);
    *out = p;
    return AZK_OK;
}

extern "C" void azk_sumcheck_chunk_plan_destroy(azk_sumcheck_chunk_plan *p) { destroy_on_ctx(p); }

extern "C" size_t azk_sumcheck_chunk_coefficient_words(uint64_t lane_count, uint32_t rounds)
{
    return (size_t)coefficient_words(lane_count, rounds);
}

extern "C" size_t azk_sumcheck_chunk_proof_size(uint64_t lane_count, uint32_t rounds)
{
    if (!is_pow2(lane_count) || lane_count < 2 || rounds < 1 || rounds > ilog2(lane_count)) return 0;
    return 28 + 64 + 3 * 8 + 4 * ((lane_count >> rounds) + coefficient_words(lane_count, rounds) + rounds);
}

namespace {

// The rounds on the device: `src` (device, n words) is read, never written; `host0` is the
// round-0 vector if the caller already holds it on the host.
azk_status chunk_run(azk_sumcheck_chunk_plan *p, const uint32_t *src, const uint32_t *host0, uint32_t *final_vector,
                     uint32_t *coefficients, uint32_t *challenges, uint8_t initial_digest[32], azk_stats *stats)
{
    const uint64_t n = p->n;
    const uint32_t R = p->rounds;
    hipStream_t st = p->ctx->stream;
    uint32_t *half[2] = {(uint32_t *)p->vec.p, (uint32_t *)p->vec.p + n / 2};
    host::Transcript t;
    const std::vector<uint8_t> hf = header_frame(n, R);
    t.absorb(hf.data(), hf.size());
    std::thread digest;
    auto join = [&]() { if (digest.joinable()) digest.join(); };
    const uint64_t w0 = now_ns();
    Timer tm;
    azk_status s = tm.init();
    if (s) return s;
    HIP_TRY(hipEventRecord(tm.a, st));
    const uint32_t *cur = src;
    uint64_t off = 0;
    for (uint32_t r = 0; r < R; ++r) {
        const uint64_t active = n >> r;
        uint32_t *v = coefficients + off;  // round r's vector, as the proof logs it
        if (r == 0 && host0) memcpy(v, host0, 4 * n);
        else if (hipMemcpyAsync(v, cur, 4 * active, hipMemcpyDeviceToHost, st) != hipSuccess ||
                 hipStreamSynchronize(st) != hipSuccess) {
            join();
            return fail(AZK_ERR_COMMAND_FAILED, "sum-check chunk: readback of round %u", r);
        }
        if (r == 0 && initial_digest) digest = std::thread([&]() { host::sumcheck_words_digest(coefficients, n, initial_digest); });
        const uint32_t c = round_challenge(t, r, v, active);
        challenges[r] = c;
        uint32_t *next = half[r & 1];
        const uint64_t pairs = active / 2;
        hipLaunchKernelGGL(k_chunk_fold, dim3((uint32_t)((pairs + 255) / 256)), dim3(256), 0, st, (const uint2 *)cur, next,
                           pairs, c);
        cur = next;
        off += active;
    }
    const bool ok = hipGetLastError() == hipSuccess &&
                    hipMemcpyAsync(final_vector, cur, 4 * (n >> R), hipMemcpyDeviceToHost, st) == hipSuccess &&
                    hipEventRecord(tm.b, st) == hipSuccess && hipStreamSynchronize(st) == hipSuccess;
    join();
    if (!ok) return fail(AZK_ERR_COMMAND_FAILED, "sum-check chunk: device rounds failed");
    if (stats) {
        stats->gpu_ns = tm.elapsed_ns();
        stats->wall_ns = now_ns() - w0;
        stats->submit_ns = 0;
    }
    return AZK_OK;
}

azk_status check_outputs(const uint32_t *final_vector, const uint32_t *coefficients, const uint32_t *challenges)
{
    if (!final_vector || !coefficients || !challenges) return fail(AZK_ERR_INVALID_ARGUMENT, "null output array");
    return AZK_OK;
}

azk_status check_host_vector(const azk_sumcheck_chunk_plan *p, const uint32_t *evals, size_t count)
{
    if (!evals || count != p->n) return fail(AZK_ERR_INVALID_LAYOUT, "sum-check chunk: %zu evaluations for %llu lanes", count,
                                             (unsigned long long)p->n);
    for (size_t i = 0; i < count; ++i)  // validateCanonicalEvaluations (:483-488)
        if (evals[i] >= kP) return fail(AZK_ERR_INVALID_LAYOUT, "non-canonical evaluation at %zu", i);
    return AZK_OK;
}

// Uploads a host vector into the plan's staging buffer (on the context stream, ahead of the folds).
azk_status upload_input(azk_sumcheck_chunk_plan *p, const uint32_t *evals, const uint32_t **dev)
{
    azk_status s = p->staging.ensure(4 * p->n);
    if (s) return s;
    HIP_TRY(hipMemcpyAsync(p->staging.p, evals, 4 * p->n, hipMemcpyHostToDevice, p->ctx->stream));
    *dev = (const uint32_t *)p->staging.p;
    return AZK_OK;
}

} // namespace

extern "C" azk_status azk_sumcheck_chunk_execute(azk_sumcheck_chunk_plan *p, const uint32_t *evals, size_t count,
                                                 uint32_t *final_vector, uint32_t *coefficients, uint32_t *challenges,
                                                 azk_stats *stats)
{
    CHECK_ARG(p);
    azk_status s = check_host_vector(p, evals, count);
    if (!s) s = check_outputs(final_vector, coefficients, challenges);
    if (s) return s;
    std::lock_guard<std::mutex> g(p->mu);
    HIP_TRY(hipSetDevice(p->ctx->device));
    std::lock_guard<std::mutex> gc(p->ctx->mu);
    const uint32_t *dev = nullptr;
    s = upload_input(p, evals, &dev);
    if (s) return s;
    return chunk_run(p, dev, evals, final_vector, coefficients, challenges, nullptr, stats);
}

extern "C" azk_status azk_sumcheck_chunk_execute_resident(azk_sumcheck_chunk_plan *p, const azk_buffer *evals, size_t offset,
                                                          uint32_t *final_vector, uint32_t *coefficients,
                                                          uint32_t *challenges, azk_stats *stats)
{
    CHECK_ARG(p && evals);
    azk_status s = check_range(evals, offset, 4 * p->n);
    if (!s) s = check_outputs(final_vector, coefficients, challenges);
    if (s) return s;
    if (((uintptr_t)evals->ptr + offset) % 8) return fail(AZK_ERR_INVALID_LAYOUT, "evaluations must be 8-byte aligned");
    std::lock_guard<std::mutex> g(p->mu);
    HIP_TRY(hipSetDevice(p->ctx->device));
    std::lock_guard<std::mutex> gc(p->ctx->mu);
    return chunk_run(p, (const uint32_t *)((const uint8_t *)evals->ptr + offset), nullptr, final_vector, coefficients,
                     challenges, nullptr, stats);
}

extern "C" azk_status azk_sumcheck_chunk_execute_verified(azk_sumcheck_chunk_plan *p, const uint32_t *evals, size_t count,
                                                          uint32_t *final_vector, uint32_t *coefficients,
                                                          uint32_t *challenges, azk_stats *stats)
{
    azk_status s = azk_sumcheck_chunk_execute(p, evals, count, final_vector, coefficients, challenges, stats);
    if (s) return s;
    // executeVerified (:223-233): the whole chunk recomputed on the host
    try {
        std::vector<uint32_t> cur(evals, evals + count), next;
        host::Transcript t;
        const std::vector<uint8_t> hf = header_frame(p->n, p->rounds);
        t.absorb(hf.data(), hf.size());
        uint64_t off = 0;
        for (uint32_t r = 0; r < p->rounds; ++r) {
            if (memcmp(coefficients + off, cur.data(), 4 * cur.size())) goto mismatch;
            const uint32_t c = round_challenge(t, r, cur.data(), cur.size());
            if (c != challenges[r]) goto mismatch;
            next.resize(cur.size() / 2);
            for (size_t i = 0; i < next.size(); ++i) next[i] = host::fadd(cur[2 * i], host::fmul(cur[2 * i + 1], c));
            off += cur.size();
            cur.swap(next);
        }
        if (memcmp(final_vector, cur.data(), 4 * cur.size())) goto mismatch;
    } catch (const std::bad_alloc &) {
        return fail(AZK_ERR_ALLOCATION, "sum-check chunk gate: out of host memory");
    }
    return AZK_OK;
mismatch:
    return fail(AZK_ERR_CORRECTNESS, "M31 sum-check GPU chunk did not match the CPU gate");
}

extern "C" azk_status azk_sumcheck_chunk_plan_clear(azk_sumcheck_chunk_plan *p)
{
    CHECK_ARG(p);
    std::lock_guard<std::mutex> g(p->mu);
    HIP_TRY(hipSetDevice(p->ctx->device));
    HIP_TRY(hipMemset(p->vec.p, 0, p->vec.n));
    if (p->staging.p) HIP_TRY(hipMemset(p->staging.p, 0, p->staging.n));
    return AZK_OK;
}

extern "C" azk_status azk_sumcheck_chunk_prove(azk_sumcheck_chunk_plan *p, const uint32_t *evals, size_t count,
                                               uint8_t *proof, size_t capacity, size_t *len, azk_stats *stats)
{
    CHECK_ARG(p && proof);
    const size_t need = azk_sumcheck_chunk_proof_size(p->n, p->rounds);
    if (capacity < need) return fail(AZK_ERR_INVALID_LAYOUT, "proof capacity %zu < %zu", capacity, need);
    azk_status s = check_host_vector(p, evals, count);
    if (s) return s;
    const uint64_t n = p->n, fin = n >> p->rounds, nc = coefficient_words(n, p->rounds);
    // the words go straight into their places in the proof (aligned copies of 4-byte words)
    std::vector<uint32_t> final_vector(fin), coeffs(nc), ch(p->rounds);
    uint8_t d0[32], d1[32];
    {
        std::lock_guard<std::mutex> g(p->mu);
        HIP_TRY(hipSetDevice(p->ctx->device));
        std::lock_guard<std::mutex> gc(p->ctx->mu);
        const uint32_t *dev = nullptr;
        s = upload_input(p, evals, &dev);
        if (!s) s = chunk_run(p, dev, evals, final_vector.data(), coeffs.data(), ch.data(), d0, stats);
        if (s) return s;
    }
    host::sumcheck_words_digest(final_vector.data(), fin, d1);
    // M31SumcheckProofCodecV1.encode (:1111-1124)
    std::vector<uint8_t> b;
    b.reserve(need);
    b.insert(b.end(), kChunkMagic, kChunkMagic + 8);
    put32(b, 1);
    put32(b, 1);
    put64(b, n);
    put32(b, p->rounds);
    b.insert(b.end(), d0, d0 + 32);
    b.insert(b.end(), d1, d1 + 32);
    for (const std::vector<uint32_t> *w : {&final_vector, &coeffs, &ch}) {
        put64(b, w->size());
        const size_t at = b.size();
        b.resize(at + 4 * w->size());
        memcpy(b.data() + at, w->data(), 4 * w->size());
    }
    memcpy(proof, b.data(), b.size());
    if (len) *len = b.size();
    return AZK_OK;
}

// M31SumcheckProofCodecV1.decode + M31SumcheckVerifierV1.verify(proof:statement:) with the
// proof's own statement, scope revealedEvaluationVectorFoldingTrace (SumcheckOracle.swift:558-700).
extern "C" azk_status azk_sumcheck_chunk_verify(const uint8_t *proof, size_t len, int *accepted)
{
    CHECK_ARG(proof && accepted);
    *accepted = 0;
    auto bad = [](const char *what) { return fail(AZK_ERR_INVALID_LAYOUT, "sum-check chunk proof: %s", what); };
    size_t at = 0;
    auto take = [&](size_t k) -> const uint8_t * {
        if (k > len - at) return nullptr;
        const uint8_t *q = proof + at;
        at += k;
        return q;
    };
    auto u32 = [&](uint32_t &v) { const uint8_t *q = take(4); if (q) memcpy(&v, q, 4); return q != nullptr; };
    auto u64 = [&](uint64_t &v) { const uint8_t *q = take(8); if (q) memcpy(&v, q, 8); return q != nullptr; };
    const uint8_t *m = take(8);
    if (!m || memcmp(m, kChunkMagic, 8)) return bad("magic");
    uint32_t pv = 0, sv = 0, rounds = 0;
    uint64_t n = 0;
    if (!u32(pv) || !u32(sv) || !u64(n) || !u32(rounds)) return bad("truncated header");
    if (pv != 1 || sv != 1) return bad("version");
    if (!is_pow2(n) || n < 2 || n > (1ull << 40) || rounds < 1 || rounds > ilog2(n)) return bad("statement shape");
    const uint8_t *d0 = take(32), *d1 = take(32);
    if (!d0 || !d1) return bad("truncated digests");
    const uint32_t *lists[3];
    uint64_t counts[3];
    for (int k = 0; k < 3; ++k) {  // readWords: u64 count, then the words (canonical)
        if (!u64(counts[k]) || counts[k] > (len - at) / 4) return bad("word list");
        lists[k] = reinterpret_cast<const uint32_t *>(proof + at);
        std::vector<uint32_t> tmp(counts[k]);
        memcpy(tmp.data(), proof + at, 4 * counts[k]);
        for (uint32_t w : tmp)
            if (w >= kP) return bad("non-canonical word");
        at += 4 * counts[k];
    }
    if (at != len) return bad("trailing bytes");
    if (counts[0] != (n >> rounds) || counts[1] != coefficient_words(n, rounds) || counts[2] != rounds)
        return bad("proof shape");  // M31SumcheckProofV1.init
    try {
        std::vector<uint32_t> fin(counts[0]), coeffs(counts[1]), ch(counts[2]);
        memcpy(fin.data(), lists[0], 4 * counts[0]);
        memcpy(coeffs.data(), lists[1], 4 * counts[1]);
        memcpy(ch.data(), lists[2], 4 * counts[2]);
        uint8_t dg[32];
        host::sumcheck_words_digest(coeffs.data(), n, dg);
        bool ok = memcmp(dg, d0, 32) == 0;  // initialEvaluationDigestMatchesRevealedVector
        host::sumcheck_words_digest(fin.data(), fin.size(), dg);
        ok = ok && memcmp(dg, d1, 32) == 0;  // finalVectorDigestMatches
        host::Transcript t;
        const std::vector<uint8_t> hf = header_frame(n, rounds);
        t.absorb(hf.data(), hf.size());
        uint64_t off = 0, active = n;
        for (uint32_t r = 0; ok && r < rounds; ++r) {  // replayTranscriptAndFolds
            const uint32_t *v = coeffs.data() + off;
            const uint32_t c = round_challenge(t, r, v, active);
            ok = c == ch[r];
            const uint32_t *next = r + 1 == rounds ? fin.data() : coeffs.data() + off + active;
            for (uint64_t i = 0; ok && i < active / 2; ++i) ok = next[i] == host::fadd(v[2 * i], host::fmul(v[2 * i + 1], c));
            off += active;
            active >>= 1;
        }
        *accepted = ok ? 1 : 0;
    } catch (const std::bad_alloc &) {
        return fail(AZK_ERR_ALLOCATION, "sum-check chunk proof: out of host memory");
    }
    return AZK_OK;
}
