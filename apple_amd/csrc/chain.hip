// Merkle-transcript FRI fold chain: per round commit the layer (SHA3 Merkle over 16-byte
// QM31 leaves), absorb the root into the Circle FRI transcript, squeeze the challenge and
// fold; output the roots, the challenges and the final layer.
//
// Reference: CircleFRIMerkleTranscriptFoldChainPlan (Sources/AppleZKProver/
// CircleFRIFoldPlan.swift:301-649; execute :390-408, executeVerified :410-431,
// executeResident :433-455) over QM31FRIFoldChainPlan.executeMerkleTranscriptDerivedLocked
// (QM31FRIFold.swift:2203-2531).  On this device it is the prover's program A in evaluation
// mode (first-layer commitment, device transcript rounds, fused folds and commitments),
// stopped before the host step; the verified form checks it against host code.
#include <vector>

#include "circle_frames.hpp"
#include "host_gates.hpp"
#include "host_sha3.hpp"
#include "prover_internal.hpp"

using namespace azk;
using namespace azk::rt;
using namespace azk::prv;

struct azk_chain_plan {
    azk_context *ctx = nullptr;
    azk_circle_prover *p = nullptr;
    DeviceMem staging;  // host-variant evaluations
    std::mutex mu;
};

namespace {

// commitAndFoldOracle (CircleFRIFoldPlan.swift:593-647) on the host: host Merkle roots,
// host transcript (prefix frames, then per round commitment frame, root, challenge frame,
// squeeze 4 mod p) and host folds.
bool host_chain(const azk_circle_prover_desc &d, const uint32_t *evals, uint8_t *roots, uint32_t *challenges,
                std::vector<uint32_t> &final_layer)
{
    host::Transcript t;
    for (const auto &m : frames::prefix_messages(d.log_n, d.log_blowup, d.query_count, d.folding_step, d.grinding_bits,
                                                 d.public_input_digest))
        t.absorb(m.data(), m.size());
    const std::vector<host::Pt> pts = host::Domain::canonical_of(d.log_n).storage_points();
    std::vector<uint32_t> cur(evals, evals + ((size_t)4 << d.log_n)), next;
    for (uint32_t r = 0; r < d.rounds; ++r) {
        host::qm31_merkle_root(cur.data(), cur.size() / 4, roots + 32 * r);
        const auto cf = frames::commitment_frame(r), xf = frames::challenge_frame(r);
        t.absorb(cf.data(), cf.size());
        t.absorb(roots + 32 * r, 32);
        t.absorb(xf.data(), xf.size());
        t.squeeze_u32(4, kP, challenges + 4 * r);
        next.assign(cur.size() / 2, 0);
        if (!host::fold_round(pts, r, cur.data(), cur.size() / 8, challenges + 4 * r, next.data())) return false;
        cur.swap(next);
    }
    final_layer.swap(cur);
    return true;
}

} // namespace

extern "C" {

azk_status azk_chain_plan_create(azk_context *ctx, const azk_circle_prover_desc *desc, azk_chain_plan **out)
{
    CHECK_ARG(ctx && desc && out);
    azk_circle_prover *p = nullptr;
    azk_status s = azk_circle_prover_create(ctx, desc, &p);
    if (s) return s;
    auto *c = new azk_chain_plan{};
    c->ctx = ctx;
    c->p = p;
    *out = c;
    return AZK_OK;
}

void azk_chain_plan_destroy(azk_chain_plan *c)
{
    if (!c) return;
    azk_circle_prover *p = c->p;
    (void)hipSetDevice(c->ctx->device);
    delete c;                      // frees the staging while the prover still holds the context
    azk_circle_prover_destroy(p);  // drops the context reference
}

// executeResident: evaluations (2^log_n QM31) in HBM; roots and challenges to host memory,
// the final layer (2^(log_n - rounds) QM31) into `out`.
azk_status azk_chain_execute_resident(azk_chain_plan *c, const azk_buffer *evals, size_t evals_offset,
                                      uint8_t *commitments, uint32_t *challenges, azk_buffer *out, size_t out_offset,
                                      azk_stats *stats)
{
    CHECK_ARG(c && evals && commitments && challenges && out);
    std::lock_guard<std::mutex> g(c->mu);
    azk_circle_prover *p = c->p;
    const size_t ib = (size_t)16 << p->d.log_n, ob = 16 * p->nfin;
    azk_status s = check_range(evals, evals_offset, ib);
    if (!s) s = check_range(out, out_offset, ob);
    if (s) return s;
    if (overlaps(evals, evals_offset, ib, out, out_offset, ob))
        return fail(AZK_ERR_INVALID_LAYOUT, "evaluation and output ranges alias");
    if (((uintptr_t)evals->ptr + evals_offset) % 16 || ((uintptr_t)out->ptr + out_offset) % 16)
        return fail(AZK_ERR_INVALID_LAYOUT, "QM31 buffers must be 16-byte aligned");
    HIP_TRY(hipSetDevice(c->ctx->device));
    std::lock_guard<std::mutex> gp(p->mu);
    std::lock_guard<std::mutex> gc(c->ctx->mu);
    return chain_single(p, (uint8_t *)evals->ptr + evals_offset, 1, commitments, challenges,
                        (uint8_t *)out->ptr + out_offset, hipMemcpyDeviceToDevice, stats);
}

// execute(evaluations:): host arrays in and out.
azk_status azk_chain_execute(azk_chain_plan *c, const uint32_t *evals, size_t eval_words, uint8_t *commitments,
                             uint32_t *challenges, uint32_t *final_layer, size_t final_words, azk_stats *stats)
{
    CHECK_ARG(c && evals && commitments && challenges && final_layer);
    std::lock_guard<std::mutex> g(c->mu);
    azk_circle_prover *p = c->p;
    const size_t iw = (size_t)4 << p->d.log_n, ow = 4 * p->nfin;
    if (eval_words < iw || final_words < ow) return fail(AZK_ERR_INVALID_LAYOUT, "host arrays shorter than the layers");
    for (size_t i = 0; i < iw; ++i)
        if (evals[i] >= kP) return fail(AZK_ERR_INVALID_LAYOUT, "non-canonical evaluation");
    HIP_TRY(hipSetDevice(c->ctx->device));
    azk_status s = c->staging.ensure(4 * iw);
    if (s) return s;
    HIP_TRY(hipMemcpy(c->staging.p, evals, 4 * iw, hipMemcpyHostToDevice));
    std::lock_guard<std::mutex> gp(p->mu);
    std::lock_guard<std::mutex> gc(c->ctx->mu);
    return chain_single(p, c->staging.p, 1, commitments, challenges, final_layer, hipMemcpyDeviceToHost, stats);
}

// executeVerified: AZK_ERR_CORRECTNESS unless roots, challenges and final layer equal the
// host computation.
azk_status azk_chain_execute_verified(azk_chain_plan *c, const uint32_t *evals, size_t eval_words,
                                      uint8_t *commitments, uint32_t *challenges, uint32_t *final_layer,
                                      size_t final_words, azk_stats *stats)
{
    azk_status s = azk_chain_execute(c, evals, eval_words, commitments, challenges, final_layer, final_words, stats);
    if (s) return s;
    const azk_circle_prover_desc &d = c->p->d;
    std::vector<uint8_t> roots(32 * (size_t)d.rounds);
    std::vector<uint32_t> ch(4 * (size_t)d.rounds), fin;
    if (!host_chain(d, evals, roots.data(), ch.data(), fin))
        return fail(AZK_ERR_INVALID_LAYOUT, "zero domain coordinate");
    if (memcmp(roots.data(), commitments, roots.size()) || memcmp(ch.data(), challenges, 4 * ch.size()) ||
        memcmp(fin.data(), final_layer, 4 * fin.size()))
        return fail(AZK_ERR_CORRECTNESS, "Merkle-transcript fold chain device result did not match the CPU gate");
    return AZK_OK;
}

} // extern "C"
