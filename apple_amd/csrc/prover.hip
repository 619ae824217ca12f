// Circle PCS/FRI commit superstep on one device, and its batched multi-job form.
//
// Reference: CircleCodewordPCSFRIProverV1 (Sources/AppleZKProver/CircleCodeword.swift:920-1419),
// CirclePCSFRIResidentProverV1 + CircleFRIResidentQueryExtractorV1
// (CircleFRIFoldPlan.swift:693-1231), transcript CircleFRITranscriptV1 and codec
// CirclePCSFRIProofCodecV1 (CircleProofFormat.swift:475-983).
//
// Flow of one prove (two captured graphs around one host step):
//   graph A: Circle FFT -> layer-0 leaf hashes + tile subtrees -> upper tree ->
//            per round r: transcript A(r) {root r, challenge frame r, squeeze 4}
//                         || transcript B(r+1) {commitment frame r+1} on a side stream,
//                         fold r fused with layer r+1 hashing/tiles -> upper tree
//            -> D2H of the transcript state and the final layer
//   host:    final-layer frame + final layer + query frame absorbed, Q indices
//            squeezed (the reference also derives these on the CPU after readback,
//            CircleFRIFoldPlan.swift:875-897; a single-wave device sponge costs ~9 us
//            per permutation, the 32 KiB final layer of the 2^21 shape is 241 of them)
//   graph B: H2D of the indices -> proof gather (roots, final layer, openings) -> D2H
// The transcript prefix (header, domain, security, public-input frames, digest,
// commitment frame 0) is constant per plan and absorbed once on the host at creation.
#include <thread>

#include "host_sha3.hpp"
#include "runtime_common.hpp"

using namespace azk;
using namespace azk::rt;

namespace {

void put32(std::vector<uint8_t> &v, uint32_t x) { for (int k = 0; k < 4; ++k) v.push_back((uint8_t)(x >> (8 * k))); }
void put64(std::vector<uint8_t> &v, uint64_t x) { for (int k = 0; k < 8; ++k) v.push_back((uint8_t)(x >> (8 * k))); }

// Circle transcript V1 base frame (CircleProofFormat.swift:966-973).
std::vector<uint8_t> base_frame(uint8_t type)
{
    static const char dom[] = "AppleZKProver.CircleFRI.PCS.V1";
    std::vector<uint8_t> f;
    put32(f, (uint32_t)(sizeof dom - 1));
    f.insert(f.end(), dom, dom + sizeof dom - 1);
    put32(f, 1);
    f.push_back(type);
    return f;
}

// 44-byte domain descriptor (CircleDomainDescriptorCodecV1.encode, CircleProofFormat.swift:75-93).
std::vector<uint8_t> descriptor(uint32_t log_n)
{
    std::vector<uint8_t> d = {0x41, 0x5a, 0x4b, 0x43, 0x44, 0x56, 0x31, 0x00};
    for (uint32_t w : {1u, kP, 31u, 2u, 1268011823u, log_n, 1u << (31 - (log_n + 1)), log_n - 1, 1u}) put32(d, w);
    return d;
}

struct PhaseClock {
    std::vector<hipEvent_t> ev;
    std::vector<int> tag;
    hipStream_t st;
    bool on;
    PhaseClock(hipStream_t s, bool enabled) : st(s), on(enabled) {}
    ~PhaseClock() { for (auto e : ev) (void)hipEventDestroy(e); }
    void mark(int t)
    {
        if (!on) return;
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return;
        (void)hipEventRecord(e, st);
        ev.push_back(e);
        tag.push_back(t);
    }
};
enum { PH_START = 0, PH_FFT, PH_COMMIT0, PH_FOLD, PH_TREE, PH_TRANSCRIPT, PH_TAIL, PH_HOST };

} // namespace

struct azk_circle_prover {
    azk_context *ctx = nullptr;
    azk_circle_prover_desc d{};
    uint32_t log_t = kDefaultLogTile;
    uint64_t n = 0, nfin = 0;
    DeviceMem fft_tw, fold_tw, arena, frames, ops, proof, coeff_staging;
    PinnedMem host_proof, host_tail, host_qidx;
    std::vector<uint4 *> layers;    // rounds + 1 layers (layer 0 = codeword, layer R = final)
    std::vector<uint32_t *> trees;  // one full tree per committed layer
    uint32_t *challenges = nullptr, *qidx = nullptr;
    uint64_t *tstate = nullptr, *tstate_init = nullptr;
    std::vector<std::pair<uint32_t, uint32_t>> prog_a, prog_b;  // (first op, count) per round
    std::vector<uint8_t> final_frame, query_frame;
    uint64_t roots_off = 0, final_off = 0, queries_off = 0, query_words = 0, proof_words = 0;
    std::vector<uint32_t> layer_off;
    hipStream_t own = nullptr, side = nullptr;
    std::vector<hipEvent_t> ev_a, ev_b;
    bool use_graph = true, timing = false;
    hipGraphExec_t graph_a = nullptr, graph_b = nullptr;
    const void *graph_src = nullptr;
    int graph_mode = -1;
    azk_phase_times phases{};
    std::mutex mu;
    ~azk_circle_prover()
    {
        if (graph_a) (void)hipGraphExecDestroy(graph_a);
        if (graph_b) (void)hipGraphExecDestroy(graph_b);
        for (auto e : ev_a) (void)hipEventDestroy(e);
        for (auto e : ev_b) (void)hipEventDestroy(e);
        if (own) (void)hipStreamDestroy(own);
        if (side) (void)hipStreamDestroy(side);
    }
};

namespace {

azk_status prover_build(azk_circle_prover *p)
{
    const azk_circle_prover_desc &d = p->d;
    const uint64_t n = p->n, nfin = p->nfin;
    azk_context *ctx = p->ctx;
    azk_status s = materialize_twiddles(ctx, d.log_n, d.rounds, p->fft_tw, &p->fold_tw);
    if (s) return s;
    HIP_TRY(hipStreamCreateWithFlags(&p->own, hipStreamNonBlocking));
    HIP_TRY(hipStreamCreateWithFlags(&p->side, hipStreamNonBlocking));
    p->ev_a.resize(d.rounds);
    p->ev_b.resize(d.rounds);
    for (uint32_t r = 0; r < d.rounds; ++r) {
        HIP_TRY(hipEventCreateWithFlags(&p->ev_a[r], hipEventDisableTiming));
        HIP_TRY(hipEventCreateWithFlags(&p->ev_b[r], hipEventDisableTiming));
    }

    // ---- arena: layers, full trees, transcript scratch
    std::vector<size_t> layer_at(d.rounds + 1), tree_at(d.rounds);
    size_t at = 0;
    for (uint32_t r = 0; r <= d.rounds; ++r) {
        layer_at[r] = at;
        at = align_up(at + ((size_t)16 << (d.log_n - r)), 256);
    }
    for (uint32_t r = 0; r < d.rounds; ++r) {
        tree_at[r] = at;
        at = align_up(at + (((size_t)2 << (d.log_n - r)) - 1) * 32, 256);
    }
    const size_t ch_at = at;
    at = align_up(at + 16 * (size_t)d.rounds, 256);
    const size_t q_at = at;
    at = align_up(at + 4 * (size_t)d.query_count, 256);
    const size_t st_at = at;
    at = align_up(at + 200, 256);
    const size_t st0_at = at;
    at = align_up(at + 200, 256);
    s = p->arena.ensure(at);
    if (s) return s;
    uint8_t *base = (uint8_t *)p->arena.p;
    p->layers.resize(d.rounds + 1);
    p->trees.resize(d.rounds);
    for (uint32_t r = 0; r <= d.rounds; ++r) p->layers[r] = (uint4 *)(base + layer_at[r]);
    for (uint32_t r = 0; r < d.rounds; ++r) p->trees[r] = (uint32_t *)(base + tree_at[r]);
    p->challenges = (uint32_t *)(base + ch_at);
    p->qidx = (uint32_t *)(base + q_at);
    p->tstate = (uint64_t *)(base + st_at);
    p->tstate_init = (uint64_t *)(base + st0_at);

    // ---- constant transcript prefix (CircleProofFormat.swift:796-818) + commitment frame 0
    std::vector<uint8_t> f;
    host::Transcript t;
    f = base_frame(0);
    for (uint32_t w : {1u, 1u, kP, 2u, 1268011823u}) put32(f, w);
    t.absorb(f.data(), f.size());
    auto desc = descriptor(d.log_n);
    t.absorb(desc.data(), desc.size());
    f = base_frame(1);
    for (uint32_t w : {d.log_blowup, d.query_count, d.folding_step, d.grinding_bits, d.log_blowup * d.query_count + d.grinding_bits})
        put32(f, w);
    t.absorb(f.data(), f.size());
    f = base_frame(2);
    put32(f, 32);
    t.absorb(f.data(), f.size());
    t.absorb(d.public_input_digest, 32);
    auto commitment_frame = [](uint32_t r) {
        std::vector<uint8_t> c = base_frame(3);
        put32(c, r);
        put32(c, 32);
        return c;
    };
    f = commitment_frame(0);
    t.absorb(f.data(), f.size());
    HIP_TRY(hipMemcpy(p->tstate_init, t.s, 200, hipMemcpyHostToDevice));

    // ---- per-round frames (8-byte aligned, zero padded) and device programs
    std::vector<uint8_t> blob;
    auto add = [&](const std::vector<uint8_t> &b) {
        size_t o = blob.size();
        blob.insert(blob.end(), b.begin(), b.end());
        blob.resize(align_up(blob.size(), 8) + 8, 0);
        return std::make_pair(o, b.size());
    };
    std::vector<std::pair<size_t, size_t>> cfr(d.rounds), chf(d.rounds);
    for (uint32_t r = 0; r < d.rounds; ++r) {
        cfr[r] = add(commitment_frame(r));
        f = base_frame(4);
        put32(f, r);
        put32(f, kP);
        put32(f, 4);
        chf[r] = add(f);
    }
    s = p->frames.ensure(blob.size());
    if (s) return s;
    HIP_TRY(hipMemcpy(p->frames.p, blob.data(), blob.size(), hipMemcpyHostToDevice));
    const uint8_t *fb = (const uint8_t *)p->frames.p;

    std::vector<TranscriptOp> ops;
    for (uint32_t r = 0; r < d.rounds; ++r) {  // A(r): root, challenge frame, squeeze 4 (:819-828)
        const uint32_t first = (uint32_t)ops.size();
        if (r == 0) ops.push_back({TOP_LOAD_STATE, 0, 0, 0, 200, p->tstate_init, nullptr});
        const uint64_t m = n >> r;
        ops.push_back({TOP_ABSORB, 0x06, 0, 0, 32, p->trees[r] + 8 * (2 * m - 2), nullptr});
        ops.push_back({TOP_ABSORB, 0x06, 0, 0, chf[r].second, fb + chf[r].first, nullptr});
        ops.push_back({TOP_SQUEEZE_U32, 0, kP, 0, 4, nullptr, p->challenges + 4 * r});
        p->prog_a.push_back({first, (uint32_t)ops.size() - first});
    }
    for (uint32_t r = 0; r < d.rounds; ++r) {  // B(r): commitment frame r (r >= 1)
        const uint32_t first = (uint32_t)ops.size();
        if (r > 0) ops.push_back({TOP_ABSORB, 0x06, 0, 0, cfr[r].second, fb + cfr[r].first, nullptr});
        p->prog_b.push_back({first, (uint32_t)ops.size() - first});
    }
    s = p->ops.ensure(ops.size() * sizeof(TranscriptOp));
    if (s) return s;
    HIP_TRY(hipMemcpy(p->ops.p, ops.data(), ops.size() * sizeof(TranscriptOp), hipMemcpyHostToDevice));

    // host-side final frames (:852-859, :708-717)
    p->final_frame = base_frame(5);
    put64(p->final_frame, nfin);
    put64(p->final_frame, nfin * 16);
    p->query_frame = base_frame(6);
    put32(p->query_frame, d.query_count);
    put64(p->query_frame, n / 2);

    // ---- proof layout, all fields 4-byte multiples (CircleProofFormat.swift:478-535)
    std::vector<uint8_t> head = {0x41, 0x5a, 0x4b, 0x50, 0x46, 0x56, 0x31, 0x00};
    put32(head, 1);
    put32(head, 1);
    put32(head, 44);
    head.insert(head.end(), desc.begin(), desc.end());
    for (uint32_t w : {d.log_blowup, d.query_count, d.folding_step, d.grinding_bits}) put32(head, w);
    put32(head, 32);
    head.insert(head.end(), d.public_input_digest, d.public_input_digest + 32);
    put32(head, d.rounds);
    p->roots_off = head.size() / 4;
    const uint64_t nfin_at = p->roots_off + 8 * (uint64_t)d.rounds;
    p->final_off = nfin_at + 2;
    const uint64_t q_count_at = p->final_off + 4 * nfin;
    p->queries_off = q_count_at + 1;
    p->layer_off.assign(d.rounds, 0);
    uint64_t qw = 3;
    for (uint32_t r = 0; r < d.rounds; ++r) {
        p->layer_off[r] = (uint32_t)qw;
        qw += 3 + 2 * (7 + 8 * (uint64_t)(d.log_n - r));
    }
    p->query_words = qw;
    p->proof_words = p->queries_off + (uint64_t)d.query_count * qw;
    std::vector<uint32_t> tmpl(p->proof_words, 0);
    memcpy(tmpl.data(), head.data(), head.size());
    tmpl[nfin_at] = (uint32_t)nfin;
    tmpl[nfin_at + 1] = (uint32_t)(nfin >> 32);
    tmpl[q_count_at] = d.query_count;
    s = p->proof.ensure(4 * p->proof_words);
    if (!s) s = p->host_proof.ensure(4 * p->proof_words);
    if (!s) s = p->host_tail.ensure(200 + 16 * nfin);
    if (!s) s = p->host_qidx.ensure(4 * (size_t)d.query_count);
    if (s) return s;
    HIP_TRY(hipMemcpy(p->proof.p, tmpl.data(), 4 * p->proof_words, hipMemcpyHostToDevice));
    return AZK_OK;
}

EmitParams emit_params(azk_circle_prover *p, const uint4 *layer0)
{
    EmitParams e{};
    e.proof = (uint32_t *)p->proof.p;
    e.roots_off = p->roots_off;
    e.final_off = p->final_off;
    e.queries_off = p->queries_off;
    e.query_words = p->query_words;
    e.rounds = p->d.rounds;
    e.log_n = p->d.log_n;
    e.final_count = p->nfin;
    for (uint32_t r = 0; r < p->d.rounds; ++r) {
        e.layer_off[r] = p->layer_off[r];
        e.roots[r] = p->trees[r] + 8 * (2 * (p->n >> r) - 2);
        e.trees[r] = p->trees[r];
        e.layers[r] = r == 0 ? layer0 : p->layers[r];
    }
    e.final_layer = (const uint32_t *)p->layers[p->d.rounds];
    e.query_pairs = p->qidx;
    return e;
}

const uint4 *layer0_of(azk_circle_prover *p, const void *src, int mode)
{
    return mode == 0 ? p->layers[0] : (const uint4 *)src;
}

// Graph A (see the file header).  mode 0: src = line-basis coefficients; mode 1: src = codeword.
azk_status enqueue_a(azk_circle_prover *p, hipStream_t st, hipStream_t side, const void *src, int mode, PhaseClock &clk)
{
    const azk_circle_prover_desc &d = p->d;
    const uint32_t lt0 = std::min(p->log_t, d.log_n);
    const uint4 *layer0 = layer0_of(p, src, mode);
    const TranscriptOp *ops = (const TranscriptOp *)p->ops.p;
    clk.mark(PH_START);
    if (mode == 0) {
        const void *cur = src;
        for (const FftPass &fp : fft_passes(d.log_n, p->log_t)) {
            launch_fft_pass(st, cur, p->layers[0], (const uint32_t *)p->fft_tw.p, d.log_n, fp.lo, fp.k, fp.m);
            cur = p->layers[0];
        }
        clk.mark(PH_FFT);
        launch_fft_final(st, cur, p->layers[0], (const uint32_t *)p->fft_tw.p, d.log_n, lt0, true, p->trees[0]);
    } else {
        clk.mark(PH_FFT);
        launch_fft_final(st, src, nullptr, nullptr, d.log_n, lt0, false, p->trees[0]);
    }
    clk.mark(PH_COMMIT0);
    launch_tree_upper(st, p->trees[0], p->n, lt0, kUpperLogChunk);
    clk.mark(PH_TREE);
    uint64_t off = 0, m = p->n;
    for (uint32_t r = 0; r < d.rounds; ++r) {
        if (r > 0) HIP_TRY(hipStreamWaitEvent(st, p->ev_b[r], 0));
        launch_transcript(st, ops + p->prog_a[r].first, p->prog_a[r].second, p->tstate);
        clk.mark(PH_TRANSCRIPT);
        const bool commit_next = r + 1 < d.rounds;
        if (commit_next) {
            HIP_TRY(hipEventRecord(p->ev_a[r], st));
            HIP_TRY(hipStreamWaitEvent(side, p->ev_a[r], 0));
            launch_transcript(side, ops + p->prog_b[r + 1].first, p->prog_b[r + 1].second, p->tstate);
            HIP_TRY(hipEventRecord(p->ev_b[r + 1], side));
        }
        const uint64_t n_out = m / 2;
        const uint32_t lt = enqueue_fold(st, r == 0 ? (const void *)layer0 : (const void *)p->layers[r], p->layers[r + 1],
                                         (const uint32_t *)p->fold_tw.p + off, p->challenges + 4 * r, n_out, p->log_t,
                                         commit_next ? p->trees[r + 1] : nullptr);
        clk.mark(PH_FOLD);
        if (commit_next) {
            launch_tree_upper(st, p->trees[r + 1], n_out, lt, kUpperLogChunk);
            clk.mark(PH_TREE);
        }
        off += n_out;
        m = n_out;
    }
    uint8_t *tail = (uint8_t *)p->host_tail.p;
    HIP_TRY(hipMemcpyAsync(tail, p->tstate, 200, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(tail + 200, p->layers[d.rounds], 16 * p->nfin, hipMemcpyDeviceToHost, st));
    clk.mark(PH_TAIL);
    return AZK_OK;
}

// Host step: final-layer frame + final layer + query frame, then Q indices mod N/2
// (CircleFRITranscriptV1.derive, CircleProofFormat.swift:687-722; grinding is 0).
void host_final(azk_circle_prover *p)
{
    host::Transcript t;
    const uint8_t *tail = (const uint8_t *)p->host_tail.p;
    memcpy(t.s, tail, 200);
    t.absorb(p->final_frame.data(), p->final_frame.size());
    t.absorb(tail + 200, 16 * p->nfin);
    t.absorb(p->query_frame.data(), p->query_frame.size());
    t.squeeze_u32(p->d.query_count, (uint32_t)(p->n / 2), (uint32_t *)p->host_qidx.p);
}

// Graph B: indices up, proof gather, proof down.
azk_status enqueue_b(azk_circle_prover *p, hipStream_t st, const void *src, int mode)
{
    HIP_TRY(hipMemcpyAsync(p->qidx, p->host_qidx.p, 4 * (size_t)p->d.query_count, hipMemcpyHostToDevice, st));
    launch_emit_proof(st, emit_params(p, layer0_of(p, src, mode)), p->d.query_count);
    HIP_TRY(hipMemcpyAsync(p->host_proof.p, p->proof.p, 4 * p->proof_words, hipMemcpyDeviceToHost, st));
    return AZK_OK;
}

template <class F>
azk_status capture(hipStream_t st, hipGraphExec_t *out, F &&enqueue)
{
    hipGraph_t gr = nullptr;
    HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    azk_status s = enqueue();
    hipError_t e = hipStreamEndCapture(st, &gr);
    if (s) {
        if (gr) (void)hipGraphDestroy(gr);
        return s;
    }
    if (e != hipSuccess) return fail(AZK_ERR_COMMAND_FAILED, "hipStreamEndCapture: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(out, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    if (e != hipSuccess) return fail(AZK_ERR_COMMAND_FAILED, "hipGraphInstantiate: %s", hipGetErrorString(e));
    return AZK_OK;
}

azk_status ensure_graphs(azk_circle_prover *p, hipStream_t st, const void *src, int mode)
{
    if (!p->graph_a || p->graph_src != src || p->graph_mode != mode) {
        if (p->graph_a) (void)hipGraphExecDestroy(p->graph_a);
        if (p->graph_b) (void)hipGraphExecDestroy(p->graph_b);
        p->graph_a = p->graph_b = nullptr;
        PhaseClock none(st, false);
        azk_status s = capture(st, &p->graph_a, [&]() { return enqueue_a(p, st, p->side, src, mode, none); });
        if (!s) s = capture(st, &p->graph_b, [&]() { return enqueue_b(p, st, src, mode); });
        if (s) return s;
        p->graph_src = src;
        p->graph_mode = mode;
    }
    return AZK_OK;
}

azk_status prove_locked(azk_circle_prover *p, const void *src, int mode, uint8_t *proof, size_t cap, size_t *len,
                        azk_stats *stats)
{
    const size_t bytes = 4 * p->proof_words;
    if (cap < bytes || !proof) return fail(AZK_ERR_INVALID_LAYOUT, "proof capacity %zu < %zu", cap, bytes);
    hipStream_t st = p->ctx->stream;
    const bool graph = p->use_graph && !p->timing;
    if (graph) {
        azk_status s = ensure_graphs(p, st, src, mode);
        if (s) return s;
    }
    Timer t;
    azk_status s = t.init();
    if (s) return s;
    PhaseClock clk(st, p->timing);
    const uint64_t w0 = now_ns();
    HIP_TRY(hipEventRecord(t.a, st));
    if (graph) HIP_TRY(hipGraphLaunch(p->graph_a, st));
    else if ((s = enqueue_a(p, st, p->side, src, mode, clk))) return s;
    HIP_TRY(hipGetLastError());
    const uint64_t w1 = now_ns();
    HIP_TRY(hipStreamSynchronize(st));
    const uint64_t h0 = now_ns();
    host_final(p);
    const uint64_t h1 = now_ns();
    clk.mark(PH_HOST);
    if (graph) HIP_TRY(hipGraphLaunch(p->graph_b, st));
    else if ((s = enqueue_b(p, st, src, mode))) return s;
    HIP_TRY(hipGetLastError());
    clk.mark(PH_TAIL);
    HIP_TRY(hipEventRecord(t.b, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (stats) {
        stats->submit_ns = w1 - w0;
        stats->wall_ns = now_ns() - w0;
        stats->gpu_ns = t.elapsed_ns();
    }
    if (p->timing && clk.ev.size() > 1) {
        azk_phase_times ph{};
        uint64_t *slot[] = {nullptr, &ph.fft_ns, &ph.commit0_ns, &ph.fold_ns, &ph.tree_ns, &ph.transcript_ns, &ph.emit_ns,
                            nullptr};
        for (size_t i = 1; i < clk.ev.size(); ++i) {
            float ms = 0;
            (void)hipEventElapsedTime(&ms, clk.ev[i - 1], clk.ev[i]);
            if (slot[clk.tag[i]]) *slot[clk.tag[i]] += (uint64_t)(ms * 1e6);
        }
        ph.host_ns = h1 - h0;
        float ms = 0;
        (void)hipEventElapsedTime(&ms, clk.ev.front(), clk.ev.back());
        ph.total_ns = (uint64_t)(ms * 1e6);
        p->phases = ph;
    }
    memcpy(proof, p->host_proof.p, bytes);
    if (len) *len = bytes;
    return AZK_OK;
}

} // namespace

extern "C" azk_status azk_circle_prover_create(azk_context *ctx, const azk_circle_prover_desc *desc,
                                               azk_circle_prover **out)
{
    CHECK_ARG(ctx && desc && out);
    const azk_circle_prover_desc &d = *desc;
    // CirclePCSFRIProofV1.init / CircleFRISecurityParametersV1.init (CircleProofFormat.swift:129-158,355-393)
    if (d.log_n < 2 || d.log_n > 28 || d.rounds < 1 || d.rounds > d.log_n || d.rounds > (uint32_t)kMaxRounds)
        return fail(AZK_ERR_INVALID_LAYOUT, "log_n %u / rounds %u outside the supported shape", d.log_n, d.rounds);
    if (d.log_blowup == 0 || d.query_count == 0 || d.query_count > 65536 || d.folding_step != 1 ||
        (uint64_t)d.log_blowup * d.query_count + d.grinding_bits > 0xffffffffull)
        return fail(AZK_ERR_INVALID_LAYOUT, "invalid security parameters");
    if (d.grinding_bits != 0)
        return fail(AZK_ERR_INVALID_KERNEL_CONFIG, "grinding is not implemented on the device path (grinding_bits = %u)",
                    d.grinding_bits);
    auto *p = new azk_circle_prover();
    p->ctx = ctx;
    p->d = d;
    p->n = 1ull << d.log_n;
    p->nfin = p->n >> d.rounds;
    HIP_TRY(hipSetDevice(ctx->device));
    std::lock_guard<std::mutex> gc(ctx->mu);
    azk_status s = prover_build(p);
    if (s) {
        delete p;
        return s;
    }
    *out = p;
    return AZK_OK;
}

extern "C" void azk_circle_prover_destroy(azk_circle_prover *p)
{
    if (!p) return;
    (void)hipSetDevice(p->ctx->device);
    (void)hipStreamSynchronize(p->ctx->stream);
    delete p;
}

extern "C" size_t azk_circle_prover_proof_size(const azk_circle_prover *p) { return p ? 4 * p->proof_words : 0; }

extern "C" azk_status azk_circle_prover_set_graph(azk_circle_prover *p, int enable)
{
    CHECK_ARG(p);
    std::lock_guard<std::mutex> g(p->mu);
    p->use_graph = enable != 0;
    return AZK_OK;
}

extern "C" azk_status azk_circle_prover_set_timing(azk_circle_prover *p, int enable)
{
    CHECK_ARG(p);
    std::lock_guard<std::mutex> g(p->mu);
    p->timing = enable != 0;
    return AZK_OK;
}

extern "C" azk_status azk_circle_prover_phase_times(const azk_circle_prover *p, azk_phase_times *out)
{
    CHECK_ARG(p && out);
    *out = p->phases;
    return AZK_OK;
}

extern "C" azk_status azk_circle_prove_fft_coefficients_resident(azk_circle_prover *p, const azk_buffer *coeffs,
                                                                 size_t off, uint8_t *proof, size_t cap, size_t *len,
                                                                 azk_stats *stats)
{
    CHECK_ARG(p && coeffs);
    std::lock_guard<std::mutex> g(p->mu);
    azk_status s = check_range(coeffs, off, (size_t)16 << p->d.log_n);
    if (s) return s;
    if (((uintptr_t)coeffs->ptr + off) % 16) return fail(AZK_ERR_INVALID_LAYOUT, "coefficients must be 16-byte aligned");
    HIP_TRY(hipSetDevice(p->ctx->device));
    std::lock_guard<std::mutex> gc(p->ctx->mu);
    return prove_locked(p, (uint8_t *)coeffs->ptr + off, 0, proof, cap, len, stats);
}

extern "C" azk_status azk_circle_prove_evaluations_resident(azk_circle_prover *p, const azk_buffer *evals, size_t off,
                                                            uint8_t *proof, size_t cap, size_t *len, azk_stats *stats)
{
    CHECK_ARG(p && evals);
    std::lock_guard<std::mutex> g(p->mu);
    azk_status s = check_range(evals, off, (size_t)16 << p->d.log_n);
    if (s) return s;
    if (((uintptr_t)evals->ptr + off) % 16) return fail(AZK_ERR_INVALID_LAYOUT, "evaluations must be 16-byte aligned");
    HIP_TRY(hipSetDevice(p->ctx->device));
    std::lock_guard<std::mutex> gc(p->ctx->mu);
    return prove_locked(p, (uint8_t *)evals->ptr + off, 1, proof, cap, len, stats);
}

extern "C" azk_status azk_circle_prove_fft_coefficients(azk_circle_prover *p, const void *coeffs, size_t bytes,
                                                        uint8_t *proof, size_t cap, size_t *len, azk_stats *stats)
{
    CHECK_ARG(p && coeffs);
    std::lock_guard<std::mutex> g(p->mu);
    const size_t need = (size_t)16 << p->d.log_n;
    if (bytes < need) return fail(AZK_ERR_INVALID_LAYOUT, "coefficient bytes %zu < %zu", bytes, need);
    HIP_TRY(hipSetDevice(p->ctx->device));
    azk_status s = p->coeff_staging.ensure(need);
    if (s) return s;
    std::lock_guard<std::mutex> gc(p->ctx->mu);
    HIP_TRY(hipMemcpyAsync(p->coeff_staging.p, coeffs, need, hipMemcpyHostToDevice, p->ctx->stream));
    return prove_locked(p, p->coeff_staging.p, 0, proof, cap, len, stats);
}

extern "C" azk_status azk_circle_prover_clear(azk_circle_prover *p)
{
    CHECK_ARG(p);
    std::lock_guard<std::mutex> g(p->mu);
    HIP_TRY(hipSetDevice(p->ctx->device));
    // zero every buffer that held witness-derived data; keep the constant parts
    std::vector<uint8_t> st0(200);
    HIP_TRY(hipMemcpy(st0.data(), p->tstate_init, 200, hipMemcpyDeviceToHost));
    std::vector<uint32_t> tmpl(p->proof_words);
    HIP_TRY(hipMemcpy(tmpl.data(), p->proof.p, 4 * p->proof_words, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemset(p->arena.p, 0, p->arena.n));
    HIP_TRY(hipMemcpy(p->tstate_init, st0.data(), 200, hipMemcpyHostToDevice));
    if (p->coeff_staging.p) HIP_TRY(hipMemset(p->coeff_staging.p, 0, p->coeff_staging.n));
    memset(tmpl.data() + p->roots_off, 0, 32 * (size_t)p->d.rounds);
    memset(tmpl.data() + p->final_off, 0, 16 * p->nfin);
    memset(tmpl.data() + p->queries_off, 0, 4 * (size_t)p->d.query_count * p->query_words);
    HIP_TRY(hipMemcpy(p->proof.p, tmpl.data(), 4 * p->proof_words, hipMemcpyHostToDevice));
    memset(p->host_proof.p, 0, p->host_proof.n);
    memset(p->host_tail.p, 0, p->host_tail.n);
    memset(p->host_qidx.p, 0, p->host_qidx.n);
    return AZK_OK;
}

// ====================================================================== batched form

struct azk_circle_batch {
    azk_context *ctx = nullptr;
    std::vector<azk_circle_prover *> provers;
    std::vector<const void *> srcs;
    hipGraphExec_t graph_a = nullptr, graph_b = nullptr;
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> done;
    ~azk_circle_batch()
    {
        if (graph_a) (void)hipGraphExecDestroy(graph_a);
        if (graph_b) (void)hipGraphExecDestroy(graph_b);
        if (fork) (void)hipEventDestroy(fork);
        for (auto e : done) (void)hipEventDestroy(e);
    }
};

extern "C" azk_status azk_circle_batch_create(azk_circle_prover *const *provers, const azk_buffer *const *coeffs,
                                              const size_t *offsets, uint32_t jobs, azk_circle_batch **out)
{
    CHECK_ARG(provers && coeffs && offsets && out && jobs);
    azk_context *ctx = provers[0]->ctx;
    for (uint32_t j = 0; j < jobs; ++j) {
        CHECK_ARG(provers[j] && coeffs[j]);
        if (provers[j]->ctx != ctx) return fail(AZK_ERR_INVALID_ARGUMENT, "all provers must share one context");
        for (uint32_t k = 0; k < j; ++k)
            if (provers[k] == provers[j]) return fail(AZK_ERR_INVALID_ARGUMENT, "a prover appears twice in one batch");
        azk_status s = check_range(coeffs[j], offsets[j], (size_t)16 << provers[j]->d.log_n);
        if (s) return s;
        if (((uintptr_t)coeffs[j]->ptr + offsets[j]) % 16) return fail(AZK_ERR_INVALID_LAYOUT, "job %u misaligned", j);
    }
    auto *b = new azk_circle_batch();
    b->ctx = ctx;
    b->provers.assign(provers, provers + jobs);
    for (uint32_t j = 0; j < jobs; ++j) b->srcs.push_back((const uint8_t *)coeffs[j]->ptr + offsets[j]);
    HIP_TRY(hipSetDevice(ctx->device));
    std::lock_guard<std::mutex> gc(ctx->mu);
    HIP_TRY(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
    b->done.resize(jobs);
    for (auto &e : b->done) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    hipStream_t st = ctx->stream;
    auto fork_join = [&](bool phase_a) -> azk_status {
        HIP_TRY(hipEventRecord(b->fork, st));
        for (uint32_t j = 0; j < jobs; ++j) {
            azk_circle_prover *p = b->provers[j];
            HIP_TRY(hipStreamWaitEvent(p->own, b->fork, 0));
            PhaseClock none(p->own, false);
            azk_status s = phase_a ? enqueue_a(p, p->own, p->side, b->srcs[j], 0, none) : enqueue_b(p, p->own, b->srcs[j], 0);
            if (s) return s;
            HIP_TRY(hipEventRecord(b->done[j], p->own));
            HIP_TRY(hipStreamWaitEvent(st, b->done[j], 0));
        }
        return AZK_OK;
    };
    azk_status s = capture(st, &b->graph_a, [&]() { return fork_join(true); });
    if (!s) s = capture(st, &b->graph_b, [&]() { return fork_join(false); });
    if (s) {
        delete b;
        return s;
    }
    *out = b;
    return AZK_OK;
}

extern "C" void azk_circle_batch_destroy(azk_circle_batch *b)
{
    if (!b) return;
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    delete b;
}

extern "C" azk_status azk_circle_batch_prove(azk_circle_batch *b, uint8_t *const *proofs, const size_t *caps,
                                             size_t *lens, azk_stats *stats)
{
    CHECK_ARG(b && proofs && caps);
    const uint32_t jobs = (uint32_t)b->provers.size();
    for (uint32_t j = 0; j < jobs; ++j) {
        CHECK_ARG(proofs[j]);
        if (caps[j] < 4 * b->provers[j]->proof_words) return fail(AZK_ERR_INVALID_LAYOUT, "proof capacity of job %u", j);
    }
    std::vector<std::unique_lock<std::mutex>> locks;
    for (auto *p : b->provers) locks.emplace_back(p->mu);
    HIP_TRY(hipSetDevice(b->ctx->device));
    std::lock_guard<std::mutex> gc(b->ctx->mu);
    hipStream_t st = b->ctx->stream;
    Timer t;
    azk_status s = t.init();
    if (s) return s;
    const uint64_t w0 = now_ns();
    HIP_TRY(hipEventRecord(t.a, st));
    HIP_TRY(hipGraphLaunch(b->graph_a, st));
    const uint64_t w1 = now_ns();
    HIP_TRY(hipStreamSynchronize(st));
    // host finals in parallel (each is a serial sponge over one job's final layer)
    const uint32_t hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const uint32_t nt = std::min(jobs, hw);
    std::vector<std::thread> pool;
    for (uint32_t k = 1; k < nt; ++k)
        pool.emplace_back([&, k]() {
            for (uint32_t j = k; j < jobs; j += nt) host_final(b->provers[j]);
        });
    for (uint32_t j = 0; j < jobs; j += nt) host_final(b->provers[j]);
    for (auto &th : pool) th.join();
    HIP_TRY(hipGraphLaunch(b->graph_b, st));
    HIP_TRY(hipEventRecord(t.b, st));
    HIP_TRY(hipStreamSynchronize(st));
    for (uint32_t j = 0; j < jobs; ++j) {
        azk_circle_prover *p = b->provers[j];
        memcpy(proofs[j], p->host_proof.p, 4 * p->proof_words);
        if (lens) lens[j] = 4 * p->proof_words;
    }
    if (stats) {
        stats->submit_ns = w1 - w0;
        stats->wall_ns = now_ns() - w0;
        stats->gpu_ns = t.elapsed_ns();
    }
    return AZK_OK;
}

extern "C" azk_status azk_circle_prove_batch(azk_circle_prover *const *provers, const azk_buffer *const *coeffs,
                                             const size_t *offsets, uint32_t jobs, uint8_t *const *proofs,
                                             const size_t *caps, size_t *lens, azk_stats *stats)
{
    if (!jobs) return AZK_OK;
    azk_circle_batch *b = nullptr;
    azk_status s = azk_circle_batch_create(provers, coeffs, offsets, jobs, &b);
    if (s) return s;
    s = azk_circle_batch_prove(b, proofs, caps, lens, stats);
    azk_circle_batch_destroy(b);
    return s;
}
