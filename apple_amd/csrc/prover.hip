// Circle PCS/FRI commit superstep, for one job or a batch of uneven jobs.
//
// Reference: CircleCodewordPCSFRIProverV1 (Sources/AppleZKProver/CircleCodeword.swift:920-1419),
// CirclePCSFRIResidentProverV1 + CircleFRIResidentQueryExtractorV1
// (CircleFRIFoldPlan.swift:693-1231), transcript CircleFRITranscriptV1 and codec
// CirclePCSFRIProofCodecV1 (CircleProofFormat.swift:475-983).
//
// A Runner turns a list of jobs into a fixed program of table-driven launches, where
// each launch covers that phase for every job (so small jobs' latency-bound tails
// overlap), captured as two graphs around one host step:
//   graph A: Circle FFT passes -> final pass fused with layer-0 leaf hashes + tile
//            subtrees -> upper trees -> per round r: transcript {root r, challenge
//            frame r, squeeze 4, commitment frame r+1}; fold r fused with layer r+1
//            hashing/tiles; upper trees -> gather of every job's transcript state +
//            final layer -> one D2H copy
//   host:    per job, final-layer frame + final layer + query frame absorbed and the
//            Q indices squeezed (the reference also derives these on the CPU after
//            readback, CircleFRIFoldPlan.swift:875-897; one wave runs a serial sponge
//            at ~9 us per permutation, the 2^21 shape's final layer is 241 of them)
//   graph B: one H2D copy of all indices -> proof gather for every job -> one D2H copy
// The transcript prefix (header, domain, security, public-input frames, digest,
// commitment frame 0) is constant per prover and absorbed once on the host.
#include <functional>
#include <new>
#include <memory>
#include <thread>

#include "circle_frames.hpp"

#include "host_gates.hpp"
#include "host_sha3.hpp"
#include "prover_internal.hpp"

using namespace azk;
using namespace azk::rt;
using namespace azk::prv;

namespace {

using frames::base_frame;
using frames::descriptor;
using frames::put32;
using frames::put64;

} // namespace

struct azk_circle_batch;

namespace azk::prv {

azk_status prover_build(azk_circle_prover *p)
{
    const azk_circle_prover_desc &d = p->d;
    const uint64_t n = p->n, nfin = p->nfin;
    azk_status s = materialize_twiddles(p->ctx, d.log_n, d.rounds, p->fft_tw, &p->fold_tw);
    if (s) return s;

    // ---- arena: layers, full trees, transcript state
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

    // Round r: root r, challenge frame r, squeeze 4 (:819-828).  The next round's
    // commitment frame (constant bytes) is its own program, run beside fold r: it needs only
    // the state after challenge frame r (the squeeze read a copy of the state), and root
    // r + 1 is absorbed after the fold anyway, so it leaves the round's serial chain.
    std::vector<TranscriptOp> ops;
    p->prog_a.clear();
    p->prog_c.clear();
    for (uint32_t r = 0; r < d.rounds; ++r) {
        const uint32_t first = (uint32_t)ops.size();
        if (r == 0) ops.push_back({TOP_LOAD_STATE, 0, 0, 0, 200, p->tstate_init, nullptr});
        const uint64_t m = n >> r;
        ops.push_back({TOP_ABSORB, 0x06, 0, 0, 32, p->trees[r] + 8 * (2 * m - 2), nullptr});
        ops.push_back({TOP_ABSORB, 0x06, 0, 0, chf[r].second, fb + chf[r].first, nullptr});
        ops.push_back({TOP_SQUEEZE_U32, 0, kP, 0, 4, nullptr, p->challenges + 4 * r});
        p->prog_a.push_back({first, (uint32_t)ops.size() - first});
        if (r + 1 < d.rounds) {
            p->prog_c.push_back({(uint32_t)ops.size(), 1});
            ops.push_back({TOP_ABSORB, 0x06, 0, 0, cfr[r + 1].second, fb + cfr[r + 1].first, nullptr});
        }
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
    if (d.grinding_bits) {  // grindingFrame (:956-962)
        p->grind_frame = base_frame(7);
        put32(p->grind_frame, d.grinding_bits);
        put32(p->grind_frame, 8);
    }

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
    if (d.grinding_bits) {  // tail: u32 0 claimed openings, then the u64 nonce (:516-525)
        p->nonce_off = p->proof_words + 1;
        p->proof_words += 3;
    }
    p->proof_template.assign(p->proof_words, 0);
    memcpy(p->proof_template.data(), head.data(), head.size());
    p->proof_template[nfin_at] = (uint32_t)nfin;
    p->proof_template[nfin_at + 1] = (uint32_t)(nfin >> 32);
    p->proof_template[q_count_at] = d.query_count;
    return AZK_OK;
}

// Builds the launch program of a job list.
azk_status runner_build(Runner &R, azk_context *ctx, const std::vector<Job> &jobs)
{
    R.ctx = ctx;
    R.jobs = jobs;
    for (const Job &j : jobs) R.rounds = std::max(R.rounds, j.p->d.rounds);
    // a fold launch keeps about 2^log_leafwise_grid blocks over all of its jobs: J jobs take
    // tiles 2^ceil(log2 J) times larger (a 64-job batch measured 17% faster with 256-leaf
    // tiles than with the single-job sizes)
    uint32_t lg_jobs = 0;
    while ((1ull << lg_jobs) < jobs.size()) ++lg_jobs;
    R.tree_waves = 4;
    for (const Job &j : jobs) R.tree_waves = std::max(R.tree_waves, j.p->xp.tree_waves);
    R.tr_a.resize(R.rounds);
    R.fold_big.resize(R.rounds);
    R.fold_leaf.resize(R.rounds);
    R.tree_r.resize(R.rounds);
    size_t tail_bytes = 0, qidx_bytes = 0, proof_bytes = 0;
    for (const Job &j : jobs) {
        azk_circle_prover *p = j.p;
        const uint32_t R_j = p->d.rounds;
        const TranscriptOp *ops = (const TranscriptOp *)p->ops.p;
        const ExecPlan &xp = p->xp;
        if (j.mode == 2)  // M31 words through the passes (in layer 1, unused until the first fold)
            plan_fft(R.fft_pass_m31, R.fft_final_m31, j.src, p->layers[0], (const uint32_t *)p->fft_tw.p, p->d.log_n,
                     xp.log_tile_commit, true, p->trees[0], true, p->layers[1]);
        else
            plan_fft(R.fft_pass, R.fft_final, j.src, p->layers[0], (const uint32_t *)p->fft_tw.p, p->d.log_n,
                     xp.log_tile_commit, j.mode == 0, p->trees[0]);
        // round r's transcript program runs in the block that computes root r when the
        // root comes from an upper-tree step, and as its own launch otherwise
        auto prog = [&](uint32_t r) { return TranscriptDesc{ops + p->prog_a[r].first, p->tstate, p->prog_a[r].second, 0}; };
        {
            const TranscriptDesc t0 = prog(0);
            if (!plan_tree(R.tree0, p->trees[0], p->n, std::min(xp.log_tile_commit, p->d.log_n), xp.log_tree_chunk, &t0))
                R.tr_a[0].add(t0, 1);
        }
        uint64_t off = 0, m = p->n;
        for (uint32_t r = 0; r < R_j; ++r) {
            const uint64_t n_out = m / 2;
            const bool commit_next = r + 1 < R_j;
            if (commit_next) {  // commitment frame r + 1: one block, first in the fold launch of this round
                const bool leafwise = n_out < (1ull << xp.log_leafwise_below);
                FoldDesc t{};
                t.t_ops = ops + p->prog_c[r].first;
                t.t_state = p->tstate;
                t.t_nops = p->prog_c[r].second;
                (leafwise ? R.fold_leaf[r] : R.fold_big[r]).add(t, 1);
            }
            const uint32_t lt = plan_fold(R.fold_big[r], R.fold_leaf[r], r == 0 ? (const void *)j.layer0() : (const void *)p->layers[r],
                                          p->layers[r + 1], (const uint32_t *)p->fold_tw.p + off, p->challenges + 4 * r, n_out,
                                          xp.log_tile_fold, commit_next ? p->trees[r + 1] : nullptr,
                                          xp.log_leafwise_below, xp.log_tile_leafwise,
                                          xp.log_leafwise_grid ? std::max(1u, xp.log_leafwise_grid - std::min(xp.log_leafwise_grid, lg_jobs)) : 0);
            if (commit_next) {
                const TranscriptDesc tn = prog(r + 1);
                if (!plan_tree(R.tree_r[r], p->trees[r + 1], n_out, lt, xp.log_tree_chunk, &tn)) R.tr_a[r + 1].add(tn, 1);
            }
            off += n_out;
            m = n_out;
        }
        R.tail_off.push_back(tail_bytes);
        tail_bytes = align_up(tail_bytes + 200 + 16 * p->nfin, 256);
        R.qidx_off.push_back(qidx_bytes);
        qidx_bytes = align_up(qidx_bytes + 4 * (size_t)p->d.query_count, 256);
        R.proof_off.push_back(proof_bytes);
        proof_bytes = align_up(proof_bytes + 4 * p->proof_words, 256);
    }
    R.nonce.assign(jobs.size(), 0);
    // The tail (state + final layer), the query indices and the gathered proofs live in pinned
    // host memory that the kernels read and write directly: no copy steps around the host step.
    azk_status s = R.tail_host.ensure(tail_bytes);
    if (!s) s = R.qidx_host.ensure(qidx_bytes);
    if (!s) s = R.proof_host.ensure(proof_bytes);
    if (s) return s;
    for (size_t k = 0; k < jobs.size(); ++k) {
        azk_circle_prover *p = jobs[k].p;
        uint32_t *tail = (uint32_t *)((uint8_t *)R.tail_host.p + R.tail_off[k]);
        R.tail.add(CopyDesc{(const uint32_t *)p->tstate, tail, 50, 0, 0}, 1);
        const uint64_t fw = 4 * p->nfin;
        R.tail.add(CopyDesc{(const uint32_t *)p->layers[p->d.rounds], tail + 50, fw, 0, 0}, (uint32_t)((fw + 1023) / 1024));
        // proof words start from the constant template; the gather fills the rest
        uint32_t *proof = (uint32_t *)((uint8_t *)R.proof_host.p + R.proof_off[k]);
        memcpy(proof, p->proof_template.data(), 4 * p->proof_words);
        EmitParams e{};
        e.proof = proof;
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
            e.layers[r] = r == 0 ? jobs[k].layer0() : p->layers[r];
        }
        e.final_layer = (const uint32_t *)p->layers[p->d.rounds];
        e.query_pairs = (const uint32_t *)((uint8_t *)R.qidx_host.p + R.qidx_off[k]);
        e.queries = p->d.query_count;
        R.emit.add(e, 1 + p->d.query_count * p->d.rounds);
    }
    // pack every table into one device buffer
    for (auto &l : R.fft_pass) l.off = R.tables.put(l.v);
    R.fft_final.off = R.tables.put(R.fft_final.v);
    for (auto &l : R.fft_pass_m31) l.off = R.tables.put(l.v);
    R.fft_final_m31.off = R.tables.put(R.fft_final_m31.v);
    for (auto &l : R.tree0) l.off = R.tables.put(l.v);
    for (uint32_t r = 0; r < R.rounds; ++r) {
        R.tr_a[r].off = R.tables.put(R.tr_a[r].v);
        R.fold_big[r].off = R.tables.put(R.fold_big[r].v);
        R.fold_leaf[r].off = R.tables.put(R.fold_leaf[r].v);
        for (auto &l : R.tree_r[r]) l.off = R.tables.put(l.v);
    }
    R.tail.off = R.tables.put(R.tail.v);
    R.emit.off = R.tables.put(R.emit.v);
    return R.tables.upload();
}

template <class D>
const D *tab(const Runner &R, const Launch<D> &l) { return R.tables.at<D>(l.off); }

azk_status enqueue_a(Runner &R, hipStream_t st, PhaseClock &clk)
{
    clk.mark(PH_START);
    for (auto &l : R.fft_pass) launch_fft_pass(st, tab(R, l), l.n(), l.blocks, l.max_log, false);
    for (auto &l : R.fft_pass_m31) launch_fft_pass(st, tab(R, l), l.n(), l.blocks, l.max_log, true);
    clk.mark(PH_FFT);
    launch_fft_final(st, tab(R, R.fft_final), R.fft_final.n(), R.fft_final.blocks, R.fft_final.max_log,
                     any_tree(R.fft_final), false);
    launch_fft_final(st, tab(R, R.fft_final_m31), R.fft_final_m31.n(), R.fft_final_m31.blocks, R.fft_final_m31.max_log,
                     any_tree(R.fft_final_m31), true);
    clk.mark(PH_COMMIT0);
    for (auto &l : R.tree0) launch_tree(st, tab(R, l), l.n(), l.blocks, l.max_log, R.tree_waves, l.v.data());
    clk.mark(PH_TREE);
    for (uint32_t r = 0; r < R.rounds; ++r) {
        launch_transcript(st, tab(R, R.tr_a[r]), R.tr_a[r].n());
        clk.mark(PH_TRANSCRIPT);
        auto &fb = R.fold_big[r], &fl = R.fold_leaf[r];
        launch_fold(st, tab(R, fb), fb.n(), fb.blocks, fb.max_log, fb.v.data());
        launch_fold_leafwise(st, tab(R, fl), fl.n(), fl.blocks, (uint32_t)R.ctx->props.multiProcessorCount, fl.v.data());
        clk.mark(PH_FOLD);
        for (auto &l : R.tree_r[r]) launch_tree(st, tab(R, l), l.n(), l.blocks, l.max_log, R.tree_waves, l.v.data());
        clk.mark(PH_TREE);
    }
    launch_copy(st, tab(R, R.tail), R.tail.n(), R.tail.blocks);  // into pinned host memory
    clk.mark(PH_TAIL);
    return AZK_OK;
}

// Smallest nonce whose 32-byte squeeze, after absorbing the nonce into state `s`, starts
// with `bits` zero bits, most significant bit of each byte first
// (CircleFRITranscriptV1.findGrindingNonce, CircleProofFormat.swift:752-791; digestMeetsTarget
// :663-681).  Threads take blocks of 4096 nonces in increasing order and stop at the best hit,
// so every nonce below the result has been tested.
uint64_t find_nonce(const uint64_t s[25], uint32_t bits, unsigned threads)
{
    auto hit = [&](uint64_t nonce) {
        uint64_t st[25];
        memcpy(st, s, 200);
        st[0] ^= nonce;                     // 8 LE nonce bytes
        st[1] ^= 0x06;                      // SHA3 suffix at byte 8
        st[16] ^= 0x8000000000000000ull;    // final padding bit at byte 135
        host::keccak_f(st);
        uint32_t rem = bits, i = 0;
        for (; rem >= 8; rem -= 8, ++i)
            if ((st[i / 8] >> (8 * (i % 8))) & 0xff) return false;
        const uint32_t byte = (uint32_t)(st[i / 8] >> (8 * (i % 8))) & 0xff;
        return !rem || (byte & (0xffu << (8 - rem)) & 0xff) == 0;
    };
    constexpr uint64_t kBlock = 4096;
    std::atomic<uint64_t> best{~0ull}, next{0};
    auto worker = [&]() {
        for (;;) {
            const uint64_t lo = next.fetch_add(1) * kBlock;
            if (lo >= best.load()) return;
            for (uint64_t n = lo; n < lo + kBlock && n < best.load(std::memory_order_relaxed); ++n)
                if (hit(n)) {
                    uint64_t cur = best.load();
                    while (n < cur && !best.compare_exchange_weak(cur, n)) {}
                    break;
                }
        }
    };
    std::vector<std::thread> pool;
    for (unsigned t = 1; t < threads; ++t) pool.emplace_back(worker);
    worker();
    for (auto &th : pool) th.join();
    return best.load();
}

// Nonce search on the device: batches of 2^22 candidates in increasing order until a batch
// has a hit; that batch's minimum is the smallest nonce.  The stream is the runner's, idle
// between the two programs.
constexpr uint32_t kHostGrindBits = 12;  // expected 2^bits candidates: up to 4096 stay on host threads
constexpr uint32_t kGrindBatch = 1u << 22;

azk_status grind_device(Runner &R, const uint64_t s[25], uint32_t bits, uint64_t &nonce)
{
    std::lock_guard<std::mutex> g(R.grind_mu);
    hipStream_t st = R.ctx->stream;
    azk_status e = R.grind_dev.ensure(256);
    if (e) return e;
    uint64_t *state = (uint64_t *)R.grind_dev.p;
    unsigned long long *best = (unsigned long long *)((uint8_t *)R.grind_dev.p + 200);
    unsigned long long none = ~0ull, got = ~0ull;
    HIP_TRY(hipMemcpyAsync(state, s, 200, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(best, &none, 8, hipMemcpyHostToDevice, st));
    for (uint64_t base = 0; got == ~0ull; base += kGrindBatch) {
        launch_grind(st, state, bits, base, kGrindBatch, best);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(&got, best, 8, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
    }
    nonce = got;
    return AZK_OK;
}

// Host step of one job: final-layer frame + final layer (+ grinding frame and nonce) +
// query frame, then Q indices mod N/2 (CircleFRITranscriptV1.derive,
// CircleProofFormat.swift:687-722, grinding :753-794).
void host_final(Runner &R, size_t k, unsigned threads)
{
    const azk_circle_prover *p = R.jobs[k].p;
    host::Transcript t;
    const uint8_t *tail = (const uint8_t *)R.tail_host.p + R.tail_off[k];
    memcpy(t.s, tail, 200);
    t.absorb(p->final_frame.data(), p->final_frame.size());
    t.absorb(tail + 200, 16 * p->nfin);
    if (p->d.grinding_bits) {
        t.absorb(p->grind_frame.data(), p->grind_frame.size());
        uint64_t nonce = 0;
        if (p->d.grinding_bits <= kHostGrindBits) {
            nonce = find_nonce(t.s, p->d.grinding_bits, threads);
        } else {
            const azk_status e = grind_device(R, t.s, p->d.grinding_bits, nonce);
            if (e) {
                std::lock_guard<std::mutex> g(R.grind_mu);
                R.grind_err = e;
                return;
            }
        }
        uint8_t nb[8];
        memcpy(nb, &nonce, 8);
        t.absorb(nb, 8);
        R.nonce[k] = nonce;
    }
    t.absorb(p->query_frame.data(), p->query_frame.size());
    t.squeeze_u32(p->d.query_count, (uint32_t)(p->n / 2), (uint32_t *)((uint8_t *)R.qidx_host.p + R.qidx_off[k]));
}

azk_status host_finals(Runner &R)
{
    R.grind_err = AZK_OK;
    const size_t jobs = R.jobs.size();
    const size_t hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const size_t nt = std::min(jobs, hw);
    const unsigned search = (unsigned)std::max<size_t>(1, hw / std::max<size_t>(nt, 1));  // nonce threads per job
    if (nt <= 1) {
        for (size_t k = 0; k < jobs; ++k) host_final(R, k, search);
        return R.grind_err;
    }
    std::vector<std::thread> pool;
    for (size_t t = 1; t < nt; ++t)
        pool.emplace_back([&R, t, nt, jobs, search]() {
            for (size_t k = t; k < jobs; k += nt) host_final(R, k, search);
        });
    for (size_t k = 0; k < jobs; k += nt) host_final(R, k, search);
    for (auto &th : pool) th.join();
    return R.grind_err;
}

// The grinding nonces go into the proofs after the gather (the emit kernel does not know them).
void patch_nonces(Runner &R)
{
    for (size_t k = 0; k < R.jobs.size(); ++k) {
        const azk_circle_prover *p = R.jobs[k].p;
        if (!p->d.grinding_bits) continue;
        uint32_t *w = (uint32_t *)((uint8_t *)R.proof_host.p + R.proof_off[k]) + p->nonce_off;
        w[0] = (uint32_t)R.nonce[k];
        w[1] = (uint32_t)(R.nonce[k] >> 32);
    }
}

azk_status enqueue_b(Runner &R, hipStream_t st)
{
    // reads the query indices from, and writes the proofs into, pinned host memory
    launch_emit_proof(st, tab(R, R.emit), R.emit.n(), R.emit.blocks);
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

azk_status ensure_graphs(Runner &R, hipStream_t st)
{
    if (R.graph_a) return AZK_OK;
    PhaseClock none(st, false);
    azk_status s = capture(st, &R.graph_a, [&]() { return enqueue_a(R, st, none); });
    if (!s) s = capture(st, &R.graph_b, [&]() { return enqueue_b(R, st); });
    return s;
}

// Runs a program (graph replay or eager launches with optional phase timing).
azk_status run(Runner &R, bool graph, azk_phase_times *phases, azk_stats *stats)
{
    hipStream_t st = R.ctx->stream;
    if (graph) {
        azk_status s = ensure_graphs(R, st);
        if (s) return s;
    }
    Timer t;
    azk_status s = t.init();
    if (s) return s;
    PhaseClock clk(st, phases != nullptr);
    const uint64_t w0 = now_ns();
    HIP_TRY(hipEventRecord(t.a, st));
    if (graph) HIP_TRY(hipGraphLaunch(R.graph_a, st));
    else if ((s = enqueue_a(R, st, clk))) return s;
    HIP_TRY(hipGetLastError());
    const uint64_t w1 = now_ns();
    HIP_TRY(hipStreamSynchronize(st));
    const uint64_t h0 = now_ns();
    if ((s = host_finals(R))) return s;
    const uint64_t h1 = now_ns();
    clk.mark(PH_HOST);
    if (graph) HIP_TRY(hipGraphLaunch(R.graph_b, st));
    else if ((s = enqueue_b(R, st))) return s;
    HIP_TRY(hipGetLastError());
    clk.mark(PH_TAIL);
    HIP_TRY(hipEventRecord(t.b, st));
    HIP_TRY(hipStreamSynchronize(st));
    patch_nonces(R);
    if (stats) {
        stats->submit_ns = w1 - w0;
        stats->wall_ns = now_ns() - w0;
        stats->gpu_ns = t.elapsed_ns();
    }
    if (phases && clk.ev.size() > 1) {
        azk_phase_times ph{};
        uint64_t *slot[] = {nullptr, &ph.fft_ns, &ph.commit0_ns, &ph.fold_ns, &ph.tree_ns, &ph.transcript_ns, &ph.emit_ns,
                            nullptr};
        for (size_t i = 1; i < clk.ev.size(); ++i) {
            float ms = 0;
            (void)hipEventElapsedTime(&ms, clk.ev[i - 1], clk.ev[i]);
            if (slot[clk.tag[i]]) *slot[clk.tag[i]] += (uint64_t)(ms * 1e6);
        }
        ph.host_ns = h1 - h0;
        auto count = [&](const std::vector<Launch<TreeDesc>> &steps) {
            for (const auto &l : steps)
                if (l.blocks) {
                    ++ph.tree_launches;
                    ph.tree_levels += l.v.empty() ? 0 : l.v[0].log_c;
                }
        };
        count(R.tree0);
        for (const auto &t : R.tree_r) count(t);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, clk.ev.front(), clk.ev.back());
        ph.total_ns = (uint64_t)(ms * 1e6);
        *phases = ph;
    }
    return AZK_OK;
}

const uint8_t *proof_of(const Runner &R, size_t k) { return (const uint8_t *)R.proof_host.p + R.proof_off[k]; }

azk_status prove_single(azk_circle_prover *p, const void *src, int mode, uint8_t *proof, size_t cap, size_t *len,
                        azk_stats *stats)
{
    const size_t bytes = 4 * p->proof_words;
    if (cap < bytes || !proof) return fail(AZK_ERR_INVALID_LAYOUT, "proof capacity %zu < %zu", cap, bytes);
    if (!p->runner || p->runner->jobs[0].src != src || p->runner->jobs[0].mode != mode) {
        p->runner.reset();
        auto r = std::make_unique<Runner>();
        azk_status s = runner_build(*r, p->ctx, {Job{p, src, mode}});
        if (s) return s;
        p->runner = std::move(r);
    }
    azk_stats local{};
    azk_stats *st = stats ? stats : &local;
    azk_status s = run(*p->runner, p->use_graph && !p->timing, p->timing ? &p->phases : nullptr, st);
    if (s) return s;
    memcpy(proof, proof_of(*p->runner, 0), bytes);
    if (len) *len = bytes;
    if (p->drift_on) {  // PlanDatabase.recordLiveObservation: EMA with alpha 0.2
        const double x = (double)st->gpu_ns;
        p->drift_ema_ns = p->drift_samples ? 0.2 * x + 0.8 * p->drift_ema_ns : x;
        ++p->drift_samples;
    }
    return AZK_OK;
}

// The commit / transcript / fold program of one job without the host step and the proof:
// roots (R x 32 bytes), challenges (R x 4 words) and the final layer (N >> R QM31, copied with
// `final_kind` to `final_dst`).  Reference: CircleFRIMerkleTranscriptFoldChainPlan
// (CircleFRIFoldPlan.swift:301-649).
azk_status chain_single(azk_circle_prover *p, const void *src, int mode, uint8_t *roots, uint32_t *challenges,
                        void *final_dst, hipMemcpyKind final_kind, azk_stats *stats)
{
    if (!p->runner || p->runner->jobs[0].src != src || p->runner->jobs[0].mode != mode) {
        p->runner.reset();
        auto r = std::make_unique<Runner>();
        azk_status s = runner_build(*r, p->ctx, {Job{p, src, mode}});
        if (s) return s;
        p->runner = std::move(r);
    }
    Runner &R = *p->runner;
    hipStream_t st = p->ctx->stream;
    const bool graph = p->use_graph && !p->timing;
    if (graph) {
        azk_status s = ensure_graphs(R, st);
        if (s) return s;
    }
    return run_timed(p->ctx, stats, [&]() -> azk_status {
        if (graph) {
            HIP_TRY(hipGraphLaunch(R.graph_a, st));
        } else {
            PhaseClock none(st, false);
            azk_status s = enqueue_a(R, st, none);
            if (s) return s;
        }
        const uint32_t rounds = p->d.rounds;
        for (uint32_t r = 0; r < rounds; ++r)
            HIP_TRY(hipMemcpyAsync(roots + 32 * r, p->trees[r] + 8 * (2 * (p->n >> r) - 2), 32, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipMemcpyAsync(challenges, p->challenges, 16 * (size_t)rounds, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipMemcpyAsync(final_dst, p->layers[rounds], 16 * p->nfin, final_kind, st));
        return AZK_OK;
    });
}

void runner_reset(Runner &R)
{
    R.jobs.clear();
    R.tables.host.clear();
    R.fft_pass.clear();
    R.fft_final = {};
    R.fft_pass_m31.clear();
    R.fft_final_m31 = {};
    R.tree0.clear();
    R.rounds = 0;
    R.tr_a.clear();
    R.fold_big.clear();
    R.fold_leaf.clear();
    R.tree_r.clear();
    R.tail = {};
    R.emit = {};
    R.tail_off.clear();
    R.qidx_off.clear();
    R.proof_off.clear();
    R.nonce.clear();
}

size_t prover_footprint(const azk_circle_prover_desc &d)
{
    const size_t n = (size_t)1 << d.log_n;
    size_t layers = 0, trees = 0;
    for (uint32_t r = 0; r <= d.rounds; ++r) layers += align_up((size_t)16 * (n >> r), 256);
    for (uint32_t r = 0; r < d.rounds; ++r) trees += align_up((2 * (n >> r) - 1) * 32, 256);
    const size_t twiddles = 4 * n + 4 * n;  // FFT stages + inverted fold stages
    const size_t staging = 16 * n;          // device copy of host coefficients
    return layers + trees + twiddles + staging + 64 * 1024;
}

azk_status validate_desc(const azk_circle_prover_desc &d)
{
    // CirclePCSFRIProofV1.init / CircleFRISecurityParametersV1.init (CircleProofFormat.swift:129-158,355-393)
    if (d.log_n < 2 || d.log_n > 28 || d.rounds < 1 || d.rounds > d.log_n || d.rounds > (uint32_t)kMaxRounds)
        return fail(AZK_ERR_INVALID_LAYOUT, "log_n %u / rounds %u outside the supported shape", d.log_n, d.rounds);
    if (d.log_blowup == 0 || d.query_count == 0 || d.query_count > 65536 || d.folding_step != 1 ||
        (uint64_t)d.log_blowup * d.query_count + d.grinding_bits > 0xffffffffull)
        return fail(AZK_ERR_INVALID_LAYOUT, "invalid security parameters");
    if (d.grinding_bits > kMaxGrindingBits)  // CircleFRIGrindingV1.maximumLocalSearchBits (:654)
        return fail(AZK_ERR_INVALID_LAYOUT, "grinding_bits %u above %u", d.grinding_bits, kMaxGrindingBits);
    return AZK_OK;
}

azk_status prover_create(azk_context *ctx, const azk_circle_prover_desc &d, azk_circle_prover **out)
{
    azk_status s = validate_desc(d);
    if (s) return s;
    auto *p = new azk_circle_prover();
    p->ctx = ctx;
    p->d = d;
    p->n = 1ull << d.log_n;
    p->nfin = p->n >> d.rounds;
    HIP_TRY(hipSetDevice(ctx->device));
    std::lock_guard<std::mutex> gc(ctx->mu);
    s = prover_build(p);
    if (s) {
        delete p;
        return s;
    }
    ctx_retain(p->ctx);
    *out = p;
    return AZK_OK;
}

std::vector<std::unique_lock<std::mutex>> batch_lock_provers(azk_circle_batch *b)
{
    // Prover locks in one global order (by address): two batches that share provers, listed
    // in different orders, then cannot each hold one lock while waiting for the other.
    std::vector<azk_circle_prover *> order(b->provers);
    std::sort(order.begin(), order.end(), std::less<azk_circle_prover *>());
    std::vector<std::unique_lock<std::mutex>> locks;
    for (auto *p : order) locks.emplace_back(p->mu);
    return locks;
}

azk_status batch_rebuild(azk_circle_batch *b)
{
    b->runner = std::make_unique<Runner>();
    azk_status s = runner_build(*b->runner, b->ctx, b->jobs);
    if (!s) s = ensure_graphs(*b->runner, b->ctx->stream);
    return s;
}

} // namespace azk::prv

azk_circle_prover::~azk_circle_prover() = default;

extern "C" azk_status azk_circle_prover_create(azk_context *ctx, const azk_circle_prover_desc *desc,
                                               azk_circle_prover **out)
{
    CHECK_ARG(ctx && desc && out);
    return prover_create(ctx, *desc, out);
}

extern "C" void azk_circle_prover_destroy(azk_circle_prover *p)
{
    if (!p) return;
    (void)hipSetDevice(p->ctx->device);
    (void)hipStreamSynchronize(p->ctx->stream);
    destroy_on_ctx(p);
}

extern "C" size_t azk_circle_prover_proof_size(const azk_circle_prover *p) { return p ? 4 * p->proof_words : 0; }

extern "C" azk_status azk_circle_prover_set_graph(azk_circle_prover *p, int enable)
{
    CHECK_ARG(p);
    std::lock_guard<std::mutex> g(p->mu);
    p->use_graph = enable != 0;
    return AZK_OK;
}

extern "C" azk_status azk_circle_prover_set_exec_plan(azk_circle_prover *p, const azk_circle_exec_plan *plan)
{
    CHECK_ARG(p && plan);
    azk_status s = check_exec_plan(*plan);
    if (s) return s;
    std::lock_guard<std::mutex> g(p->mu);
    HIP_TRY(hipSetDevice(p->ctx->device));
    (void)hipStreamSynchronize(p->ctx->stream);
    p->xp = to_exec_plan(*plan);
    p->runner.reset();  // the cached program was built for the previous tiles
    return AZK_OK;
}

extern "C" azk_status azk_circle_prover_get_exec_plan(const azk_circle_prover *p, azk_circle_exec_plan *plan)
{
    CHECK_ARG(p && plan);
    *plan = from_exec_plan(p->xp);
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
    return prove_single(p, (uint8_t *)coeffs->ptr + off, 0, proof, cap, len, stats);
}

// The same proof from M31 line-basis coefficients: N words, coefficient i's QM31 lift is
// (coeffs[i], 0, 0, 0).  The FFT runs on M31 words (a quarter of the QM31 FFT's arithmetic
// and bytes); the first layer, its leaves and everything after are the QM31 path's.
extern "C" azk_status azk_circle_prove_fft_coefficients_m31_resident(azk_circle_prover *p, const azk_buffer *coeffs,
                                                                     size_t off, uint8_t *proof, size_t cap, size_t *len,
                                                                     azk_stats *stats)
{
    CHECK_ARG(p && coeffs);
    std::lock_guard<std::mutex> g(p->mu);
    azk_status s = check_range(coeffs, off, (size_t)4 << p->d.log_n);
    if (s) return s;
    if (((uintptr_t)coeffs->ptr + off) % 4) return fail(AZK_ERR_INVALID_LAYOUT, "coefficients must be 4-byte aligned");
    HIP_TRY(hipSetDevice(p->ctx->device));
    std::lock_guard<std::mutex> gc(p->ctx->mu);
    return prove_single(p, (uint8_t *)coeffs->ptr + off, 2, proof, cap, len, stats);
}

extern "C" azk_status azk_circle_prove_fft_coefficients_m31(azk_circle_prover *p, const uint32_t *coeffs, size_t words,
                                                            uint8_t *proof, size_t cap, size_t *len, azk_stats *stats)
{
    CHECK_ARG(p && coeffs);
    std::lock_guard<std::mutex> g(p->mu);
    const size_t need = (size_t)4 << p->d.log_n;
    if (words * 4 < need) return fail(AZK_ERR_INVALID_LAYOUT, "coefficient words %zu < %zu", words, need / 4);
    HIP_TRY(hipSetDevice(p->ctx->device));
    azk_status s = p->coeff_staging.ensure(need);
    if (s) return s;
    std::lock_guard<std::mutex> gc(p->ctx->mu);
    HIP_TRY(hipMemcpyAsync(p->coeff_staging.p, coeffs, need, hipMemcpyHostToDevice, p->ctx->stream));
    return prove_single(p, p->coeff_staging.p, 2, proof, cap, len, stats);
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
    return prove_single(p, (uint8_t *)evals->ptr + off, 1, proof, cap, len, stats);
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
    return prove_single(p, p->coeff_staging.p, 0, proof, cap, len, stats);
}

// proveCircleFFTCoefficientsResidentVerified (CircleCodeword.swift:1082-1113): the proof
// must verify and every opened first-layer value must equal the host Circle FFT of the
// coefficients (the reference rebuilds the whole proof on the CPU).
// proveVerified (CircleCodeword.swift:1152-1177): the whole proof rebuilt on the host from the
// coefficients must equal the device's byte for byte, and the verifier must accept it.
static azk_status gate_proof(azk_circle_prover *p, const uint32_t *coeffs, const uint8_t *proof, size_t len)
{
    std::vector<uint8_t> want;
    try {
        if (!host::circle_prove(coeffs, p->d, want))
            return fail(AZK_ERR_CORRECTNESS, "the CPU gate could not fold this domain");
    } catch (const std::bad_alloc &) {
        return fail(AZK_ERR_ALLOCATION, "the CPU gate ran out of host memory");
    }
    if (want.size() != len || memcmp(want.data(), proof, len))
        return fail(AZK_ERR_CORRECTNESS, "Circle PCS proof differs from the CPU gate's rebuild");
    int ok = 0;
    azk_status s = azk_circle_verify(proof, len, p->d.public_input_digest, &ok);
    if (s) return s;
    if (!ok) return fail(AZK_ERR_CORRECTNESS, "Circle PCS proof rejected by the CPU verifier");
    return AZK_OK;
}

extern "C" azk_status azk_circle_prove_fft_coefficients_verified(azk_circle_prover *p, const void *coeffs, size_t bytes,
                                                                 uint8_t *proof, size_t cap, size_t *len,
                                                                 azk_stats *stats)
{
    size_t n = 0;
    azk_status s = azk_circle_prove_fft_coefficients(p, coeffs, bytes, proof, cap, &n, stats);
    if (s) return s;
    if (len) *len = n;
    return gate_proof(p, (const uint32_t *)coeffs, proof, n);
}

extern "C" azk_status azk_circle_prove_fft_coefficients_resident_verified(azk_circle_prover *p,
                                                                          const azk_buffer *coeffs, size_t offset,
                                                                          uint8_t *proof, size_t cap, size_t *len,
                                                                          azk_stats *stats)
{
    size_t n = 0;
    azk_status s = azk_circle_prove_fft_coefficients_resident(p, coeffs, offset, proof, cap, &n, stats);
    if (s) return s;
    if (len) *len = n;
    std::vector<uint32_t> host((size_t)4 << p->d.log_n);
    s = azk_buffer_download(coeffs, offset, host.data(), 4 * host.size());
    if (s) return s;
    return gate_proof(p, host.data(), proof, n);
}

extern "C" azk_status azk_circle_prover_clear(azk_circle_prover *p)
{
    CHECK_ARG(p);
    std::lock_guard<std::mutex> g(p->mu);
    HIP_TRY(hipSetDevice(p->ctx->device));
    std::lock_guard<std::mutex> gc(p->ctx->mu);
    HIP_TRY(hipStreamSynchronize(p->ctx->stream));
    // zero every buffer that held witness-derived data, then restore the constant prefix state
    std::vector<uint8_t> st0(200);
    HIP_TRY(hipMemcpy(st0.data(), p->tstate_init, 200, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemset(p->arena.p, 0, p->arena.n));
    HIP_TRY(hipMemcpy(p->tstate_init, st0.data(), 200, hipMemcpyHostToDevice));
    if (p->coeff_staging.p) HIP_TRY(hipMemset(p->coeff_staging.p, 0, p->coeff_staging.n));
    p->runner.reset();  // its staging held proof bytes
    return AZK_OK;
}

// ====================================================================== batched form


// mode 0: QM31 line-basis coefficients (16 bytes each); mode 2: M31 words (4 bytes each).
static azk_status batch_create(azk_circle_prover *const *provers, const azk_buffer *const *coeffs, const size_t *offsets,
                               uint32_t jobs, int mode, azk_circle_batch **out)
{
    CHECK_ARG(provers && coeffs && offsets && out && jobs);
    azk_context *ctx = provers[0]->ctx;
    const size_t word = mode == 2 ? 4 : 16;
    std::vector<Job> list;
    for (uint32_t j = 0; j < jobs; ++j) {
        CHECK_ARG(provers[j] && coeffs[j]);
        if (provers[j]->ctx != ctx) return fail(AZK_ERR_INVALID_ARGUMENT, "all provers must share one context");
        for (uint32_t k = 0; k < j; ++k)
            if (provers[k] == provers[j]) return fail(AZK_ERR_INVALID_ARGUMENT, "a prover appears twice in one batch");
        azk_status s = check_range(coeffs[j], offsets[j], word << provers[j]->d.log_n);
        if (s) return s;
        if (((uintptr_t)coeffs[j]->ptr + offsets[j]) % word) return fail(AZK_ERR_INVALID_LAYOUT, "job %u misaligned", j);
        list.push_back(Job{provers[j], (const uint8_t *)coeffs[j]->ptr + offsets[j], mode});
    }
    HIP_TRY(hipSetDevice(ctx->device));
    auto *b = new azk_circle_batch();
    b->ctx = ctx;
    b->provers.assign(provers, provers + jobs);
    b->jobs = list;
    std::lock_guard<std::mutex> gc(ctx->mu);
    azk_status s = batch_rebuild(b);
    if (s) {
        delete b;
        return s;
    }
    *out = b;
    return AZK_OK;
}

extern "C" azk_status azk_circle_batch_create(azk_circle_prover *const *provers, const azk_buffer *const *coeffs,
                                              const size_t *offsets, uint32_t jobs, azk_circle_batch **out)
{
    return batch_create(provers, coeffs, offsets, jobs, 0, out);
}

extern "C" azk_status azk_circle_batch_create_m31(azk_circle_prover *const *provers, const azk_buffer *const *coeffs,
                                                  const size_t *offsets, uint32_t jobs, azk_circle_batch **out)
{
    return batch_create(provers, coeffs, offsets, jobs, 2, out);
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
    const size_t jobs = b->provers.size();
    for (size_t j = 0; j < jobs; ++j) {
        CHECK_ARG(proofs[j]);
        if (caps[j] < 4 * b->provers[j]->proof_words) return fail(AZK_ERR_INVALID_LAYOUT, "proof capacity of job %zu", j);
    }
    // Prover locks in one global order (by address): two batches that share provers, listed
    // in different orders, then cannot each hold one lock while waiting for the other.
    std::vector<std::unique_lock<std::mutex>> locks = batch_lock_provers(b);
    HIP_TRY(hipSetDevice(b->ctx->device));
    std::lock_guard<std::mutex> gc(b->ctx->mu);
    azk_status s = run(*b->runner, true, nullptr, stats);
    if (s) return s;
    for (size_t j = 0; j < jobs; ++j) {
        memcpy(proofs[j], proof_of(*b->runner, j), 4 * b->provers[j]->proof_words);
        if (lens) lens[j] = 4 * b->provers[j]->proof_words;
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
