// Internal interface of the Circle PCS prover, shared by the plain and batched entry
// points (prover.hip) and the serving loop (server.hip).  Not part of the C ABI.
#pragma once
#include <memory>

#include "runtime_common.hpp"

namespace azk::prv {

using namespace azk::rt;

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

struct Runner;

} // namespace azk::prv

struct azk_circle_prover {
    azk_context *ctx = nullptr;
    azk_circle_prover_desc d{};
    azk::rt::ExecPlan xp;  // tile shapes of the program (default or tuned)
    uint64_t n = 0, nfin = 0;
    azk::rt::DeviceMem fft_tw, fold_tw, arena, frames, ops, coeff_staging;
    std::vector<uint4 *> layers;    // rounds + 1 layers (layer 0 = codeword, layer R = final)
    std::vector<uint32_t *> trees;  // one full tree per committed layer
    uint32_t *challenges = nullptr;
    uint64_t *tstate = nullptr, *tstate_init = nullptr;
    std::vector<std::pair<uint32_t, uint32_t>> prog_a;  // transcript program of round r: (first op, count)
    std::vector<std::pair<uint32_t, uint32_t>> prog_c;  // commitment frame r + 1, run beside fold r
    std::vector<uint8_t> final_frame, query_frame, grind_frame;
    std::vector<uint32_t> proof_template;
    uint64_t roots_off = 0, final_off = 0, queries_off = 0, query_words = 0, proof_words = 0;
    uint64_t nonce_off = 0;  // word of the grinding nonce in the proof (grinding_bits > 0)
    std::vector<uint32_t> layer_off;
    bool use_graph = true, timing = false;
    std::unique_ptr<azk::prv::Runner> runner;  // cached single-job program
    azk_phase_times phases{};
    // drift of a tuned plan: EMA of live prove times against the tuned median
    bool drift_on = false;
    uint32_t drift_samples = 0;
    double drift_baseline_ns = 0, drift_ema_ns = 0;
    std::mutex mu;
    ~azk_circle_prover();
};

namespace azk::prv {

struct Job {
    azk_circle_prover *p;
    const void *src;
    int mode;  // 0: src = line-basis coefficients, 1: src = codeword, 2: src = M31 line-basis coefficients
    const uint4 *layer0() const { return mode == 1 ? (const uint4 *)src : p->layers[0]; }
};

struct Runner {
    azk_context *ctx = nullptr;
    std::vector<Job> jobs;
    TableArena tables;
    std::vector<Launch<FftPassDesc>> fft_pass, fft_pass_m31;
    Launch<FftFinalDesc> fft_final, fft_final_m31;
    std::vector<Launch<TreeDesc>> tree0;
    uint32_t rounds = 0;
    std::vector<Launch<TranscriptDesc>> tr_a;
    std::vector<Launch<FoldDesc>> fold_big, fold_leaf;
    std::vector<std::vector<Launch<TreeDesc>>> tree_r;
    Launch<CopyDesc> tail;
    uint32_t tree_waves = 4;  // largest tree_waves of the jobs' plans
    Launch<EmitParams> emit;
    PinnedMem tail_host, qidx_host, proof_host;  // written / read by the kernels directly
    std::vector<size_t> tail_off, qidx_off, proof_off;  // bytes, per job
    std::vector<uint64_t> nonce;                         // grinding nonce per job (grinding_bits > 0)
    DeviceMem grind_dev;                                 // device nonce search: state + best
    std::mutex grind_mu;                                 // one device search at a time
    azk_status grind_err = AZK_OK;
    hipGraphExec_t graph_a = nullptr, graph_b = nullptr;
    ~Runner()
    {
        if (graph_a) (void)hipGraphExecDestroy(graph_a);
        if (graph_b) (void)hipGraphExecDestroy(graph_b);
    }
};

// Builds the launch program of a job list into R (which must be fresh or reset).
azk_status runner_build(Runner &R, azk_context *ctx, const std::vector<Job> &jobs);
// Forgets the job list and launch tables of R; keeps its device/pinned staging for the
// next build.  R must not hold captured graphs.
void runner_reset(Runner &R);
azk_status enqueue_a(Runner &R, hipStream_t st, PhaseClock &clk);
// The host step of every job (final-layer absorb, grinding nonce search, query indices).
azk_status host_finals(Runner &R);
// Writes the grinding nonces into the gathered proofs (after enqueue_b has completed).
void patch_nonces(Runner &R);
azk_status enqueue_b(Runner &R, hipStream_t st);
const uint8_t *proof_of(const Runner &R, size_t k);
// One prove through the prover's cached single-job program (rebuilt when the source,
// the mode or the execution plan changed).  Caller holds p->mu and the context lock.
azk_status prove_single(azk_circle_prover *p, const void *src, int mode, uint8_t *proof, size_t cap, size_t *len,
                        azk_stats *stats);
// The same program up to the final layer (no host step, no proof): roots, challenges and the
// final layer.  Caller holds p->mu and the context lock.
azk_status chain_single(azk_circle_prover *p, const void *src, int mode, uint8_t *roots, uint32_t *challenges,
                        void *final_dst, hipMemcpyKind final_kind, azk_stats *stats);

// Graphs of R's program (captured on first use) and one run of it: graph A, the host step,
// graph B (or eager launches with optional phase timing when `graph` is false).
azk_status ensure_graphs(Runner &R, hipStream_t st);
azk_status run(Runner &R, bool graph, azk_phase_times *phases, azk_stats *stats);

// Descriptor rules of CircleFRISecurityParametersV1 / CirclePCSFRIProofV1 plus the
// device path's limits (folding step 1, no grinding).
azk_status validate_desc(const azk_circle_prover_desc &d);
// Validates the descriptor and builds a prover; the caller holds no locks.
azk_status prover_create(azk_context *ctx, const azk_circle_prover_desc &d, azk_circle_prover **out);
// Device bytes a prover of this shape holds (twiddles, layers, trees, staging).
size_t prover_footprint(const azk_circle_prover_desc &d);

} // namespace azk::prv

// A batch program over several provers (the batched form, prover.hip).
struct azk_circle_batch {
    azk_context *ctx = nullptr;
    std::vector<azk_circle_prover *> provers;
    std::vector<azk::prv::Job> jobs;
    std::unique_ptr<azk::prv::Runner> runner;
};

namespace azk::prv {

// The provers' locks, taken in address order.
std::vector<std::unique_lock<std::mutex>> batch_lock_provers(azk_circle_batch *b);
// Rebuilds the batch's program from its jobs with each prover's current execution plan.
// Caller holds the provers' locks and the context lock.
azk_status batch_rebuild(azk_circle_batch *b);

} // namespace azk::prv
