// Autotuner of the commit program's tile shapes, with persisted plan records and drift.
//
// Reference: MetalProofPlanner.tuneMerkleCommitExecutionPlan (Planner/MetalProofPlanner.swift:
// 20-117: differential gate, warmups, measured runs, objective median GPU + median CPU
// submit with p95 tie break :604-614, randomized winner validation before persistence),
// PlanDatabase (Planner/PlanDatabase.swift: all race rows stored, winner lookup keyed by
// device, workload, shader hash and protocol hash; drift EMA alpha 0.2, threshold 25%,
// 8 samples :15-36).  Here the gate compares a candidate's proof bytes with the default
// plan's proof (every plan computes the same proof), the race is a coordinate descent
// over the five knobs of azk_circle_exec_plan, and the store is a JSON-lines file.
#include <dlfcn.h>

#include <algorithm>
#include <cmath>
#include <fstream>
#include <random>
#include <sstream>

#include "host_sha3.hpp"
#include "prover_internal.hpp"

using namespace azk;
using namespace azk::rt;
using namespace azk::prv;

namespace {

constexpr const char *kProtocol = "AppleZKProver.CircleFRI.PCS.V1";
std::mutex g_db_mu;

// SHA3-256 of the loaded library file: changes whenever the kernels are rebuilt
// differently (the reference keys records by its shader source hash).
std::string build_hash()
{
    static std::string h;
    static std::once_flag once;
    std::call_once(once, [] {
        Dl_info info{};
        std::string path = dladdr((void *)&build_hash, &info) && info.dli_fname ? info.dli_fname : "";
        std::ifstream f(path, std::ios::binary);
        std::vector<uint8_t> bytes((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
        uint8_t d[32];
        host::sha3_256(bytes.data(), bytes.size(), d);
        char hex[65];
        for (int i = 0; i < 32; ++i) snprintf(hex + 2 * i, 3, "%02x", d[i]);
        h.assign(hex, 16);
    });
    return h;
}

std::string device_key(const azk_context *ctx)
{
    std::ostringstream o;
    o << "{\"arch\": \"" << ctx->props.gcnArchName << "\", \"cus\": " << ctx->props.multiProcessorCount
      << ", \"memory_mib\": " << (ctx->props.totalGlobalMem >> 20) << "}";
    return o.str();
}

std::string workload_key(const azk_circle_prover_desc &d)
{
    std::ostringstream o;
    o << "{\"stage\": \"circle_pcs_commit\", \"log_n\": " << d.log_n << ", \"rounds\": " << d.rounds
      << ", \"log_blowup\": " << d.log_blowup << ", \"queries\": " << d.query_count << "}";
    return o.str();
}

std::string plan_json(const ExecPlan &x)
{
    std::ostringstream o;
    o << "{\"log_tile_commit\": " << x.log_tile_commit << ", \"log_tile_fold\": " << x.log_tile_fold
      << ", \"log_leafwise_below\": " << x.log_leafwise_below << ", \"log_tile_leafwise\": " << x.log_tile_leafwise
      << ", \"log_tree_chunk\": " << x.log_tree_chunk << ", \"log_leafwise_grid\": " << x.log_leafwise_grid
      << ", \"tree_waves\": " << x.tree_waves << "}";
    return o.str();
}

bool read_field(const std::string &line, const char *key, double &v)
{
    const std::string k = std::string("\"") + key + "\": ";
    size_t at = line.find(k);
    if (at == std::string::npos) return false;
    v = strtod(line.c_str() + at + k.size(), nullptr);
    return true;
}
bool read_field(const std::string &line, const char *key, uint32_t &v)
{
    double d;
    if (!read_field(line, key, d) || d < 0 || d > 4294967295.0) return false;
    v = (uint32_t)d;
    return true;
}

struct Measurement {
    ExecPlan plan;
    double median_gpu = 0, p95_gpu = 0, median_submit = 0;
    bool ok = false;
    double objective() const { return median_gpu + median_submit; }
};

bool better(const Measurement &a, const Measurement &b)
{
    if (a.objective() != b.objective()) return a.objective() < b.objective();
    return a.p95_gpu < b.p95_gpu;
}

bool same_plan(const ExecPlan &a, const ExecPlan &b)
{
    return a.log_tile_commit == b.log_tile_commit && a.log_tile_fold == b.log_tile_fold &&
           a.log_leafwise_below == b.log_leafwise_below && a.log_tile_leafwise == b.log_tile_leafwise &&
           a.log_tree_chunk == b.log_tree_chunk && a.log_leafwise_grid == b.log_leafwise_grid &&
           a.tree_waves == b.tree_waves;
}

double quantile(std::vector<double> v, double q)
{
    std::sort(v.begin(), v.end());
    const double pos = q * (v.size() - 1);
    const size_t i = (size_t)pos;
    return i + 1 < v.size() ? v[i] + (pos - i) * (v[i + 1] - v[i]) : v[i];
}

// Proves once with `plan` (the ungated first run also builds and captures the program).
azk_status prove_with(azk_circle_prover *p, const ExecPlan &plan, const void *src, std::vector<uint8_t> &out,
                      azk_stats *st)
{
    if (!same_plan(p->xp, plan)) {
        p->xp = plan;
        p->runner.reset();
    }
    out.resize(4 * p->proof_words);
    size_t len = 0;
    return prove_single(p, src, 0, out.data(), out.size(), &len, st);
}

azk_status measure(azk_circle_prover *p, const ExecPlan &plan, const void *src, const std::vector<uint8_t> &want,
                   const azk_tuning_config &c, Measurement &m)
{
    m.plan = plan;
    std::vector<uint8_t> got;
    azk_stats st{};
    azk_status s = prove_with(p, plan, src, got, &st);  // differential gate
    if (s) return s;
    if (got != want) return AZK_OK;  // rejected: m.ok stays false
    for (uint32_t k = 0; k < c.warmup_runs; ++k)
        if ((s = prove_with(p, plan, src, got, &st))) return s;
    std::vector<double> gpu, sub;
    for (uint32_t k = 0; k < c.measured_runs; ++k) {
        if ((s = prove_with(p, plan, src, got, &st))) return s;
        if (got != want) return AZK_OK;
        gpu.push_back((double)st.gpu_ns);
        sub.push_back((double)st.submit_ns);
    }
    m.median_gpu = quantile(gpu, 0.5);
    m.p95_gpu = quantile(gpu, 0.95);
    m.median_submit = quantile(sub, 0.5);
    m.ok = true;
    return AZK_OK;
}

// Canonical random QM31 coefficients (SplitMix64 words masked to 31 bits, p rejected).
void random_coefficients(uint64_t seed, size_t words, std::vector<uint32_t> &out)
{
    out.resize(words);
    uint64_t x = seed;
    for (size_t i = 0; i < words;) {
        x += 0x9e3779b97f4a7c15ull;
        uint64_t z = x;
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        z ^= z >> 31;
        for (int h = 0; h < 2 && i < words; ++h) {
            const uint32_t w = (uint32_t)(z >> (32 * h)) & kP;
            if (w != kP) out[i++] = w;
        }
    }
}

azk_status persist(const azk_tuning_config &c, const std::string &dev, const std::string &wl,
                   const std::vector<Measurement> &ms, const Measurement &win)
{
    if (!c.database_path) return AZK_OK;
    std::lock_guard<std::mutex> g(g_db_mu);
    std::ofstream f(c.database_path, std::ios::app);
    if (!f) return fail(AZK_ERR_COMMAND_FAILED, "cannot open plan database %s", c.database_path);
    const std::string bh = build_hash();
    const long long now = (long long)time(nullptr);
    for (const Measurement &m : ms) {
        if (!m.ok) continue;
        f << "{\"created_at\": " << now << ", \"device\": " << dev << ", \"workload\": " << wl
          << ", \"build_hash\": \"" << bh << "\", \"protocol_hash\": \"" << kProtocol << "\", \"plan\": "
          << plan_json(m.plan) << ", \"median_gpu_ns\": " << (long long)m.median_gpu
          << ", \"p95_gpu_ns\": " << (long long)m.p95_gpu << ", \"median_submit_ns\": " << (long long)m.median_submit
          << ", \"is_winner\": " << (same_plan(m.plan, win.plan) ? "true" : "false") << "}\n";
    }
    return f ? AZK_OK : fail(AZK_ERR_COMMAND_FAILED, "write to plan database %s failed", c.database_path);
}

// latestWinner: the last winning row for this device, workload, build and protocol.
bool lookup(const azk_tuning_config &c, const std::string &dev, const std::string &wl, ExecPlan &plan, double &median)
{
    if (!c.database_path) return false;
    std::lock_guard<std::mutex> g(g_db_mu);
    std::ifstream f(c.database_path);
    if (!f) return false;
    const std::string key = "\"device\": " + dev + ", \"workload\": " + wl + ", \"build_hash\": \"" + build_hash() +
                            "\", \"protocol_hash\": \"" + kProtocol + "\"";
    bool found = false;
    std::string line;
    while (std::getline(f, line)) {
        if (line.find(key) == std::string::npos || line.find("\"is_winner\": true") == std::string::npos) continue;
        ExecPlan x;
        double med = 0;
        if (read_field(line, "log_tile_commit", x.log_tile_commit) && read_field(line, "log_tile_fold", x.log_tile_fold) &&
            read_field(line, "log_leafwise_below", x.log_leafwise_below) &&
            read_field(line, "log_tile_leafwise", x.log_tile_leafwise) &&
            read_field(line, "log_tree_chunk", x.log_tree_chunk) &&
            read_field(line, "log_leafwise_grid", x.log_leafwise_grid) && read_field(line, "tree_waves", x.tree_waves) &&
            !check_exec_plan(from_exec_plan(x))) {
            plan = x;
            median = read_field(line, "median_gpu_ns", med) ? med : 0;
            found = true;
        }
    }
    return found;
}

struct Knob { uint32_t ExecPlan::*field; std::vector<uint32_t> values; };
const std::vector<Knob> &race_knobs()
{
    static const std::vector<Knob> knobs = {
        {&ExecPlan::log_tree_chunk, {3, 4, 5, 6, 8, 10}},
        {&ExecPlan::tree_waves, {8, 16}},
        {&ExecPlan::log_tile_fold, {9, 10, 11}},
        {&ExecPlan::log_tile_commit, {9, 10, 11}},
        {&ExecPlan::log_leafwise_below, {17, 18, 19, 20, 21}},
        {&ExecPlan::log_tile_leafwise, {6, 7, 8}},
        {&ExecPlan::log_leafwise_grid, {0, 7, 9, 10}},
    };
    return knobs;
}

} // namespace

extern "C" azk_status azk_circle_prover_tune(azk_circle_prover *p, const azk_buffer *coeffs, size_t off,
                                             const azk_tuning_config *config, azk_tuning_result *res)
{
    CHECK_ARG(p && coeffs && res);
    azk_tuning_config c{};
    if (config) c = *config;
    if (!c.warmup_runs) c.warmup_runs = 2;
    if (!c.measured_runs) c.measured_runs = 5;
    if (!c.max_candidates) c.max_candidates = 28;
    if (!c.validation_runs) c.validation_runs = 3;
    std::lock_guard<std::mutex> g(p->mu);
    azk_status s = check_range(coeffs, off, (size_t)16 << p->d.log_n);
    if (s) return s;
    if (((uintptr_t)coeffs->ptr + off) % 16) return fail(AZK_ERR_INVALID_LAYOUT, "coefficients must be 16-byte aligned");
    HIP_TRY(hipSetDevice(p->ctx->device));
    std::lock_guard<std::mutex> gc(p->ctx->mu);
    const void *src = (const uint8_t *)coeffs->ptr + off;
    *res = azk_tuning_result{};

    ExecPlan stored;
    double stored_median = 0;
    if (!c.force && lookup(c, device_key(p->ctx), workload_key(p->d), stored, stored_median)) {
        p->xp = stored;
        p->runner.reset();
        p->drift_on = stored_median > 0;
        p->drift_baseline_ns = stored_median;
        p->drift_samples = 0;
        res->winner = from_exec_plan(stored);
        res->median_gpu_ns = stored_median;
        res->from_database = 1;
        return AZK_OK;
    }

    p->drift_on = false;  // the race's own proves are not live observations
    const ExecPlan def;
    std::vector<uint8_t> want;
    s = prove_with(p, def, src, want, nullptr);
    if (s) return s;

    std::vector<Measurement> ms;
    Measurement best;
    s = measure(p, def, src, want, c, best);
    if (s) return s;
    if (!best.ok) return fail(AZK_ERR_CORRECTNESS, "the default plan is not deterministic");
    ms.push_back(best);
    res->default_median_gpu_ns = best.median_gpu;

    // coordinate descent: one knob at a time, the others at the best plan so far
    const std::vector<Knob> &knobs = race_knobs();
    uint32_t raced = 1;
    for (const Knob &k : knobs) {
        const ExecPlan base = best.plan;
        for (uint32_t v : k.values) {
            if (raced >= c.max_candidates) break;
            ExecPlan cand = base;
            cand.*(k.field) = v;
            if (std::any_of(ms.begin(), ms.end(), [&](const Measurement &m) { return same_plan(m.plan, cand); }))
                continue;
            Measurement m;
            s = measure(p, cand, src, want, c, m);
            if (s) return s;
            ++raced;
            ms.push_back(m);
            if (!m.ok) {
                ++res->candidates_rejected;
                continue;
            }
            if (better(m, best)) best = m;
        }
    }

    // randomized validation of the winner against the default plan before persistence
    if (!same_plan(best.plan, def) && c.validation_runs) {
        DeviceMem scratch;
        s = scratch.ensure((size_t)16 << p->d.log_n);
        if (s) return s;
        std::vector<uint32_t> words;
        std::vector<uint8_t> a, b;
        for (uint32_t k = 0; k < c.validation_runs; ++k) {
            random_coefficients(0x5eed0000ull + k, (size_t)4 << p->d.log_n, words);
            HIP_TRY(hipMemcpy(scratch.p, words.data(), words.size() * 4, hipMemcpyHostToDevice));
            if ((s = prove_with(p, def, scratch.p, a, nullptr)) || (s = prove_with(p, best.plan, scratch.p, b, nullptr)))
                return s;
            if (a != b) {
                p->xp = def;
                p->runner.reset();
                return fail(AZK_ERR_CORRECTNESS, "tuned plan disagrees with the default plan on validation input %u", k);
            }
        }
        p->runner.reset();  // the cached program points at the scratch input
    }
    s = persist(c, device_key(p->ctx), workload_key(p->d), ms, best);
    if (s) return s;
    if (!same_plan(p->xp, best.plan)) {
        p->xp = best.plan;
        p->runner.reset();
    }
    p->drift_on = true;
    p->drift_baseline_ns = best.median_gpu;
    p->drift_samples = 0;
    p->drift_ema_ns = 0;
    res->winner = from_exec_plan(best.plan);
    res->median_gpu_ns = best.median_gpu;
    res->p95_gpu_ns = best.p95_gpu;
    res->median_submit_ns = best.median_submit;
    res->candidates_measured = raced;
    return AZK_OK;
}

extern "C" azk_status azk_circle_prover_drift(const azk_circle_prover *p, azk_drift_status *out)
{
    CHECK_ARG(p && out);
    *out = azk_drift_status{};
    out->samples = p->drift_samples;
    out->ema_gpu_ns = p->drift_ema_ns;
    out->baseline_gpu_ns = p->drift_baseline_ns;
    if (p->drift_on && p->drift_baseline_ns > 0 && p->drift_samples) {
        out->relative_drift = std::fabs(p->drift_ema_ns - p->drift_baseline_ns) / p->drift_baseline_ns;
        out->stale = p->drift_samples >= 8 && out->relative_drift > 0.25;
    }
    return AZK_OK;
}

// ====================================================================== batch programs

namespace {

// The batch's job shapes in order, as a workload key: {"stage": "circle_pcs_batch", "jobs": J,
// "shapes": first 16 hex digits of SHA3-256 over (log N, rounds, blowup, queries) per job}.
std::string batch_workload_key(const azk_circle_batch *b)
{
    std::vector<uint8_t> bytes;
    for (const azk_circle_prover *p : b->provers)
        for (uint32_t w : {p->d.log_n, p->d.rounds, p->d.log_blowup, p->d.query_count})
            for (int k = 0; k < 4; ++k) bytes.push_back((uint8_t)(w >> (8 * k)));
    uint8_t d[32];
    host::sha3_256(bytes.data(), bytes.size(), d);
    char hex[17];
    for (int i = 0; i < 8; ++i) snprintf(hex + 2 * i, 3, "%02x", d[i]);
    std::ostringstream o;
    o << "{\"stage\": \"circle_pcs_batch\", \"jobs\": " << b->provers.size() << ", \"shapes\": \"" << hex << "\"}";
    return o.str();
}

// Every prover of the batch on `plan`, the batch program rebuilt for it.
azk_status batch_set_plan(azk_circle_batch *b, const ExecPlan &plan)
{
    for (azk_circle_prover *p : b->provers)
        if (!same_plan(p->xp, plan)) {
            p->xp = plan;
            p->runner.reset();  // their own single-job programs follow the plan
        }
    return batch_rebuild(b);
}

azk_status batch_run(azk_circle_batch *b, std::vector<uint8_t> &proofs, azk_stats *st)
{
    azk_status s = run(*b->runner, true, nullptr, st);
    if (s) return s;
    proofs.clear();
    for (size_t j = 0; j < b->provers.size(); ++j) {
        const uint8_t *pr = proof_of(*b->runner, j);
        proofs.insert(proofs.end(), pr, pr + 4 * b->provers[j]->proof_words);
    }
    return AZK_OK;
}

azk_status batch_measure(azk_circle_batch *b, const ExecPlan &plan, const std::vector<uint8_t> &want,
                         const azk_tuning_config &c, Measurement &m)
{
    m.plan = plan;
    azk_status s = batch_set_plan(b, plan);
    if (s) return s;
    std::vector<uint8_t> got;
    azk_stats st{};
    if ((s = batch_run(b, got, &st))) return s;
    if (got != want) return AZK_OK;  // differential gate: every job's proof
    for (uint32_t k = 0; k < c.warmup_runs; ++k)
        if ((s = batch_run(b, got, &st))) return s;
    std::vector<double> gpu, sub;
    for (uint32_t k = 0; k < c.measured_runs; ++k) {
        if ((s = batch_run(b, got, &st))) return s;
        if (got != want) return AZK_OK;
        gpu.push_back((double)st.gpu_ns);
        sub.push_back((double)st.submit_ns);
    }
    m.median_gpu = quantile(gpu, 0.5);
    m.p95_gpu = quantile(gpu, 0.95);
    m.median_submit = quantile(sub, 0.5);
    m.ok = true;
    return AZK_OK;
}

} // namespace

extern "C" azk_status azk_circle_batch_tune(azk_circle_batch *b, const azk_tuning_config *config, azk_tuning_result *res)
{
    CHECK_ARG(b && res);
    azk_tuning_config c{};
    if (config) c = *config;
    if (!c.warmup_runs) c.warmup_runs = 2;
    if (!c.measured_runs) c.measured_runs = 5;
    if (!c.max_candidates) c.max_candidates = 28;
    std::vector<std::unique_lock<std::mutex>> locks = batch_lock_provers(b);
    HIP_TRY(hipSetDevice(b->ctx->device));
    std::lock_guard<std::mutex> gc(b->ctx->mu);
    *res = azk_tuning_result{};
    const std::string dev = device_key(b->ctx), wl = batch_workload_key(b);
    for (azk_circle_prover *p : b->provers) p->drift_on = false;

    ExecPlan stored;
    double stored_median = 0;
    if (!c.force && lookup(c, dev, wl, stored, stored_median)) {
        azk_status s = batch_set_plan(b, stored);
        if (s) return s;
        res->winner = from_exec_plan(stored);
        res->median_gpu_ns = stored_median;
        res->from_database = 1;
        return AZK_OK;
    }
    const ExecPlan def;
    std::vector<uint8_t> want;
    azk_status s = batch_set_plan(b, def);
    if (!s) s = batch_run(b, want, nullptr);
    if (s) return s;
    std::vector<Measurement> ms;
    Measurement best;
    if ((s = batch_measure(b, def, want, c, best))) return s;
    if (!best.ok) return fail(AZK_ERR_CORRECTNESS, "the default batch program is not deterministic");
    ms.push_back(best);
    res->default_median_gpu_ns = best.median_gpu;
    uint32_t raced = 1;
    for (const Knob &k : race_knobs()) {
        const ExecPlan base = best.plan;
        for (uint32_t v : k.values) {
            if (raced >= c.max_candidates) break;
            ExecPlan cand = base;
            cand.*(k.field) = v;
            if (std::any_of(ms.begin(), ms.end(), [&](const Measurement &m) { return same_plan(m.plan, cand); }))
                continue;
            Measurement m;
            if ((s = batch_measure(b, cand, want, c, m))) return s;
            ++raced;
            ms.push_back(m);
            if (!m.ok) {
                ++res->candidates_rejected;
                continue;
            }
            if (better(m, best)) best = m;
        }
    }
    if ((s = batch_set_plan(b, best.plan))) return s;
    if ((s = persist(c, dev, wl, ms, best))) return s;
    res->winner = from_exec_plan(best.plan);
    res->median_gpu_ns = best.median_gpu;
    res->p95_gpu_ns = best.p95_gpu;
    res->median_submit_ns = best.median_submit;
    res->candidates_measured = raced;
    return AZK_OK;
}
