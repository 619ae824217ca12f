// Serving loop: Circle PCS commits and M31 multilinear sum-checks arriving at any time,
// admitted against an HBM budget, commits run by continuous batching.
//
// No reference counterpart: the reference runs one synchronous command buffer per call
// (MetalContext.swift, CircleCodeword.swift:1007-1018).  Every job's result is the same
// bytes its plain call returns (the batched program is byte-identical per job).
//
//  * commit lane (one thread, one stream): each iteration takes every admitted commit
//    (up to max_batch) and runs them as one table-driven program (graph A eager ->
//    host final-layer step -> graph B eager).  While the device runs A, the lane admits
//    the next arrivals and starts their coefficient uploads on a separate copy stream,
//    so the next iteration starts on resident inputs.  Provers (their layers, trees,
//    twiddles and input staging) stay cached per shape and are reused.
//  * auxiliary lanes (threads, one stream each) run sum-check jobs (the words are hashed
//    on the host, the serial initial digest, while the device rounds run, unless the caller
//    supplied that digest), sum-check chunk jobs and fold-chain jobs.  A lane keeps the plan
//    and buffers of every shape it has run for later jobs of that shape, and gives them
//    back (least recently used first) when another job waits for memory.
//  * commit inputs are QM31 line-basis coefficients or M31 words (the FFT then runs on the
//    M31 words, a quarter of the upload); a batch may hold both kinds.
//  * admission: a job is admitted when the bytes of live device state (cached provers,
//    sum-check plans) plus its own need fit the budget; idle cached state is evicted
//    first when it does not.  A job larger than the whole budget fails with
//    AZK_ERR_ALLOCATION.
#include <condition_variable>
#include <deque>
#include <map>
#include <string>
#include <thread>

#include "prover_internal.hpp"

using namespace azk;
using namespace azk::rt;
using namespace azk::prv;

namespace {

enum { KIND_COMMIT = 0, KIND_SUMCHECK = 1, KIND_FOLD = 2, KIND_CHUNK = 3 };

struct SJob {
    uint64_t ticket = 0;
    uint32_t kind = 0;
    azk_circle_prover_desc desc{};
    const void *input = nullptr;
    int mode = 0;      // commit input: 0 = QM31 line-basis coefficients, 2 = M31 words
    size_t count = 0;  // sum-check words
    uint32_t log_n = 0, rounds = 0;   // sum-check and fold shapes
    std::vector<uint32_t> challenges;  // fold: 4 words per round, copied at submit
    bool have_digest = false;          // sum-check: initial digest supplied by the caller
    uint8_t digest[32] = {};
    size_t need = 0;   // device bytes of the state it needs
    azk_status status = AZK_OK;
    std::string err;
    std::vector<uint8_t> result;
    uint64_t t_submit = 0, t_admit = 0, t_start = 0, t_done = 0;
    uint32_t batch_jobs = 0;
    bool done = false;
    azk_circle_prover *prover = nullptr;
    hipEvent_t uploaded = nullptr;
};

struct ShapeKey {
    azk_circle_prover_desc d;
    bool operator<(const ShapeKey &o) const { return memcmp(&d, &o.d, sizeof d) < 0; }
};

size_t sumcheck_need(uint32_t log_n) { return ((size_t)4 << log_n) * 7 / 4 + 64 * 1024; }
// sum-check chunk plan: the ping-pong vectors and the input staging
size_t chunk_need(uint32_t log_n) { return ((size_t)8 << log_n) + 64 * 1024; }
// fold chain plan: FFT + fold twiddles, two ping-pong layers, input and output buffers
size_t fold_need(uint32_t log_n, uint32_t rounds)
{
    const size_t n = (size_t)1 << log_n;
    return 8 * n + 16 * n + 16 * n + 16 * (n >> rounds) + 64 * 1024;
}

// The device state an auxiliary lane keeps between jobs of the same shape.  A lane keeps one
// per shape it has run (sum-check plans are small next to the budget), so a mix of shapes
// does not free and re-allocate device memory (a hipFree waits for the whole device) on
// every change; under memory pressure the least recently used ones go first.
struct AuxCache {
    uint32_t kind = ~0u, log_n = 0, rounds = 0;
    size_t bytes = 0;
    uint64_t last_use = 0;
    azk_sumcheck_plan *sc = nullptr;
    azk_sumcheck_chunk_plan *chunk = nullptr;
    azk_fold_plan *fold = nullptr;
    azk_buffer *in = nullptr, *out = nullptr;
    bool holds(const SJob *j) const { return kind == j->kind && log_n == j->log_n && rounds == j->rounds; }
    bool empty() const { return kind == ~0u; }
    void release()
    {
        if (sc) azk_sumcheck_plan_destroy(sc);
        if (chunk) azk_sumcheck_chunk_plan_destroy(chunk);
        if (fold) azk_fold_plan_destroy(fold);
        if (in) azk_buffer_free(in);
        if (out) azk_buffer_free(out);
        *this = AuxCache{};
    }
};

struct AuxKey {
    uint32_t kind, log_n, rounds;
    bool operator<(const AuxKey &o) const
    {
        return kind != o.kind ? kind < o.kind : log_n != o.log_n ? log_n < o.log_n : rounds < o.rounds;
    }
};

} // namespace

struct azk_server {
    azk_server_desc d{};
    azk_context *ctx = nullptr;               // commit lane
    std::vector<azk_context *> aux_ctx;        // one per sum-check lane
    hipStream_t upload = nullptr;
    std::mutex mu;
    std::condition_variable cv_commit, cv_aux, cv_done;
    std::deque<SJob *> pend_commit, pend_aux;
    std::map<uint64_t, std::unique_ptr<SJob>> jobs;
    uint64_t next_ticket = 1;
    size_t budget = 0, held = 0;  // held: bytes of live provers and plans
    size_t aux_waiting = 0;        // bytes a blocked sum-check job needs (eviction request)
    bool commit_blocked = false;  // the head commit waits for memory (idle plans are released)
    bool stop = false;
    // commit lane state (commit thread only, except `held` under mu)
    std::multimap<ShapeKey, azk_circle_prover *> idle;
    std::deque<SJob *> ready;  // admitted, upload enqueued
    Runner runner;
    std::thread commit_thr;
    std::vector<std::thread> aux_thr;
};

namespace {

void finish(azk_server *s, SJob *j, azk_status st, const char *msg)
{
    std::lock_guard<std::mutex> g(s->mu);
    j->status = st;
    if (st && msg) j->err = msg;
    j->t_done = now_ns();
    j->done = true;
    s->cv_done.notify_all();
}

// Evicts idle provers until `want` more bytes fit (commit thread).
void evict_idle(azk_server *s, size_t want)
{
    while (!s->idle.empty()) {
        size_t held;
        {
            std::lock_guard<std::mutex> g(s->mu);
            held = s->held;
        }
        if (held + want <= s->budget) return;
        auto it = s->idle.begin();
        azk_circle_prover *p = it->second;
        const size_t fp = prover_footprint(p->d);
        s->idle.erase(it);
        azk_circle_prover_destroy(p);
        std::lock_guard<std::mutex> g(s->mu);
        s->held -= fp;
    }
}

// Admits pending commits (first fit in arrival order) and enqueues their uploads, up to one
// device program ahead: the jobs of the next batch get their provers and uploads while the
// current one runs, and later arrivals wait unadmitted (no provers built or bytes moved for
// jobs that cannot start yet).
void admit_commits(azk_server *s)
{
    while (s->ready.size() < s->d.max_batch) {
        SJob *j = nullptr;
        azk_circle_prover *p = nullptr;
        bool create = false;
        {
            std::lock_guard<std::mutex> g(s->mu);
            for (auto it = s->pend_commit.begin(); it != s->pend_commit.end(); ++it) {
                SJob *c = *it;
                auto idle = s->idle.find(ShapeKey{c->desc});
                if (idle != s->idle.end()) {
                    p = idle->second;
                    s->idle.erase(idle);
                } else if (s->held + c->need <= s->budget) {
                    s->held += c->need;  // reserved; the prover is built below
                    create = true;
                } else {
                    continue;
                }
                j = c;
                s->pend_commit.erase(it);
                break;
            }
        }
        if (!j) return;
        if (create) {
            azk_status st = prover_create(s->ctx, j->desc, &p);
            if (st) {
                std::string m = azk_last_error();
                {
                    std::lock_guard<std::mutex> g(s->mu);
                    s->held -= j->need;
                }
                finish(s, j, st, m.c_str());
                continue;
            }
        }
        j->prover = p;
        const size_t bytes = (size_t)(j->mode == 2 ? 4 : 16) << j->desc.log_n;
        azk_status st = p->coeff_staging.ensure(bytes);
        hipError_t e = hipSuccess;
        if (!st) e = hipMemcpyAsync(p->coeff_staging.p, j->input, bytes, hipMemcpyHostToDevice, s->upload);
        if (!st && e == hipSuccess && !j->uploaded) e = hipEventCreateWithFlags(&j->uploaded, hipEventDisableTiming);
        if (!st && e == hipSuccess) e = hipEventRecord(j->uploaded, s->upload);
        if (st || e != hipSuccess) {
            s->idle.emplace(ShapeKey{j->desc}, p);
            j->prover = nullptr;
            finish(s, j, st ? st : AZK_ERR_COMMAND_FAILED, st ? azk_last_error() : hipGetErrorString(e));
            continue;
        }
        j->t_admit = now_ns();
        s->ready.push_back(j);
    }
}

azk_status run_batch(azk_server *s, std::vector<SJob *> &batch)
{
    hipStream_t st = s->ctx->stream;
    std::vector<Job> list;
    for (SJob *j : batch) list.push_back(Job{j->prover, j->prover->coeff_staging.p, j->mode});
    runner_reset(s->runner);
    azk_status r = runner_build(s->runner, s->ctx, list);
    if (r) return r;
    for (SJob *j : batch) HIP_TRY(hipStreamWaitEvent(st, j->uploaded, 0));
    const uint64_t t0 = now_ns();
    for (SJob *j : batch) j->t_start = t0;
    PhaseClock none(st, false);
    r = enqueue_a(s->runner, st, none);
    if (r) return r;
    admit_commits(s);  // the next batch's uploads overlap this one
    HIP_TRY(hipStreamSynchronize(st));
    r = host_finals(s->runner);
    if (r) return r;
    r = enqueue_b(s->runner, st);
    if (r) return r;
    HIP_TRY(hipStreamSynchronize(st));
    patch_nonces(s->runner);
    for (size_t k = 0; k < batch.size(); ++k) {
        SJob *j = batch[k];
        const uint8_t *pr = proof_of(s->runner, k);
        j->result.assign(pr, pr + 4 * j->prover->proof_words);
    }
    return AZK_OK;
}

void commit_lane(azk_server *s)
{
    (void)hipSetDevice(s->d.device);
    for (;;) {
        {
            std::unique_lock<std::mutex> g(s->mu);
            s->cv_commit.wait(g, [&] { return s->stop || !s->pend_commit.empty() || !s->ready.empty() || s->aux_waiting; });
            if (s->stop && s->pend_commit.empty() && s->ready.empty()) break;
        }
        admit_commits(s);
        size_t sc_want;
        {
            std::lock_guard<std::mutex> g(s->mu);
            sc_want = s->aux_waiting;
        }
        if (sc_want) {
            evict_idle(s, sc_want);
            std::lock_guard<std::mutex> g(s->mu);
            s->aux_waiting = 0;  // served as far as idle provers allow; the lane asks again if still blocked
        }
        if (s->ready.empty()) {
            // Nothing admissible: free idle provers for the head job, or fail it if it can
            // never fit.
            SJob *head = nullptr;
            {
                std::lock_guard<std::mutex> g(s->mu);
                if (!s->pend_commit.empty()) head = s->pend_commit.front();
            }
            if (!head) continue;
            evict_idle(s, head->need);
            admit_commits(s);
            if (s->ready.empty()) {
                bool never;
                {
                    std::lock_guard<std::mutex> g(s->mu);
                    never = head->need > s->budget;
                    if (never) s->pend_commit.pop_front();
                }
                if (never) {
                    finish(s, head, AZK_ERR_ALLOCATION, "job exceeds the server's HBM budget");
                } else {
                    std::unique_lock<std::mutex> g(s->mu);  // sum-check lanes hold the memory
                    s->commit_blocked = true;
                    s->cv_aux.notify_all();
                    s->cv_commit.wait_for(g, std::chrono::milliseconds(1));
                }
                continue;
            }
        }
        {
            std::lock_guard<std::mutex> g(s->mu);
            s->commit_blocked = false;
        }
        std::vector<SJob *> batch;
        while (!s->ready.empty() && batch.size() < s->d.max_batch) {
            batch.push_back(s->ready.front());
            s->ready.pop_front();
        }
        azk_status r = run_batch(s, batch);
        std::string m = r ? azk_last_error() : "";
        for (SJob *j : batch) {
            j->batch_jobs = (uint32_t)batch.size();
            s->idle.emplace(ShapeKey{j->desc}, j->prover);
            j->prover = nullptr;
            finish(s, j, r, m.c_str());
        }
    }
}

// Runs one auxiliary job with the lane's cached state (built when the shape changes).
azk_status run_aux(azk_server *s, size_t lane, AuxCache &c, SJob *j)
{
    azk_context *ctx = s->aux_ctx[lane];
    azk_status st = AZK_OK;
    if (!c.holds(j)) {
        c.release();
        c.kind = j->kind;
        c.log_n = j->log_n;
        c.rounds = j->rounds;
        c.bytes = j->need;
        if (j->kind == KIND_SUMCHECK) {
            st = azk_sumcheck_plan_create(ctx, j->log_n, &c.sc);
        } else if (j->kind == KIND_CHUNK) {
            st = azk_sumcheck_chunk_plan_create(ctx, (uint64_t)1 << j->log_n, j->rounds, &c.chunk);
        } else {
            st = azk_fold_plan_create(ctx, j->log_n, j->rounds, &c.fold);
            if (!st) st = azk_buffer_alloc(ctx, (size_t)16 << j->log_n, &c.in);
            if (!st) st = azk_buffer_alloc(ctx, (size_t)16 << (j->log_n - j->rounds), &c.out);
        }
        if (st) {
            const std::string m = azk_last_error();
            c.release();
            return fail(st, "%s", m.c_str());
        }
    }
    j->t_start = now_ns();
    if (j->kind == KIND_SUMCHECK) {
        j->result.resize(azk_sumcheck_proof_size(j->log_n));
        size_t len = 0;
        st = azk_sumcheck_prove_with_digest(c.sc, (const uint32_t *)j->input, j->count, j->have_digest ? j->digest : nullptr,
                                            nullptr, j->result.data(), j->result.size(), &len, nullptr);
        j->result.resize(st ? 0 : len);
        return st;
    }
    if (j->kind == KIND_CHUNK) {
        j->result.resize(azk_sumcheck_chunk_proof_size(j->count, j->rounds));
        size_t len = 0;
        st = azk_sumcheck_chunk_prove(c.chunk, (const uint32_t *)j->input, j->count, j->result.data(), j->result.size(),
                                      &len, nullptr);
        j->result.resize(st ? 0 : len);
        return st;
    }
    const size_t ib = (size_t)16 << j->log_n, ob = (size_t)16 << (j->log_n - j->rounds);
    st = azk_buffer_upload(c.in, 0, j->input, ib);
    if (!st) st = azk_fold_execute_resident(c.fold, c.in, 0, j->challenges.data(), c.out, 0, nullptr);
    if (!st) {
        j->result.resize(ob);
        st = azk_buffer_download(c.out, 0, j->result.data(), ob);
    }
    return st;
}

AuxKey key_of(const SJob *j) { return AuxKey{j->kind, j->log_n, j->rounds}; }

// The lane's least recently used cached state other than `keep` (end() if none).
std::map<AuxKey, AuxCache>::iterator lru_victim(std::map<AuxKey, AuxCache> &caches, const AuxKey *keep)
{
    auto best = caches.end();
    for (auto it = caches.begin(); it != caches.end(); ++it) {
        if (keep && !(it->first < *keep) && !(*keep < it->first)) continue;
        if (best == caches.end() || it->second.last_use < best->second.last_use) best = it;
    }
    return best;
}

void aux_lane(azk_server *s, size_t lane)
{
    (void)hipSetDevice(s->d.device);
    std::map<AuxKey, AuxCache> caches;  // this lane's only: read and changed by this thread alone
    uint64_t uses = 0;
    // Frees one cached state (outside the server lock) and gives its bytes back.
    auto drop = [&](std::unique_lock<std::mutex> &g, std::map<AuxKey, AuxCache>::iterator it) {
        const size_t b = it->second.bytes;
        g.unlock();
        it->second.release();
        caches.erase(it);
        g.lock();
        s->held -= b;
    };
    for (;;) {
        SJob *j = nullptr;
        {
            std::unique_lock<std::mutex> g(s->mu);
            for (;;) {
                if (!s->pend_aux.empty()) {
                    SJob *c = s->pend_aux.front();
                    const AuxKey k = key_of(c);
                    const bool same = caches.count(k) != 0;
                    const size_t add = same ? 0 : c->need;
                    if (s->held + add <= s->budget || c->need > s->budget) {
                        j = c;
                        s->pend_aux.pop_front();
                        if (s->aux_waiting == c->need) s->aux_waiting = 0;
                        if (c->need <= s->budget) s->held += add;
                        break;
                    }
                    auto victim = lru_victim(caches, &k);
                    if (victim != caches.end()) {  // this lane's other shapes go first
                        drop(g, victim);
                        continue;
                    }
                    s->aux_waiting = c->need;  // ask the commit lane to evict idle provers
                    s->cv_commit.notify_one();
                    s->cv_aux.wait_for(g, std::chrono::milliseconds(1));
                    continue;
                }
                if (s->stop) break;
                if (!caches.empty() && (s->commit_blocked || s->aux_waiting)) {
                    // idle while another job waits for memory: give the cached bytes back
                    drop(g, lru_victim(caches, nullptr));
                    continue;
                }
                if (!caches.empty()) s->cv_aux.wait_for(g, std::chrono::milliseconds(2));
                else s->cv_aux.wait(g);
            }
        }
        if (!j) break;
        if (j->need > s->budget) {
            finish(s, j, AZK_ERR_ALLOCATION, "job exceeds the server's HBM budget");
            continue;
        }
        j->t_admit = now_ns();
        const AuxKey k = key_of(j);
        AuxCache &c = caches[k];
        const bool had = !c.empty();
        azk_status st = run_aux(s, lane, c, j);
        c.last_use = ++uses;
        if (c.empty()) {  // not built: the reservation goes back
            caches.erase(k);
            if (!had) {
                std::lock_guard<std::mutex> g(s->mu);
                s->held -= j->need;
            }
        }
        j->batch_jobs = 1;
        std::string m = st ? azk_last_error() : "";
        finish(s, j, st, m.c_str());
    }
    size_t b = 0;
    for (auto &kv : caches) {
        b += kv.second.bytes;
        kv.second.release();
    }
    std::lock_guard<std::mutex> g(s->mu);
    s->held -= b;
}

} // namespace

namespace {

// mode 0: `bytes` of QM31 line-basis coefficients (16 per element); mode 2: M31 words (4 each).
azk_status submit_commit(azk_server *s, const azk_circle_prover_desc *desc, const void *coeffs, size_t bytes, int mode,
                         uint64_t *ticket)
{
    azk_status st = validate_desc(*desc);
    if (st) return st;
    const size_t want = (size_t)(mode == 2 ? 4 : 16) << desc->log_n;
    if (bytes < want) return fail(AZK_ERR_INVALID_LAYOUT, "coefficient bytes %zu < %zu", bytes, want);
    auto j = std::make_unique<SJob>();
    j->kind = KIND_COMMIT;
    j->desc = *desc;
    j->input = coeffs;
    j->mode = mode;
    j->need = prover_footprint(*desc);
    j->t_submit = now_ns();
    {
        std::lock_guard<std::mutex> g(s->mu);
        if (s->stop) return fail(AZK_ERR_INVALID_ARGUMENT, "server is stopping");
        j->ticket = s->next_ticket++;
        *ticket = j->ticket;
        s->pend_commit.push_back(j.get());
        s->jobs.emplace(j->ticket, std::move(j));
    }
    s->cv_commit.notify_one();
    return AZK_OK;
}

azk_status submit_aux(azk_server *s, std::unique_ptr<SJob> j, uint64_t *ticket)
{
    j->t_submit = now_ns();
    {
        std::lock_guard<std::mutex> g(s->mu);
        if (s->stop) return fail(AZK_ERR_INVALID_ARGUMENT, "server is stopping");
        j->ticket = s->next_ticket++;
        *ticket = j->ticket;
        s->pend_aux.push_back(j.get());
        s->jobs.emplace(j->ticket, std::move(j));
    }
    s->cv_aux.notify_one();
    return AZK_OK;
}

std::unique_ptr<SJob> sumcheck_job(const uint32_t *evals, size_t count, const uint8_t *digest)
{
    auto j = std::make_unique<SJob>();
    j->kind = KIND_SUMCHECK;
    j->input = evals;
    j->count = count;
    j->log_n = ilog2(count);
    j->need = sumcheck_need(j->log_n);
    if (digest) {
        j->have_digest = true;
        memcpy(j->digest, digest, 32);
    }
    return j;
}

} // namespace

extern "C" {

uint64_t azk_now_ns(void) { return now_ns(); }

azk_status azk_host_alloc(size_t bytes, void **out)
{
    CHECK_ARG(out);
    *out = nullptr;
    if (hipHostMalloc(out, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess)
        return fail(AZK_ERR_ALLOCATION, "hipHostMalloc(%zu) failed", bytes);
    return AZK_OK;
}

void azk_host_free(void *p)
{
    if (p) (void)hipHostFree(p);
}

azk_status azk_server_create(const azk_server_desc *desc, azk_server **out)
{
    CHECK_ARG(desc && out);
    auto *s = new azk_server();
    s->d = *desc;
    if (!s->d.max_batch) s->d.max_batch = 64;
    if (!s->d.aux_lanes) s->d.aux_lanes = 4;
    if (s->d.aux_lanes > 16) s->d.aux_lanes = 16;
    azk_status st = azk_context_create(s->d.device, &s->ctx);
    if (st) { delete s; return st; }
    s->budget = s->d.hbm_budget;
    if (!s->budget) {
        size_t free_b = 0, total = 0;
        if (hipMemGetInfo(&free_b, &total) != hipSuccess) free_b = (size_t)8 << 30;
        s->budget = free_b / 2;
    }
    for (uint32_t k = 0; k < s->d.aux_lanes && !st; ++k) {
        azk_context *c = nullptr;
        st = azk_context_create(s->d.device, &c);
        if (!st) s->aux_ctx.push_back(c);
    }
    if (!st && hipStreamCreateWithFlags(&s->upload, hipStreamNonBlocking) != hipSuccess)
        st = fail(AZK_ERR_ALLOCATION, "hipStreamCreate failed");
    if (st) {
        for (auto *c : s->aux_ctx) azk_context_destroy(c);
        azk_context_destroy(s->ctx);
        delete s;
        return st;
    }
    s->commit_thr = std::thread(commit_lane, s);
    for (size_t k = 0; k < s->aux_ctx.size(); ++k) s->aux_thr.emplace_back(aux_lane, s, k);
    *out = s;
    return AZK_OK;
}

void azk_server_destroy(azk_server *s)
{
    if (!s) return;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
    }
    s->cv_commit.notify_all();
    s->cv_aux.notify_all();
    s->commit_thr.join();
    for (auto &t : s->aux_thr) t.join();
    (void)hipSetDevice(s->d.device);
    (void)hipStreamSynchronize(s->upload);
    for (auto &kv : s->idle) azk_circle_prover_destroy(kv.second);
    for (auto &kv : s->jobs)
        if (kv.second->uploaded) (void)hipEventDestroy(kv.second->uploaded);
    (void)hipStreamDestroy(s->upload);
    for (auto *c : s->aux_ctx) azk_context_destroy(c);
    azk_context *ctx = s->ctx;
    delete s;  // the Runner's staging and streams go first
    azk_context_destroy(ctx);
}

azk_status azk_server_submit_commit(azk_server *s, const azk_circle_prover_desc *desc, const void *coeffs, size_t bytes,
                                    uint64_t *ticket)
{
    CHECK_ARG(s && desc && coeffs && ticket);
    return submit_commit(s, desc, coeffs, bytes, 0, ticket);
}

azk_status azk_server_submit_commit_m31(azk_server *s, const azk_circle_prover_desc *desc, const uint32_t *coeffs,
                                        size_t words, uint64_t *ticket)
{
    CHECK_ARG(s && desc && coeffs && ticket);
    if (words > SIZE_MAX / 4) return fail(AZK_ERR_INVALID_LAYOUT, "coefficient word count %zu", words);
    return submit_commit(s, desc, coeffs, words * 4, 2, ticket);
}

azk_status azk_server_submit_sumcheck_digest(azk_server *s, const uint32_t *evals, size_t count, const uint8_t digest[32],
                                             uint64_t *ticket)
{
    CHECK_ARG(s && evals && digest && ticket);
    if (count < 2 || (count & (count - 1)) || count > ((size_t)1 << 30))
        return fail(AZK_ERR_INVALID_LAYOUT, "sum-check lane count %zu is not a power of two in [2, 2^30]", count);
    return submit_aux(s, sumcheck_job(evals, count, digest), ticket);
}

azk_status azk_server_submit_sumcheck_chunk(azk_server *s, const uint32_t *evals, size_t count, uint32_t rounds,
                                            uint64_t *ticket)
{
    CHECK_ARG(s && evals && ticket);
    if (count < 2 || (count & (count - 1)) || count > ((size_t)1 << 30) || rounds < 1 || rounds > ilog2(count))
        return fail(AZK_ERR_INVALID_LAYOUT, "sum-check chunk: %zu lanes (power of two, 2..2^30), %u rounds", count, rounds);
    auto j = std::make_unique<SJob>();
    j->kind = KIND_CHUNK;
    j->input = evals;
    j->count = count;
    j->log_n = ilog2(count);
    j->rounds = rounds;
    j->need = chunk_need(j->log_n);
    return submit_aux(s, std::move(j), ticket);
}

azk_status azk_server_submit_sumcheck(azk_server *s, const uint32_t *evals, size_t count, uint64_t *ticket)
{
    CHECK_ARG(s && evals && ticket);
    if (count < 2 || (count & (count - 1)) || count > ((size_t)1 << 30))
        return fail(AZK_ERR_INVALID_LAYOUT, "sum-check lane count %zu is not a power of two in [2, 2^30]", count);
    return submit_aux(s, sumcheck_job(evals, count, nullptr), ticket);
}

azk_status azk_server_submit_fold(azk_server *s, uint32_t log_n, uint32_t rounds, const void *evals, size_t bytes,
                                  const uint32_t *challenges, uint64_t *ticket)
{
    CHECK_ARG(s && evals && challenges && ticket);
    if (log_n < 1 || log_n > 28 || rounds < 1 || rounds > log_n || rounds > 30)
        return fail(AZK_ERR_INVALID_LAYOUT, "fold chain log_n %u rounds %u", log_n, rounds);
    if (bytes < ((size_t)16 << log_n)) return fail(AZK_ERR_INVALID_LAYOUT, "evaluation bytes %zu < %zu", bytes, (size_t)16 << log_n);
    for (uint32_t i = 0; i < 4 * rounds; ++i)
        if (challenges[i] >= kP) return fail(AZK_ERR_INVALID_LAYOUT, "non-canonical challenge limb");
    auto j = std::make_unique<SJob>();
    j->kind = KIND_FOLD;
    j->input = evals;
    j->log_n = log_n;
    j->rounds = rounds;
    j->challenges.assign(challenges, challenges + 4 * rounds);
    j->need = fold_need(log_n, rounds);
    return submit_aux(s, std::move(j), ticket);
}

azk_status azk_server_wait(azk_server *s, uint64_t ticket, azk_job_info *info)
{
    CHECK_ARG(s && info);
    std::unique_lock<std::mutex> g(s->mu);
    auto it = s->jobs.find(ticket);
    if (it == s->jobs.end()) return fail(AZK_ERR_INVALID_ARGUMENT, "unknown ticket %llu", (unsigned long long)ticket);
    SJob *j = it->second.get();
    s->cv_done.wait(g, [&] { return j->done; });
    info->status = j->status;
    info->kind = j->kind;
    info->batch_jobs = j->batch_jobs;
    info->result_bytes = j->result.size();
    info->submit_ns = j->t_submit;
    info->admit_ns = j->t_admit;
    info->start_ns = j->t_start;
    info->done_ns = j->t_done;
    if (j->status) return fail(j->status, "job %llu: %s", (unsigned long long)ticket, j->err.c_str());
    return AZK_OK;
}

azk_status azk_server_take(azk_server *s, uint64_t ticket, uint8_t *out, size_t cap, size_t *len)
{
    CHECK_ARG(s);
    std::unique_lock<std::mutex> g(s->mu);
    auto it = s->jobs.find(ticket);
    if (it == s->jobs.end()) return fail(AZK_ERR_INVALID_ARGUMENT, "unknown ticket %llu", (unsigned long long)ticket);
    SJob *j = it->second.get();
    s->cv_done.wait(g, [&] { return j->done; });
    if (out && j->status == AZK_OK) {
        if (cap < j->result.size()) return fail(AZK_ERR_INVALID_LAYOUT, "result capacity %zu < %zu", cap, j->result.size());
        memcpy(out, j->result.data(), j->result.size());
    }
    if (len) *len = j->result.size();
    if (j->uploaded) (void)hipEventDestroy(j->uploaded);
    s->jobs.erase(it);
    return AZK_OK;
}

} // extern "C"
