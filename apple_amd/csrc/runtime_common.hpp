// Shared host-side pieces of the runtime: errors, device memory, timing, layout checks,
// the context/buffer objects and the Circle FFT pass schedule.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/azk.h"
#include "kernels.hpp"

#include <atomic>

struct azk_context {
    int device = 0;
    hipStream_t stream = nullptr;
    hipDeviceProp_t props{};
    std::mutex mu;  // serializes use of `stream` across plans
    // One reference for the owner (dropped by azk_context_destroy) and one per object
    // built on the context (buffers, plans, provers): the context outlives them all.
    std::atomic<int> refs{1};
};

struct azk_buffer {
    azk_context *ctx = nullptr;
    void *ptr = nullptr;
    size_t bytes = 0;
    bool owned = false;
};

namespace azk::rt {

azk_status fail(azk_status s, const char *fmt, ...);

void ctx_retain(azk_context *c);
void ctx_release(azk_context *c);
// Deletes an object built on a context, then drops its reference to the context.
template <class T>
void destroy_on_ctx(T *p)
{
    if (!p) return;
    azk_context *c = p->ctx;
    delete p;
    ctx_release(c);
}

#define HIP_TRY(expr)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess) return ::azk::rt::fail(AZK_ERR_COMMAND_FAILED, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

#define CHECK_ARG(c)                                                                               \
    do {                                                                                           \
        if (!(c)) return ::azk::rt::fail(AZK_ERR_INVALID_ARGUMENT, "invalid argument: %s", #c);    \
    } while (0)

inline uint64_t now_ns()
{
    return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}
inline bool is_pow2(uint64_t v) { return v && !(v & (v - 1)); }
inline uint32_t ilog2(uint64_t v)
{
    uint32_t r = 0;
    while (v > 1) { v >>= 1; ++r; }
    return r;
}
inline size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

constexpr uint32_t kP = 0x7fffffffu;
constexpr uint32_t kDefaultLogTile = 11;    // 2048-element tiles: 40 KB LDS (final FFT pass), 32 KB (folds)
// Final-pass tiles of a codeword that is not hashed: 20 KB of LDS, eight blocks (eight waves
// per SIMD) per CU.  At 2^26 the final pass takes 0.59-0.65 ms against 0.68-0.75 ms with
// 2^11-element tiles (four waves per SIMD); the passes above take 16 stages instead of 15.
constexpr uint32_t kFftOnlyLogTile = 10;
constexpr uint32_t kUpperLogChunk = 10;     // largest upper-tree reduction chunk
// Commit trees: 8-node chunks, so every level of a step runs in the one-state-per-wave
// form with one parent per wave (a 2^21 commit measured 2% faster than 16-node chunks).
constexpr uint32_t kCommitLogTreeChunk = 3;
constexpr uint32_t kLeafwiseLogTile = 8;    // latency form: one leaf per thread, 256 per block
constexpr uint32_t kLeafwiseMinLogTile = 3;
// Leafwise tiles of 2^(k - 8) leaves for a 2^k layer (at least 8, at most 256): small layers
// keep only levels of a few parents inside the fold block, large ones keep 256-leaf tiles for
// throughput.  Measured on 2^21 commits with the two-lanes-per-state permutation: grid 8
// 2.085 ms, 7 2.095, 9 2.110, fixed 256-leaf tiles 2.104.
constexpr uint32_t kLeafwiseLogGrid = 8;

struct DeviceMem {
    void *p = nullptr;
    size_t n = 0;
    DeviceMem() = default;
    DeviceMem(const DeviceMem &) = delete;
    DeviceMem &operator=(const DeviceMem &) = delete;
    ~DeviceMem() { if (p) (void)hipFree(p); }
    azk_status ensure(size_t bytes)
    {
        if (bytes <= n && p) return AZK_OK;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess) return fail(AZK_ERR_ALLOCATION, "hipMalloc(%zu) failed", bytes);
        n = bytes;
        return AZK_OK;
    }
};

struct PinnedMem {
    void *p = nullptr;
    size_t n = 0;
    PinnedMem() = default;
    PinnedMem(const PinnedMem &) = delete;
    PinnedMem &operator=(const PinnedMem &) = delete;
    ~PinnedMem() { if (p) (void)hipHostFree(p); }
    azk_status ensure(size_t bytes)
    {
        if (bytes <= n && p) return AZK_OK;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        n = 0;
        if (hipHostMalloc(&p, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess)
            return fail(AZK_ERR_ALLOCATION, "hipHostMalloc(%zu) failed", bytes);
        n = bytes;
        return AZK_OK;
    }
};

struct Timer {
    hipEvent_t a = nullptr, b = nullptr;
    ~Timer()
    {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
    azk_status init()
    {
        if (!a) HIP_TRY(hipEventCreate(&a));
        if (!b) HIP_TRY(hipEventCreate(&b));
        return AZK_OK;
    }
    uint64_t elapsed_ns()
    {
        float ms = 0;
        if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0;
        return (uint64_t)(ms * 1e6);
    }
};

// Range check of a (buffer, offset, bytes) triple: validateBufferRange,
// CircleCodeword.swift:845-857.
azk_status check_range(const azk_buffer *b, size_t off, size_t bytes);
// Aliasing of two ranges: rangesOverlap, CircleCodeword.swift:863-879.
bool overlaps(const azk_buffer *a, size_t ao, size_t an, const azk_buffer *b, size_t bo, size_t bn);

// Runs `enqueue` on the context stream bracketed by HIP events and waits.
template <class F>
azk_status run_timed(azk_context *ctx, azk_stats *stats, F &&enqueue)
{
    Timer t;
    azk_status s = t.init();
    if (s) return s;
    uint64_t w0 = now_ns();
    HIP_TRY(hipEventRecord(t.a, ctx->stream));
    s = enqueue();
    if (s) return s;
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(t.b, ctx->stream));
    uint64_t w1 = now_ns();
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (stats) {
        stats->submit_ns = w1 - w0;
        stats->wall_ns = now_ns() - w0;
        stats->gpu_ns = t.elapsed_ns();
    }
    return AZK_OK;
}

struct FftPass { uint32_t lo, k, m; };

// Stages log_n-1 .. lt (lt = min(log_t_final, log_n)) split into the fewest balanced
// passes of at most 10 stages.  A QM31 pass of k stages runs on tiles of
// 2^min(13, max(12, k + 4)) elements: 2^k strided rows x >= 8 contiguous columns, 16 (256
// bytes) below the cap -- at 2^26 the first pass takes 0.50 ms against 0.59 ms with 128-byte
// rows -- and the 10 stages above a 2^21 codeword's final pass take one 2^13-element pass
// (132 KB of LDS, one block per CU) instead of two passes over HBM.  An M31 pass (m31)
// runs on 2^min(14, max(13, k + 6)) words.
constexpr uint32_t kMaxPassLogTile = 13, kMaxPassLogTileM31 = 14;
std::vector<FftPass> fft_passes(uint32_t log_n, uint32_t log_t_final, bool m31 = false);

// FFT twiddles (and, with `fold_tw`, the first `rounds` stages inverted = FRI fold twiddles).
azk_status materialize_twiddles(azk_context *ctx, uint32_t log_n, uint32_t rounds, DeviceMem &fft_tw, DeviceMem *fold_tw);

// Descriptor tables of many launches packed into one device buffer.
struct TableArena {
    std::vector<uint8_t> host;
    DeviceMem dev;
    template <class D>
    size_t put(const std::vector<D> &v)
    {
        size_t off = align_up(host.size(), 64);
        host.resize(off + v.size() * sizeof(D));
        if (!v.empty()) memcpy(host.data() + off, v.data(), v.size() * sizeof(D));
        return off;
    }
    azk_status upload()
    {
        azk_status s = dev.ensure(std::max<size_t>(host.size(), 64));
        if (s) return s;
        if (!host.empty()) HIP_TRY(hipMemcpy(dev.p, host.data(), host.size(), hipMemcpyHostToDevice));
        return AZK_OK;
    }
    template <class D>
    const D *at(size_t off) const { return (const D *)((const uint8_t *)dev.p + off); }
};

// One table-driven launch: descriptors (first_block assigned in order) and blocks.
template <class D>
struct Launch {
    std::vector<D> v;
    uint32_t blocks = 0;
    uint32_t max_log = 0;  // largest tile / chunk log of the entries (sets LDS)
    size_t off = 0;
    void add(D d, uint32_t nblocks, uint32_t log_size = 0)
    {
        if (!nblocks) return;
        d.first_block = blocks;
        v.push_back(d);
        blocks += nblocks;
        max_log = std::max(max_log, log_size);
    }
    bool empty() const { return v.empty(); }
    uint32_t n() const { return (uint32_t)v.size(); }
};

// Whether some final-FFT entry of a launch hashes its tile.
inline bool any_tree(const Launch<FftFinalDesc> &l)
{
    for (const FftFinalDesc &d : l.v)
        if (d.tree) return true;
    return false;
}

// Upper-tree schedule of one tree: launches of <= 2^max_log_c-node chunks from `level`
// to the root; step k of several trees can share one launch.
struct TreeStep { uint32_t level, log_c; uint64_t chunks; };
std::vector<TreeStep> tree_schedule(uint64_t m, uint32_t level, uint32_t max_log_c);

// Tile shapes of the commit program (azk_circle_exec_plan); the defaults are the
// shapes used when no tuned plan is set.
struct ExecPlan {
    uint32_t log_tile_commit = kDefaultLogTile;    // final FFT pass + first-layer treelets
    uint32_t log_tile_fold = kDefaultLogTile;      // committed folds, big-tile form
    uint32_t log_leafwise_below = 19;              // layers < 2^this: one leaf per thread
    uint32_t log_tile_leafwise = kLeafwiseLogTile; // latency-form tile
    uint32_t log_tree_chunk = kCommitLogTreeChunk; // upper-tree reduction chunk
    uint32_t log_leafwise_grid = kLeafwiseLogGrid; // leafwise tiles shrink to keep ~2^this blocks
    uint32_t tree_waves = 4;                       // waves per upper-tree block (4, 8, 16)
};
azk_status check_exec_plan(const azk_circle_exec_plan &x);
ExecPlan to_exec_plan(const azk_circle_exec_plan &x);
azk_circle_exec_plan from_exec_plan(const ExecPlan &x);

// Appends the FFT of one job (src -> dst, optional leaf-hash tree) to pass/final launches.
// m31: src is N M31 words (the coefficients' first components; the QM31 lift is (a, 0, 0, 0)),
// the passes work in `scratch` (N M31 words, dst if null) and the final pass writes dst:
// the QM31 first layer with a tree, N M31 words without.  The launches take only one kind.
void plan_fft(std::vector<Launch<FftPassDesc>> &passes, Launch<FftFinalDesc> &final_, const void *src, void *dst,
              const uint32_t *tw, uint32_t log_n, uint32_t log_t, bool run_fft, uint32_t *tree, bool m31 = false,
              void *scratch = nullptr);

// Appends one fold round of one job to the big-tile or latency-form launch; returns the
// log tile size (the level whose nodes remain for the upper tree).
uint32_t plan_fold(Launch<FoldDesc> &big, Launch<FoldDesc> &leaf, const void *in, void *out, const uint32_t *tw,
                   const uint32_t *ch, uint64_t n_out, uint32_t log_t, uint32_t *tree,
                   uint32_t log_leafwise_below = 19, uint32_t log_t_leaf = kLeafwiseLogTile,
                   uint32_t log_grid = kLeafwiseLogGrid);

// Appends the upper-tree steps of one tree.  With `after`, the block of the last step
// runs that transcript program once the root is written (returns false, and runs
// nothing, when the tree needs no step: the caller then launches the program itself).
bool plan_tree(std::vector<Launch<TreeDesc>> &steps, uint32_t *tree, uint64_t m, uint32_t level,
               uint32_t log_chunk = kUpperLogChunk, const TranscriptDesc *after = nullptr);

// Default raw-leaf treelet of a 2^log_m-leaf commitment.
uint32_t merkle_default_log_tile(uint32_t log_m);

// Leaf tiles, then the upper tree, of one raw-leaf commitment into `tree` (all 2m - 1 nodes,
// level by level).  Every launch takes its descriptor by value: nothing is uploaded,
// allocated or waited for.
void merkle_launch(hipStream_t st, const void *leaves, uint32_t *tree, uint64_t m, uint32_t stride, uint32_t len,
                   uint32_t log_t);

} // namespace azk::rt
