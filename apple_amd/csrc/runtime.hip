// Host runtime behind include/azk.h: context, resident buffers, reusable plans, the
// Circle PCS/FRI prover superstep (captured once per shape as a hipGraph and replayed),
// and the batched multi-job form.
//
// Plan semantics mirror the reference's Swift plans (cited per entry point in azk.h);
// buffers, streams and graphs are HIP-native.
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
#include "host_sha3.hpp"
#include "kernels.hpp"

using namespace azk;

namespace {

thread_local std::string g_err;

azk_status fail(azk_status s, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return s;
}

#define HIP_TRY(expr)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess) return fail(AZK_ERR_COMMAND_FAILED, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

#define CHECK_ARG(c)                                                                               \
    do {                                                                                           \
        if (!(c)) return fail(AZK_ERR_INVALID_ARGUMENT, "invalid argument: %s", #c);               \
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
constexpr uint32_t kDefaultLogTile = 11;   // 2048-element tiles: 48 KB LDS per block
constexpr uint32_t kUpperLogChunk = 10;    // upper-tree reduction chunk

} // namespace

// ====================================================================== context

struct azk_context {
    int device = 0;
    hipStream_t stream = nullptr;
    hipDeviceProp_t props{};
    std::mutex mu;  // serializes use of `stream` across plans
};

struct azk_buffer {
    azk_context *ctx = nullptr;
    void *ptr = nullptr;
    size_t bytes = 0;
    bool owned = false;
};

namespace {

// Range check of a (buffer, offset, bytes) triple: validateBufferRange,
// CircleCodeword.swift:845-857.
azk_status check_range(const azk_buffer *b, size_t off, size_t bytes)
{
    if (!b) return fail(AZK_ERR_INVALID_ARGUMENT, "null buffer");
    if (off > b->bytes || bytes > b->bytes - off)
        return fail(AZK_ERR_INVALID_LAYOUT, "range [%zu, +%zu) exceeds buffer of %zu bytes", off, bytes, b->bytes);
    return AZK_OK;
}

// Aliasing of two ranges: rangesOverlap, CircleCodeword.swift:863-879.
bool overlaps(const azk_buffer *a, size_t ao, size_t an, const azk_buffer *b, size_t bo, size_t bn)
{
    if (!a || !b || !an || !bn) return false;
    uintptr_t a0 = (uintptr_t)a->ptr + ao, b0 = (uintptr_t)b->ptr + bo;
    return a0 < b0 + bn && b0 < a0 + an;
}

struct DeviceMem {
    void *p = nullptr;
    size_t n = 0;
    ~DeviceMem() { if (p) (void)hipFree(p); }
    azk_status ensure(size_t bytes)
    {
        if (bytes <= n) return AZK_OK;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess) return fail(AZK_ERR_ALLOCATION, "hipMalloc(%zu) failed", bytes);
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

} // namespace

extern "C" {

const char *azk_last_error(void) { return g_err.c_str(); }
const char *azk_version(void) { return "azk 0.1 gfx950"; }

azk_status azk_context_create(int device, azk_context **out)
{
    CHECK_ARG(out);
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return fail(AZK_ERR_NO_DEVICE, "no HIP device");
    if (device < 0 || device >= count) return fail(AZK_ERR_NO_DEVICE, "device %d of %d", device, count);
    auto *c = new azk_context();
    c->device = device;
    HIP_TRY(hipSetDevice(device));
    HIP_TRY(hipGetDeviceProperties(&c->props, device));
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return fail(AZK_ERR_ALLOCATION, "hipStreamCreate failed");
    }
    *out = c;
    return AZK_OK;
}

void azk_context_destroy(azk_context *ctx)
{
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

azk_status azk_context_capabilities(azk_context *ctx, azk_capabilities *caps)
{
    CHECK_ARG(ctx && caps);
    memset(caps, 0, sizeof *caps);
    snprintf(caps->name, sizeof caps->name, "%s", ctx->props.name);
    snprintf(caps->arch, sizeof caps->arch, "%s", ctx->props.gcnArchName);
    caps->compute_units = ctx->props.multiProcessorCount;
    caps->wavefront_size = ctx->props.warpSize;
    caps->lds_bytes_per_workgroup = ctx->props.sharedMemPerBlock;
    caps->hbm_bytes = ctx->props.totalGlobalMem;
    caps->clock_khz = ctx->props.clockRate;
    return AZK_OK;
}

azk_status azk_context_synchronize(azk_context *ctx)
{
    CHECK_ARG(ctx);
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return AZK_OK;
}

azk_status azk_buffer_alloc(azk_context *ctx, size_t bytes, azk_buffer **out)
{
    CHECK_ARG(ctx && out);
    HIP_TRY(hipSetDevice(ctx->device));
    void *p = nullptr;
    if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess) return fail(AZK_ERR_ALLOCATION, "hipMalloc(%zu) failed", bytes);
    *out = new azk_buffer{ctx, p, bytes, true};
    return AZK_OK;
}

azk_status azk_buffer_wrap(azk_context *ctx, void *device_ptr, size_t bytes, azk_buffer **out)
{
    CHECK_ARG(ctx && out && device_ptr);
    *out = new azk_buffer{ctx, device_ptr, bytes, false};
    return AZK_OK;
}

void azk_buffer_free(azk_buffer *b)
{
    if (!b) return;
    if (b->owned && b->ptr) {
        (void)hipSetDevice(b->ctx->device);
        (void)hipFree(b->ptr);
    }
    delete b;
}

size_t azk_buffer_size(const azk_buffer *b) { return b ? b->bytes : 0; }
void *azk_buffer_device_ptr(const azk_buffer *b) { return b ? b->ptr : nullptr; }

azk_status azk_buffer_upload(azk_buffer *b, size_t off, const void *host, size_t bytes)
{
    CHECK_ARG(b && (host || !bytes));
    azk_status s = check_range(b, off, bytes);
    if (s) return s;
    HIP_TRY(hipSetDevice(b->ctx->device));
    std::lock_guard<std::mutex> g(b->ctx->mu);
    HIP_TRY(hipMemcpyAsync((uint8_t *)b->ptr + off, host, bytes, hipMemcpyHostToDevice, b->ctx->stream));
    HIP_TRY(hipStreamSynchronize(b->ctx->stream));
    return AZK_OK;
}

azk_status azk_buffer_download(const azk_buffer *b, size_t off, void *host, size_t bytes)
{
    CHECK_ARG(b && (host || !bytes));
    azk_status s = check_range(b, off, bytes);
    if (s) return s;
    HIP_TRY(hipSetDevice(b->ctx->device));
    std::lock_guard<std::mutex> g(b->ctx->mu);
    HIP_TRY(hipMemcpyAsync(host, (const uint8_t *)b->ptr + off, bytes, hipMemcpyDeviceToHost, b->ctx->stream));
    HIP_TRY(hipStreamSynchronize(b->ctx->stream));
    return AZK_OK;
}

} // extern "C"

// ====================================================================== common plan helpers

namespace {

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

} // namespace

// ====================================================================== Keccak-f batch

struct azk_keccakf_plan {
    azk_context *ctx;
    azk_keccakf_desc d;
    DeviceMem in, out;
    std::mutex mu;
};

extern "C" azk_status azk_keccakf_plan_create(azk_context *ctx, const azk_keccakf_desc *desc, azk_keccakf_plan **out)
{
    CHECK_ARG(ctx && desc && out);
    if (desc->input_stride < 200 || desc->output_stride < 200 || desc->input_stride % 8 || desc->output_stride % 8)
        return fail(AZK_ERR_INVALID_LAYOUT, "Keccak-f strides must be multiples of 8 and >= 200");
    *out = new azk_keccakf_plan{ctx, *desc, {}, {}, {}};
    return AZK_OK;
}

extern "C" void azk_keccakf_plan_destroy(azk_keccakf_plan *p) { delete p; }

extern "C" azk_status azk_keccakf_permute_resident(azk_keccakf_plan *p, const azk_buffer *in, size_t in_off,
                                                   azk_buffer *out, size_t out_off, azk_stats *stats)
{
    CHECK_ARG(p && in && out);
    std::lock_guard<std::mutex> g(p->mu);
    const size_t ib = (size_t)p->d.count * p->d.input_stride, ob = (size_t)p->d.count * p->d.output_stride;
    azk_status s = check_range(in, in_off, ib);
    if (!s) s = check_range(out, out_off, ob);
    if (s) return s;
    if (((uintptr_t)in->ptr + in_off) % 8 || ((uintptr_t)out->ptr + out_off) % 8)
        return fail(AZK_ERR_INVALID_LAYOUT, "state buffers must be 8-byte aligned");
    if (overlaps(in, in_off, ib, out, out_off, ob) && (in->ptr != out->ptr || in_off != out_off || p->d.input_stride != p->d.output_stride))
        return fail(AZK_ERR_INVALID_LAYOUT, "input and output ranges alias");
    HIP_TRY(hipSetDevice(p->ctx->device));
    std::lock_guard<std::mutex> gc(p->ctx->mu);
    return run_timed(p->ctx, stats, [&]() {
        launch_keccakf_batch(p->ctx->stream, (uint8_t *)in->ptr + in_off, p->d.input_stride, (uint8_t *)out->ptr + out_off,
                             p->d.output_stride, p->d.count);
        return AZK_OK;
    });
}

extern "C" azk_status azk_keccakf_permute(azk_keccakf_plan *p, const void *states, size_t in_bytes, void *out,
                                          size_t out_bytes, azk_stats *stats)
{
    CHECK_ARG(p && (states || !p->d.count) && (out || !p->d.count));
    std::lock_guard<std::mutex> g(p->mu);
    const size_t ib = (size_t)p->d.count * p->d.input_stride, ob = (size_t)p->d.count * p->d.output_stride;
    if (in_bytes < ib || out_bytes < ob) return fail(AZK_ERR_INVALID_LAYOUT, "host arrays shorter than count x stride");
    HIP_TRY(hipSetDevice(p->ctx->device));
    azk_status s = p->in.ensure(ib);
    if (!s) s = p->out.ensure(ob);
    if (s) return s;
    std::lock_guard<std::mutex> gc(p->ctx->mu);
    return run_timed(p->ctx, stats, [&]() -> azk_status {
        HIP_TRY(hipMemcpyAsync(p->in.p, states, ib, hipMemcpyHostToDevice, p->ctx->stream));
        launch_keccakf_batch(p->ctx->stream, p->in.p, p->d.input_stride, p->out.p, p->d.output_stride, p->d.count);
        HIP_TRY(hipMemcpyAsync(out, p->out.p, ob, hipMemcpyDeviceToHost, p->ctx->stream));
        return AZK_OK;
    });
}

extern "C" azk_status azk_keccakf_permute_verified(azk_keccakf_plan *p, const void *states, size_t in_bytes, void *out,
                                                   size_t out_bytes, azk_stats *stats)
{
    azk_status s = azk_keccakf_permute(p, states, in_bytes, out, out_bytes, stats);
    if (s) return s;
    for (uint32_t i = 0; i < p->d.count; ++i) {
        uint64_t a[25];
        memcpy(a, (const uint8_t *)states + (size_t)i * p->d.input_stride, 200);
        host::keccak_f(a);
        if (memcmp(a, (const uint8_t *)out + (size_t)i * p->d.output_stride, 200))
            return fail(AZK_ERR_CORRECTNESS, "Keccak-f GPU result %u did not match the CPU oracle", i);
    }
    return AZK_OK;
}

extern "C" azk_status azk_keccakf_plan_clear(azk_keccakf_plan *p)
{
    CHECK_ARG(p);
    std::lock_guard<std::mutex> g(p->mu);
    HIP_TRY(hipSetDevice(p->ctx->device));
    if (p->in.p) HIP_TRY(hipMemset(p->in.p, 0, p->in.n));
    if (p->out.p) HIP_TRY(hipMemset(p->out.p, 0, p->out.n));
    return AZK_OK;
}

// ====================================================================== one-block hash

struct azk_hash_plan {
    azk_context *ctx;
    azk_oneblock_desc d;
    uint32_t suffix;
    DeviceMem in, out;
    std::mutex mu;
};

static azk_status hash_plan_create(azk_context *ctx, const azk_oneblock_desc *desc, uint32_t suffix, azk_hash_plan **out)
{
    CHECK_ARG(ctx && desc && out);
    if (desc->input_length > 136)
        return fail(AZK_ERR_UNSUPPORTED_ONE_BLOCK_LENGTH, "Fixed-rate SHA3/Keccak path supports at most 136 bytes, got %u",
                    desc->input_length);
    if (desc->input_stride < desc->input_length || desc->output_stride < 32)
        return fail(AZK_ERR_INVALID_LAYOUT, "stride shorter than the message or digest");
    *out = new azk_hash_plan{ctx, *desc, suffix, {}, {}, {}};
    return AZK_OK;
}

extern "C" azk_status azk_sha3_256_plan_create(azk_context *ctx, const azk_oneblock_desc *d, azk_hash_plan **o)
{
    return hash_plan_create(ctx, d, 0x06, o);
}
extern "C" azk_status azk_keccak_256_plan_create(azk_context *ctx, const azk_oneblock_desc *d, azk_hash_plan **o)
{
    return hash_plan_create(ctx, d, 0x01, o);
}
extern "C" void azk_hash_plan_destroy(azk_hash_plan *p) { delete p; }

extern "C" azk_status azk_hash_execute_resident(azk_hash_plan *p, const azk_buffer *in, size_t in_off, azk_buffer *out,
                                                size_t out_off, azk_stats *stats)
{
    CHECK_ARG(p && in && out);
    std::lock_guard<std::mutex> g(p->mu);
    const size_t ib = (size_t)p->d.count * p->d.input_stride, ob = (size_t)p->d.count * p->d.output_stride;
    azk_status s = check_range(in, in_off, ib);
    if (!s) s = check_range(out, out_off, ob);
    if (s) return s;
    if (overlaps(in, in_off, ib, out, out_off, ob)) return fail(AZK_ERR_INVALID_LAYOUT, "input and output ranges alias");
    HIP_TRY(hipSetDevice(p->ctx->device));
    std::lock_guard<std::mutex> gc(p->ctx->mu);
    return run_timed(p->ctx, stats, [&]() {
        launch_oneblock_hash(p->ctx->stream, (uint8_t *)in->ptr + in_off, p->d.input_stride, p->d.input_length,
                             (uint8_t *)out->ptr + out_off, p->d.output_stride, p->d.count, p->suffix);
        return AZK_OK;
    });
}

extern "C" azk_status azk_hash_execute(azk_hash_plan *p, const void *msgs, size_t in_bytes, void *digests,
                                       size_t out_bytes, azk_stats *stats)
{
    CHECK_ARG(p && (msgs || !p->d.count) && (digests || !p->d.count));
    std::lock_guard<std::mutex> g(p->mu);
    const size_t ib = (size_t)p->d.count * p->d.input_stride, ob = (size_t)p->d.count * p->d.output_stride;
    if (in_bytes < ib || out_bytes < ob) return fail(AZK_ERR_INVALID_LAYOUT, "host arrays shorter than count x stride");
    HIP_TRY(hipSetDevice(p->ctx->device));
    azk_status s = p->in.ensure(ib);
    if (!s) s = p->out.ensure(ob);
    if (s) return s;
    std::lock_guard<std::mutex> gc(p->ctx->mu);
    return run_timed(p->ctx, stats, [&]() -> azk_status {
        if (ib) HIP_TRY(hipMemcpyAsync(p->in.p, msgs, ib, hipMemcpyHostToDevice, p->ctx->stream));
        launch_oneblock_hash(p->ctx->stream, p->in.p, p->d.input_stride, p->d.input_length, p->out.p, p->d.output_stride,
                             p->d.count, p->suffix);
        if (ob) HIP_TRY(hipMemcpyAsync(digests, p->out.p, ob, hipMemcpyDeviceToHost, p->ctx->stream));
        return AZK_OK;
    });
}

extern "C" azk_status azk_hash_execute_verified(azk_hash_plan *p, const void *msgs, size_t in_bytes, void *digests,
                                                size_t out_bytes, azk_stats *stats)
{
    azk_status s = azk_hash_execute(p, msgs, in_bytes, digests, out_bytes, stats);
    if (s) return s;
    for (uint32_t i = 0; i < p->d.count; ++i) {
        uint8_t d[32];
        host::sha3_256((const uint8_t *)msgs + (size_t)i * p->d.input_stride, p->d.input_length, d, (uint8_t)p->suffix);
        if (memcmp(d, (const uint8_t *)digests + (size_t)i * p->d.output_stride, 32))
            return fail(AZK_ERR_CORRECTNESS, "one-block hash GPU result %u did not match the CPU oracle", i);
    }
    return AZK_OK;
}

extern "C" azk_status azk_hash_plan_clear(azk_hash_plan *p)
{
    CHECK_ARG(p);
    std::lock_guard<std::mutex> g(p->mu);
    HIP_TRY(hipSetDevice(p->ctx->device));
    if (p->in.p) HIP_TRY(hipMemset(p->in.p, 0, p->in.n));
    if (p->out.p) HIP_TRY(hipMemset(p->out.p, 0, p->out.n));
    return AZK_OK;
}

// ====================================================================== Merkle (raw leaves)

struct azk_merkle_plan {
    azk_context *ctx;
    azk_merkle_desc d;
    uint32_t log_m, log_t;
    DeviceMem tree, staging, path;
    bool committed = false;
    std::mutex mu;
};

extern "C" azk_status azk_merkle_plan_create(azk_context *ctx, const azk_merkle_desc *desc, azk_merkle_plan **out)
{
    CHECK_ARG(ctx && desc && out);
    if (!is_pow2(desc->leaf_count))
        return fail(AZK_ERR_INVALID_LEAF_COUNT, "Leaf count must be a non-zero power of two, got %u", desc->leaf_count);
    if (desc->leaf_length > 136)
        return fail(AZK_ERR_UNSUPPORTED_ONE_BLOCK_LENGTH, "leaves are hashed as one block (<= 136 bytes), got %u",
                    desc->leaf_length);
    if (desc->leaf_stride < desc->leaf_length) return fail(AZK_ERR_INVALID_LAYOUT, "leaf stride shorter than leaf length");
    auto *p = new azk_merkle_plan{ctx, *desc, ilog2(desc->leaf_count), 0, {}, {}, {}, false, {}};
    uint32_t lt = desc->subtree_log ? desc->subtree_log : kDefaultLogTile;
    if (lt > kDefaultLogTile) lt = kDefaultLogTile;  // 24 x 2^11 bytes of LDS per block
    p->log_t = std::min(lt, p->log_m);
    HIP_TRY(hipSetDevice(ctx->device));
    azk_status s = p->tree.ensure(((size_t)2 * desc->leaf_count - 1) * 32);
    if (!s) s = p->path.ensure(32 * 32 + 32);
    if (s) { delete p; return s; }
    *out = p;
    return AZK_OK;
}

extern "C" void azk_merkle_plan_destroy(azk_merkle_plan *p) { delete p; }

static azk_status merkle_enqueue(azk_merkle_plan *p, const void *leaves)
{
    launch_merkle_raw(p->ctx->stream, leaves, p->d.leaf_stride, p->d.leaf_length, p->d.leaf_count, p->log_t,
                      (uint32_t *)p->tree.p);
    launch_tree_upper(p->ctx->stream, (uint32_t *)p->tree.p, p->d.leaf_count, p->log_t, kUpperLogChunk);
    return AZK_OK;
}

static azk_status merkle_root_readback(azk_merkle_plan *p, uint8_t root[32])
{
    HIP_TRY(hipMemcpy(root, (uint8_t *)p->tree.p + ((size_t)2 * p->d.leaf_count - 2) * 32, 32, hipMemcpyDeviceToHost));
    p->committed = true;
    return AZK_OK;
}

extern "C" azk_status azk_merkle_commit_resident(azk_merkle_plan *p, const azk_buffer *leaves, size_t off,
                                                 uint8_t root[32], azk_stats *stats)
{
    CHECK_ARG(p && leaves && root);
    std::lock_guard<std::mutex> g(p->mu);
    azk_status s = check_range(leaves, off, (size_t)p->d.leaf_count * p->d.leaf_stride);
    if (s) return s;
    HIP_TRY(hipSetDevice(p->ctx->device));
    std::lock_guard<std::mutex> gc(p->ctx->mu);
    s = run_timed(p->ctx, stats, [&]() { return merkle_enqueue(p, (uint8_t *)leaves->ptr + off); });
    return s ? s : merkle_root_readback(p, root);
}

extern "C" azk_status azk_merkle_commit(azk_merkle_plan *p, const void *leaves, size_t bytes, uint8_t root[32],
                                        azk_stats *stats)
{
    CHECK_ARG(p && root);
    std::lock_guard<std::mutex> g(p->mu);
    const size_t need = (size_t)p->d.leaf_count * p->d.leaf_stride;
    if (bytes < need) return fail(AZK_ERR_INVALID_LAYOUT, "leaf bytes %zu < count x stride %zu", bytes, need);
    CHECK_ARG(leaves || !need);
    HIP_TRY(hipSetDevice(p->ctx->device));
    azk_status s = p->staging.ensure(need + 8);
    if (s) return s;
    std::lock_guard<std::mutex> gc(p->ctx->mu);
    s = run_timed(p->ctx, stats, [&]() -> azk_status {
        if (need) HIP_TRY(hipMemcpyAsync(p->staging.p, leaves, need, hipMemcpyHostToDevice, p->ctx->stream));
        return merkle_enqueue(p, p->staging.p);
    });
    return s ? s : merkle_root_readback(p, root);
}

extern "C" azk_status azk_merkle_commit_verified(azk_merkle_plan *p, const void *leaves, size_t bytes, uint8_t root[32],
                                                 azk_stats *stats)
{
    azk_status s = azk_merkle_commit(p, leaves, bytes, root, stats);
    if (s) return s;
    std::vector<uint8_t> level((size_t)p->d.leaf_count * 32);
    for (uint32_t i = 0; i < p->d.leaf_count; ++i)
        host::sha3_256((const uint8_t *)leaves + (size_t)i * p->d.leaf_stride, p->d.leaf_length, &level[32 * (size_t)i]);
    for (size_t n = p->d.leaf_count; n > 1; n /= 2)
        for (size_t k = 0; k < n / 2; ++k) host::sha3_256(&level[64 * k], 64, &level[32 * k]);
    if (memcmp(level.data(), root, 32)) return fail(AZK_ERR_CORRECTNESS, "Merkle GPU root did not match the CPU oracle");
    return AZK_OK;
}

extern "C" azk_status azk_merkle_open(azk_merkle_plan *p, uint32_t index, uint8_t *siblings, size_t sib_bytes,
                                      uint8_t root[32])
{
    CHECK_ARG(p);
    std::lock_guard<std::mutex> g(p->mu);
    if (!p->committed) return fail(AZK_ERR_INVALID_LAYOUT, "no resident tree: commit before opening");
    if (index >= p->d.leaf_count) return fail(AZK_ERR_INVALID_LAYOUT, "leaf index %u >= %u", index, p->d.leaf_count);
    const size_t need = (size_t)32 * p->log_m;
    if (sib_bytes < need || (need && !siblings)) return fail(AZK_ERR_INVALID_LAYOUT, "sibling buffer too small");
    HIP_TRY(hipSetDevice(p->ctx->device));
    std::lock_guard<std::mutex> gc(p->ctx->mu);
    launch_merkle_path(p->ctx->stream, (const uint32_t *)p->tree.p, p->d.leaf_count, p->log_m, index, (uint32_t *)p->path.p);
    HIP_TRY(hipGetLastError());
    std::vector<uint8_t> h(need + 32);
    HIP_TRY(hipMemcpyAsync(h.data(), p->path.p, need + 32, hipMemcpyDeviceToHost, p->ctx->stream));
    HIP_TRY(hipStreamSynchronize(p->ctx->stream));
    if (need) memcpy(siblings, h.data(), need);
    if (root) memcpy(root, h.data() + need, 32);
    return AZK_OK;
}

extern "C" azk_status azk_merkle_plan_clear(azk_merkle_plan *p)
{
    CHECK_ARG(p);
    std::lock_guard<std::mutex> g(p->mu);
    HIP_TRY(hipSetDevice(p->ctx->device));
    HIP_TRY(hipMemset(p->tree.p, 0, p->tree.n));
    if (p->staging.p) HIP_TRY(hipMemset(p->staging.p, 0, p->staging.n));
    p->committed = false;
    return AZK_OK;
}

// ====================================================================== prehashed Merkle

struct azk_prehashed_plan {
    azk_context *ctx;
    uint32_t count, log_m;
    DeviceMem tree;
    std::mutex mu;
};

extern "C" azk_status azk_prehashed_plan_create(azk_context *ctx, uint32_t leaf_count, azk_prehashed_plan **out)
{
    CHECK_ARG(ctx && out);
    if (!is_pow2(leaf_count))
        return fail(AZK_ERR_INVALID_LEAF_COUNT, "Leaf count must be a non-zero power of two, got %u", leaf_count);
    auto *p = new azk_prehashed_plan{ctx, leaf_count, ilog2(leaf_count), {}, {}};
    HIP_TRY(hipSetDevice(ctx->device));
    azk_status s = p->tree.ensure(((size_t)2 * leaf_count - 1) * 32);
    if (s) { delete p; return s; }
    *out = p;
    return AZK_OK;
}
extern "C" void azk_prehashed_plan_destroy(azk_prehashed_plan *p) { delete p; }

extern "C" azk_status azk_prehashed_commit_resident(azk_prehashed_plan *p, const azk_buffer *hashes, size_t off,
                                                    uint8_t root[32], azk_stats *stats)
{
    CHECK_ARG(p && hashes && root);
    std::lock_guard<std::mutex> g(p->mu);
    azk_status s = check_range(hashes, off, (size_t)p->count * 32);
    if (s) return s;
    if (((uintptr_t)hashes->ptr + off) % 4) return fail(AZK_ERR_INVALID_LAYOUT, "leaf hashes must be 4-byte aligned");
    HIP_TRY(hipSetDevice(p->ctx->device));
    std::lock_guard<std::mutex> gc(p->ctx->mu);
    s = run_timed(p->ctx, stats, [&]() {
        launch_copy_words(p->ctx->stream, (const uint32_t *)((uint8_t *)hashes->ptr + off), (uint32_t *)p->tree.p,
                          (uint64_t)p->count * 8);
        launch_tree_upper(p->ctx->stream, (uint32_t *)p->tree.p, p->count, 0, kUpperLogChunk);
        return AZK_OK;
    });
    if (s) return s;
    HIP_TRY(hipMemcpy(root, (uint8_t *)p->tree.p + ((size_t)2 * p->count - 2) * 32, 32, hipMemcpyDeviceToHost));
    return AZK_OK;
}

// ====================================================================== Circle FFT helpers

namespace {

struct FftPass { uint32_t lo, k, m; };

// Stages log_n-1 .. lt (lt = min(log_t, log_n)) split into passes of at most
// log_t - 3 stages, each on tiles of 2^log_t elements with >= 8 contiguous columns.
std::vector<FftPass> fft_passes(uint32_t log_n, uint32_t log_t)
{
    std::vector<FftPass> v;
    uint32_t lt = std::min(log_t, log_n);
    uint32_t S = log_n - lt;
    if (!S) return v;
    uint32_t kmax = log_t - 3;
    uint32_t P = (S + kmax - 1) / kmax, hi = log_n;
    for (uint32_t p = 0; p < P; ++p) {
        uint32_t k = S / P + (p < S % P ? 1 : 0);
        uint32_t lo = hi - k;
        v.push_back({lo, k, log_t - k});
        hi = lo;
    }
    return v;
}

// Enqueues the Circle FFT of `src` (interleaved line-basis coefficients) into `dst`;
// with `tree`, the final pass also hashes the codeword and builds its tile subtrees.
void enqueue_fft(hipStream_t st, const void *src, void *dst, const uint32_t *tw, uint32_t log_n, uint32_t log_t,
                 uint32_t *tree)
{
    const void *cur = src;
    for (const FftPass &p : fft_passes(log_n, log_t)) {
        launch_fft_pass(st, cur, dst, tw, log_n, p.lo, p.k, p.m);
        cur = dst;
    }
    launch_fft_final(st, cur, dst, tw, log_n, std::min(log_t, log_n), true, tree);
}

azk_status materialize_twiddles(azk_context *ctx, uint32_t log_n, uint32_t rounds, DeviceMem &fft_tw, DeviceMem *fold_tw)
{
    const uint64_t n = 1ull << log_n;
    azk_status s = fft_tw.ensure(4 * n);
    if (s) return s;
    launch_fft_twiddles(ctx->stream, (uint32_t *)fft_tw.p, log_n);
    if (fold_tw) {
        const uint64_t total = n - (n >> rounds);
        s = fold_tw->ensure(4 * std::max<uint64_t>(total, 1));
        if (s) return s;
        launch_inverse(ctx->stream, (const uint32_t *)fft_tw.p, (uint32_t *)fold_tw->p, total);
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return AZK_OK;
}

} // namespace

// ====================================================================== codeword plan

struct azk_codeword_plan {
    azk_context *ctx;
    uint32_t log_n;
    DeviceMem tw, in, out;
    std::mutex mu;
};

extern "C" azk_status azk_codeword_plan_create(azk_context *ctx, uint32_t log_n, azk_codeword_plan **out)
{
    CHECK_ARG(ctx && out);
    if (log_n < 1 || log_n > 28) return fail(AZK_ERR_INVALID_LAYOUT, "domain log size %u outside 1..28", log_n);
    auto *p = new azk_codeword_plan{ctx, log_n, {}, {}, {}, {}};
    HIP_TRY(hipSetDevice(ctx->device));
    std::lock_guard<std::mutex> gc(ctx->mu);
    azk_status s = materialize_twiddles(ctx, log_n, 1, p->tw, nullptr);
    if (s) { delete p; return s; }
    *out = p;
    return AZK_OK;
}
extern "C" void azk_codeword_plan_destroy(azk_codeword_plan *p) { delete p; }

extern "C" azk_status azk_codeword_execute_resident(azk_codeword_plan *p, const azk_buffer *coeffs, size_t coff,
                                                    azk_buffer *out, size_t ooff, azk_stats *stats)
{
    CHECK_ARG(p && coeffs && out);
    std::lock_guard<std::mutex> g(p->mu);
    const size_t bytes = (size_t)16 << p->log_n;
    azk_status s = check_range(coeffs, coff, bytes);
    if (!s) s = check_range(out, ooff, bytes);
    if (s) return s;
    if (overlaps(coeffs, coff, bytes, out, ooff, bytes)) return fail(AZK_ERR_INVALID_LAYOUT, "coefficient and output ranges alias");
    if (((uintptr_t)coeffs->ptr + coff) % 16 || ((uintptr_t)out->ptr + ooff) % 16)
        return fail(AZK_ERR_INVALID_LAYOUT, "QM31 buffers must be 16-byte aligned");
    HIP_TRY(hipSetDevice(p->ctx->device));
    std::lock_guard<std::mutex> gc(p->ctx->mu);
    return run_timed(p->ctx, stats, [&]() {
        enqueue_fft(p->ctx->stream, (uint8_t *)coeffs->ptr + coff, (uint8_t *)out->ptr + ooff, (const uint32_t *)p->tw.p,
                    p->log_n, kDefaultLogTile, nullptr);
        return AZK_OK;
    });
}

extern "C" azk_status azk_codeword_execute(azk_codeword_plan *p, const void *coeffs, size_t in_bytes, void *evals,
                                           size_t out_bytes, azk_stats *stats)
{
    CHECK_ARG(p && coeffs && evals);
    std::lock_guard<std::mutex> g(p->mu);
    const size_t bytes = (size_t)16 << p->log_n;
    if (in_bytes < bytes || out_bytes < bytes) return fail(AZK_ERR_INVALID_LAYOUT, "host arrays shorter than N x 16 bytes");
    HIP_TRY(hipSetDevice(p->ctx->device));
    azk_status s = p->in.ensure(bytes);
    if (!s) s = p->out.ensure(bytes);
    if (s) return s;
    std::lock_guard<std::mutex> gc(p->ctx->mu);
    return run_timed(p->ctx, stats, [&]() -> azk_status {
        HIP_TRY(hipMemcpyAsync(p->in.p, coeffs, bytes, hipMemcpyHostToDevice, p->ctx->stream));
        enqueue_fft(p->ctx->stream, p->in.p, p->out.p, (const uint32_t *)p->tw.p, p->log_n, kDefaultLogTile, nullptr);
        HIP_TRY(hipMemcpyAsync(evals, p->out.p, bytes, hipMemcpyDeviceToHost, p->ctx->stream));
        return AZK_OK;
    });
}

extern "C" azk_status azk_codeword_read_twiddles(azk_codeword_plan *p, uint32_t *out, size_t words)
{
    CHECK_ARG(p && out);
    const size_t n = ((size_t)1 << p->log_n) - 1;
    if (words < n) return fail(AZK_ERR_INVALID_LAYOUT, "need %zu words", n);
    HIP_TRY(hipSetDevice(p->ctx->device));
    HIP_TRY(hipMemcpy(out, p->tw.p, 4 * n, hipMemcpyDeviceToHost));
    return AZK_OK;
}

// ====================================================================== fold chain plan

struct azk_fold_plan {
    azk_context *ctx;
    uint32_t log_n, rounds;
    DeviceMem fft_tw, fold_tw, a, b, ch;
    std::mutex mu;
};

extern "C" azk_status azk_fold_plan_create(azk_context *ctx, uint32_t log_n, uint32_t rounds, azk_fold_plan **out)
{
    CHECK_ARG(ctx && out);
    if (log_n < 1 || log_n > 28 || rounds < 1 || rounds > log_n || rounds > (uint32_t)kMaxRounds)
        return fail(AZK_ERR_INVALID_LAYOUT, "fold chain log_n %u rounds %u", log_n, rounds);
    auto *p = new azk_fold_plan{ctx, log_n, rounds, {}, {}, {}, {}, {}, {}};
    HIP_TRY(hipSetDevice(ctx->device));
    std::lock_guard<std::mutex> gc(ctx->mu);
    azk_status s = materialize_twiddles(ctx, log_n, rounds, p->fft_tw, &p->fold_tw);
    if (!s) s = p->a.ensure((size_t)16 << (log_n - 1));
    if (!s) s = p->b.ensure((size_t)16 << (log_n - 1));
    if (!s) s = p->ch.ensure(16 * rounds);
    if (s) { delete p; return s; }
    *out = p;
    return AZK_OK;
}
extern "C" void azk_fold_plan_destroy(azk_fold_plan *p) { delete p; }

static azk_status fold_enqueue(azk_fold_plan *p, const void *in, void *out)
{
    const uint64_t n = 1ull << p->log_n;
    const void *cur = in;
    uint64_t off = 0, m = n;
    for (uint32_t r = 0; r < p->rounds; ++r) {
        void *dst = (r + 1 == p->rounds) ? out : (r % 2 ? p->b.p : p->a.p);
        uint64_t n_out = m / 2;
        uint32_t lt = std::min<uint32_t>(kDefaultLogTile, ilog2(n_out));
        launch_fold(p->ctx->stream, cur, dst, (const uint32_t *)p->fold_tw.p + off, (const uint32_t *)p->ch.p + 4 * r, n_out,
                    lt, nullptr);
        cur = dst;
        off += n_out;
        m = n_out;
    }
    return AZK_OK;
}

extern "C" azk_status azk_fold_execute_resident(azk_fold_plan *p, const azk_buffer *evals, size_t eoff,
                                                const uint32_t *challenges, azk_buffer *out, size_t ooff, azk_stats *stats)
{
    CHECK_ARG(p && evals && challenges && out);
    std::lock_guard<std::mutex> g(p->mu);
    const size_t ib = (size_t)16 << p->log_n, ob = (size_t)16 << (p->log_n - p->rounds);
    azk_status s = check_range(evals, eoff, ib);
    if (!s) s = check_range(out, ooff, ob);
    if (s) return s;
    if (overlaps(evals, eoff, ib, out, ooff, ob)) return fail(AZK_ERR_INVALID_LAYOUT, "input and output ranges alias");
    for (uint32_t i = 0; i < 4 * p->rounds; ++i)
        if (challenges[i] >= kP) return fail(AZK_ERR_INVALID_LAYOUT, "non-canonical challenge limb");
    HIP_TRY(hipSetDevice(p->ctx->device));
    std::lock_guard<std::mutex> gc(p->ctx->mu);
    return run_timed(p->ctx, stats, [&]() -> azk_status {
        HIP_TRY(hipMemcpyAsync(p->ch.p, challenges, 16 * p->rounds, hipMemcpyHostToDevice, p->ctx->stream));
        return fold_enqueue(p, (uint8_t *)evals->ptr + eoff, (uint8_t *)out->ptr + ooff);
    });
}

extern "C" azk_status azk_fold_execute(azk_fold_plan *p, const void *evals, size_t in_bytes, const uint32_t *challenges,
                                       void *out, size_t out_bytes, azk_stats *stats)
{
    CHECK_ARG(p && evals && challenges && out);
    const size_t ib = (size_t)16 << p->log_n, ob = (size_t)16 << (p->log_n - p->rounds);
    if (in_bytes < ib || out_bytes < ob) return fail(AZK_ERR_INVALID_LAYOUT, "host arrays too short");
    azk_buffer *bi = nullptr, *bo = nullptr;
    azk_status s = azk_buffer_alloc(p->ctx, ib, &bi);
    if (!s) s = azk_buffer_alloc(p->ctx, ob, &bo);
    if (!s) s = azk_buffer_upload(bi, 0, evals, ib);
    if (!s) s = azk_fold_execute_resident(p, bi, 0, challenges, bo, 0, stats);
    if (!s) s = azk_buffer_download(bo, 0, out, ob);
    azk_buffer_free(bi);
    azk_buffer_free(bo);
    return s;
}

// ====================================================================== Circle PCS prover

namespace {

void put32(std::vector<uint8_t> &v, uint32_t x) { for (int k = 0; k < 4; ++k) v.push_back((uint8_t)(x >> (8 * k))); }
void put64(std::vector<uint8_t> &v, uint64_t x) { for (int k = 0; k < 8; ++k) v.push_back((uint8_t)(x >> (8 * k))); }

// Circle transcript V1 frames, CircleProofFormat.swift:906-975.
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

// 44-byte domain descriptor, CircleDomainDescriptorCodecV1.encode (CircleProofFormat.swift:75-93).
std::vector<uint8_t> descriptor(uint32_t log_n)
{
    std::vector<uint8_t> d = {0x41, 0x5a, 0x4b, 0x43, 0x44, 0x56, 0x31, 0x00};
    for (uint32_t w : {1u, kP, 31u, 2u, 1268011823u, log_n, 1u << (31 - (log_n + 1)), log_n - 1, 1u}) put32(d, w);
    return d;
}

} // namespace

struct azk_circle_prover {
    azk_context *ctx;
    azk_circle_prover_desc d;
    uint32_t log_t;
    uint64_t n;
    // device state
    DeviceMem fft_tw, fold_tw, arena, frames, ops, proof, coeff_staging;
    std::vector<uint4 *> layers;      // rounds + 1 layers (layer 0 = codeword)
    std::vector<uint32_t *> trees;    // rounds trees
    uint32_t *challenges = nullptr, *qidx = nullptr;
    uint64_t *tstate = nullptr;
    std::vector<std::pair<uint32_t, uint32_t>> programs;  // (first op, count), rounds + 1 programs
    // proof layout (words)
    uint64_t roots_off, final_off, queries_off, query_words, proof_words;
    std::vector<uint32_t> layer_off;
    uint8_t *host_proof = nullptr;   // pinned staging
    // graph cache
    bool use_graph = true, timing = false;
    hipGraphExec_t graph = nullptr;
    const void *graph_src = nullptr;
    int graph_mode = -1;
    azk_phase_times phases{};
    std::mutex mu;
    ~azk_circle_prover()
    {
        if (graph) (void)hipGraphExecDestroy(graph);
        if (host_proof) (void)hipHostFree(host_proof);
    }
};

namespace {

azk_status prover_build(azk_circle_prover *p)
{
    const azk_circle_prover_desc &d = p->d;
    const uint64_t n = p->n;
    azk_context *ctx = p->ctx;
    azk_status s = materialize_twiddles(ctx, d.log_n, d.rounds, p->fft_tw, &p->fold_tw);
    if (s) return s;

    // ---- arena: layers, trees, transcript scratch
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
    size_t ch_at = at; at = align_up(at + 16 * (size_t)d.rounds, 256);
    size_t q_at = at; at = align_up(at + 4 * (size_t)d.query_count, 256);
    size_t st_at = at; at = align_up(at + 200, 256);
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

    // ---- transcript frames (8-byte aligned, zero padded) and programs
    std::vector<uint8_t> blob;
    auto add = [&](const std::vector<uint8_t> &b) {
        size_t o = blob.size();
        blob.insert(blob.end(), b.begin(), b.end());
        blob.resize(align_up(blob.size(), 8) + 8, 0);
        return std::make_pair(o, b.size());
    };
    std::vector<uint8_t> f;
    f = base_frame(0);
    for (uint32_t w : {1u, 1u, kP, 2u, 1268011823u}) put32(f, w);
    auto hdr = add(f);
    auto desc = add(descriptor(d.log_n));
    f = base_frame(1);
    for (uint32_t w : {d.log_blowup, d.query_count, d.folding_step, d.grinding_bits, d.log_blowup * d.query_count + d.grinding_bits})
        put32(f, w);
    auto sec = add(f);
    f = base_frame(2);
    put32(f, 32);
    auto pin = add(f);
    auto dig = add(std::vector<uint8_t>(d.public_input_digest, d.public_input_digest + 32));
    std::vector<std::pair<size_t, size_t>> cfr(d.rounds), chf(d.rounds);
    for (uint32_t r = 0; r < d.rounds; ++r) {
        f = base_frame(3);
        put32(f, r);
        put32(f, 32);
        cfr[r] = add(f);
        f = base_frame(4);
        put32(f, r);
        put32(f, kP);
        put32(f, 4);
        chf[r] = add(f);
    }
    const uint64_t nfin = n >> d.rounds;
    f = base_frame(5);
    put64(f, nfin);
    put64(f, nfin * 16);
    auto ffr = add(f);
    f = base_frame(6);
    put32(f, d.query_count);
    put64(f, n / 2);
    auto qfr = add(f);
    s = p->frames.ensure(blob.size());
    if (s) return s;
    HIP_TRY(hipMemcpy(p->frames.p, blob.data(), blob.size(), hipMemcpyHostToDevice));
    const uint8_t *fb = (const uint8_t *)p->frames.p;

    std::vector<TranscriptOp> ops;
    auto absorb = [&](const void *src, uint64_t len) { ops.push_back({TOP_ABSORB, 0x06, 0, 0, len, src, nullptr}); };
    auto absorb_f = [&](std::pair<size_t, size_t> fr) { absorb(fb + fr.first, fr.second); };
    for (uint32_t r = 0; r < d.rounds; ++r) {
        uint32_t first = (uint32_t)ops.size();
        if (r == 0) {  // prefixTranscriptAndChallenges, CircleProofFormat.swift:796-818
            ops.push_back({TOP_RESET, 0, 0, 0, 0, nullptr, nullptr});
            absorb_f(hdr);
            absorb_f(desc);
            absorb_f(sec);
            absorb_f(pin);
            absorb_f(dig);
        }
        const uint64_t m = n >> r;  // per round: commitment frame, root, challenge frame, squeeze 4 (:819-828)
        absorb_f(cfr[r]);
        absorb(p->trees[r] + 8 * (2 * m - 2), 32);
        absorb_f(chf[r]);
        ops.push_back({TOP_SQUEEZE_U32, 0, kP, 0, 4, nullptr, p->challenges + 4 * r});
        p->programs.push_back({first, (uint32_t)ops.size() - first});
    }
    {   // final layer + query squeeze, CircleProofFormat.swift:699-717,852-859
        uint32_t first = (uint32_t)ops.size();
        absorb_f(ffr);
        absorb(p->layers[d.rounds], nfin * 16);
        absorb_f(qfr);
        ops.push_back({TOP_SQUEEZE_U32, 0, (uint32_t)(n / 2), 0, d.query_count, nullptr, p->qidx});
        p->programs.push_back({first, (uint32_t)ops.size() - first});
    }
    s = p->ops.ensure(ops.size() * sizeof(TranscriptOp));
    if (s) return s;
    HIP_TRY(hipMemcpy(p->ops.p, ops.data(), ops.size() * sizeof(TranscriptOp), hipMemcpyHostToDevice));

    // ---- proof layout (all fields are 4-byte multiples; CircleProofFormat.swift:478-535)
    std::vector<uint8_t> head = {0x41, 0x5a, 0x4b, 0x50, 0x46, 0x56, 0x31, 0x00};
    put32(head, 1);
    put32(head, 1);
    put32(head, 44);
    auto dd = descriptor(d.log_n);
    head.insert(head.end(), dd.begin(), dd.end());
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
    if (s) return s;
    HIP_TRY(hipMemcpy(p->proof.p, tmpl.data(), 4 * p->proof_words, hipMemcpyHostToDevice));
    HIP_TRY(hipHostMalloc((void **)&p->host_proof, 4 * p->proof_words, hipHostMallocDefault));
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
    e.final_count = p->n >> p->d.rounds;
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

// Event slots for per-phase timing (eager mode only).
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
enum { PH_START = 0, PH_FFT, PH_COMMIT0, PH_FOLD, PH_TREE, PH_TRANSCRIPT, PH_EMIT };

// The commit superstep, stream-ordered: FFT -> first-layer leaf hashes + tree ->
// per round {transcript -> fold fused with next-layer hashing -> upper tree} ->
// final-layer absorb + query squeeze -> proof gather -> one D2H copy.
// mode 0: src = line-basis coefficients (FFT first); mode 1: src = codeword.
void prover_enqueue(azk_circle_prover *p, hipStream_t st, const void *src, int mode, PhaseClock &clk)
{
    const azk_circle_prover_desc &d = p->d;
    const uint32_t lt0 = std::min(p->log_t, d.log_n);
    const uint4 *layer0 = mode == 0 ? p->layers[0] : (const uint4 *)src;
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
    const TranscriptOp *ops = (const TranscriptOp *)p->ops.p;
    uint64_t off = 0, m = p->n;
    for (uint32_t r = 0; r < d.rounds; ++r) {
        launch_transcript(st, ops + p->programs[r].first, p->programs[r].second, p->tstate);
        clk.mark(PH_TRANSCRIPT);
        const uint64_t n_out = m / 2;
        const bool commit_next = r + 1 < d.rounds;
        const uint32_t lt = std::min<uint32_t>(p->log_t, ilog2(n_out));
        launch_fold(st, r == 0 ? (const void *)layer0 : (const void *)p->layers[r], p->layers[r + 1],
                    (const uint32_t *)p->fold_tw.p + off, p->challenges + 4 * r, n_out, lt,
                    commit_next ? p->trees[r + 1] : nullptr);
        clk.mark(PH_FOLD);
        if (commit_next) {
            launch_tree_upper(st, p->trees[r + 1], n_out, lt, kUpperLogChunk);
            clk.mark(PH_TREE);
        }
        off += n_out;
        m = n_out;
    }
    launch_transcript(st, ops + p->programs[d.rounds].first, p->programs[d.rounds].second, p->tstate);
    clk.mark(PH_TRANSCRIPT);
    launch_emit_proof(st, emit_params(p, layer0), d.query_count);
    (void)hipMemcpyAsync(p->host_proof, p->proof.p, 4 * p->proof_words, hipMemcpyDeviceToHost, st);
    clk.mark(PH_EMIT);
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
    p->log_t = kDefaultLogTile;
    HIP_TRY(hipSetDevice(ctx->device));
    std::lock_guard<std::mutex> gc(ctx->mu);
    azk_status s = prover_build(p);
    if (s) { delete p; return s; }
    *out = p;
    return AZK_OK;
}

extern "C" void azk_circle_prover_destroy(azk_circle_prover *p)
{
    if (!p) return;
    (void)hipSetDevice(p->ctx->device);
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

namespace {

azk_status prove_locked(azk_circle_prover *p, const void *src, int mode, uint8_t *proof, size_t cap, size_t *len,
                        azk_stats *stats)
{
    const size_t bytes = 4 * p->proof_words;
    if (cap < bytes || !proof) return fail(AZK_ERR_INVALID_LAYOUT, "proof capacity %zu < %zu", cap, bytes);
    hipStream_t st = p->ctx->stream;
    const bool graph = p->use_graph && !p->timing;
    PhaseClock clk(st, p->timing);
    azk_status s = run_timed(p->ctx, stats, [&]() -> azk_status {
        if (graph) {
            if (!p->graph || p->graph_src != src || p->graph_mode != mode) {
                if (p->graph) (void)hipGraphExecDestroy(p->graph);
                p->graph = nullptr;
                hipGraph_t gr;
                HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
                PhaseClock none(st, false);
                prover_enqueue(p, st, src, mode, none);
                HIP_TRY(hipStreamEndCapture(st, &gr));
                hipError_t e = hipGraphInstantiate(&p->graph, gr, nullptr, nullptr, 0);
                (void)hipGraphDestroy(gr);
                if (e != hipSuccess) return fail(AZK_ERR_COMMAND_FAILED, "hipGraphInstantiate: %s", hipGetErrorString(e));
                p->graph_src = src;
                p->graph_mode = mode;
            }
            HIP_TRY(hipGraphLaunch(p->graph, st));
        } else {
            prover_enqueue(p, st, src, mode, clk);
        }
        return AZK_OK;
    });
    if (s) return s;
    if (p->timing && clk.ev.size() > 1) {
        azk_phase_times t{};
        uint64_t *slot[] = {nullptr, &t.fft_ns, &t.commit0_ns, &t.fold_ns, &t.tree_ns, &t.transcript_ns, &t.emit_ns};
        for (size_t i = 1; i < clk.ev.size(); ++i) {
            float ms = 0;
            (void)hipEventElapsedTime(&ms, clk.ev[i - 1], clk.ev[i]);
            *slot[clk.tag[i]] += (uint64_t)(ms * 1e6);
        }
        float ms = 0;
        (void)hipEventElapsedTime(&ms, clk.ev.front(), clk.ev.back());
        t.total_ns = (uint64_t)(ms * 1e6);
        p->phases = t;
    }
    memcpy(proof, p->host_proof, bytes);
    if (len) *len = bytes;
    return AZK_OK;
}

} // namespace

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
    // zero everything that held witness-derived data; the proof template is rebuilt
    std::vector<uint32_t> tmpl(p->proof_words);
    HIP_TRY(hipMemcpy(tmpl.data(), p->proof.p, 4 * p->proof_words, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemset(p->arena.p, 0, p->arena.n));
    if (p->coeff_staging.p) HIP_TRY(hipMemset(p->coeff_staging.p, 0, p->coeff_staging.n));
    for (uint32_t r = 0; r < p->d.rounds; ++r)
        memset(tmpl.data() + p->roots_off + 8 * r, 0, 32);
    memset(tmpl.data() + p->final_off, 0, 16 * (p->n >> p->d.rounds));
    memset(tmpl.data() + p->queries_off, 0, 4 * p->d.query_count * p->query_words);
    HIP_TRY(hipMemcpy(p->proof.p, tmpl.data(), 4 * p->proof_words, hipMemcpyHostToDevice));
    memset(p->host_proof, 0, 4 * p->proof_words);
    return AZK_OK;
}

// ====================================================================== batched form

extern "C" azk_status azk_circle_prove_batch(azk_circle_prover *const *provers, const azk_buffer *const *coeffs,
                                             const size_t *offsets, uint32_t jobs, uint8_t *const *proofs,
                                             const size_t *caps, size_t *lens, azk_stats *stats)
{
    CHECK_ARG(provers && coeffs && offsets && proofs && caps);
    if (!jobs) return AZK_OK;
    azk_context *ctx = provers[0]->ctx;
    for (uint32_t j = 0; j < jobs; ++j) {
        CHECK_ARG(provers[j] && coeffs[j] && proofs[j]);
        if (provers[j]->ctx != ctx) return fail(AZK_ERR_INVALID_ARGUMENT, "all provers must share one context");
        for (uint32_t k = 0; k < j; ++k)
            if (provers[k] == provers[j]) return fail(AZK_ERR_INVALID_ARGUMENT, "a prover appears twice in one batch");
        azk_status s = check_range(coeffs[j], offsets[j], (size_t)16 << provers[j]->d.log_n);
        if (s) return s;
        if (caps[j] < 4 * provers[j]->proof_words) return fail(AZK_ERR_INVALID_LAYOUT, "proof capacity of job %u", j);
    }
    std::vector<std::unique_lock<std::mutex>> locks;
    for (uint32_t j = 0; j < jobs; ++j) locks.emplace_back(provers[j]->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    std::lock_guard<std::mutex> gc(ctx->mu);
    hipStream_t st = ctx->stream;
    // One graph: a fork from the capture stream into one branch per job, joined at the end.
    std::vector<hipStream_t> side(jobs);
    std::vector<hipEvent_t> done(jobs);
    hipEvent_t fork;
    HIP_TRY(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
    for (uint32_t j = 0; j < jobs; ++j) {
        HIP_TRY(hipStreamCreateWithFlags(&side[j], hipStreamNonBlocking));
        HIP_TRY(hipEventCreateWithFlags(&done[j], hipEventDisableTiming));
    }
    hipGraph_t gr = nullptr;
    hipGraphExec_t ge = nullptr;
    azk_status s = run_timed(ctx, stats, [&]() -> azk_status {
        HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        HIP_TRY(hipEventRecord(fork, st));
        for (uint32_t j = 0; j < jobs; ++j) {
            HIP_TRY(hipStreamWaitEvent(side[j], fork, 0));
            PhaseClock none(side[j], false);
            prover_enqueue(provers[j], side[j], (const uint8_t *)coeffs[j]->ptr + offsets[j], 0, none);
            HIP_TRY(hipEventRecord(done[j], side[j]));
            HIP_TRY(hipStreamWaitEvent(st, done[j], 0));
        }
        HIP_TRY(hipStreamEndCapture(st, &gr));
        HIP_TRY(hipGraphInstantiate(&ge, gr, nullptr, nullptr, 0));
        HIP_TRY(hipGraphLaunch(ge, st));
        return AZK_OK;
    });
    if (ge) (void)hipGraphExecDestroy(ge);
    if (gr) (void)hipGraphDestroy(gr);
    for (uint32_t j = 0; j < jobs; ++j) {
        (void)hipStreamDestroy(side[j]);
        (void)hipEventDestroy(done[j]);
    }
    (void)hipEventDestroy(fork);
    if (s) return s;
    for (uint32_t j = 0; j < jobs; ++j) {
        memcpy(proofs[j], provers[j]->host_proof, 4 * provers[j]->proof_words);
        if (lens) lens[j] = 4 * provers[j]->proof_words;
    }
    return AZK_OK;
}
