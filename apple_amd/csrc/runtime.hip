// Host runtime behind include/azk.h: context, resident buffers and the reusable plans
// (Keccak-f, one-block hashes, Merkle, prehashed Merkle, Circle codeword, fold chain).
// The Circle PCS prover and the batched form live in prover.hip.
//
// Plan semantics mirror the reference's Swift plans (cited per entry point in azk.h);
// buffers, streams and graphs are HIP-native.
#include "host_gates.hpp"
#include "runtime_common.hpp"
#include "host_sha3.hpp"

using namespace azk;
using namespace azk::rt;

namespace {
thread_local std::string g_err;
}

namespace azk::rt {

void ctx_retain(azk_context *c) { c->refs.fetch_add(1, std::memory_order_relaxed); }

void ctx_release(azk_context *c)
{
    if (c->refs.fetch_sub(1, std::memory_order_acq_rel) != 1) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    (void)hipStreamDestroy(c->stream);
    delete c;
}

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

azk_status check_range(const azk_buffer *b, size_t off, size_t bytes)
{
    if (!b) return fail(AZK_ERR_INVALID_ARGUMENT, "null buffer");
    if (off > b->bytes || bytes > b->bytes - off)
        return fail(AZK_ERR_INVALID_LAYOUT, "range [%zu, +%zu) exceeds buffer of %zu bytes", off, bytes, b->bytes);
    return AZK_OK;
}

bool overlaps(const azk_buffer *a, size_t ao, size_t an, const azk_buffer *b, size_t bo, size_t bn)
{
    if (!a || !b || !an || !bn) return false;
    uintptr_t a0 = (uintptr_t)a->ptr + ao, b0 = (uintptr_t)b->ptr + bo;
    return a0 < b0 + bn && b0 < a0 + an;
}
std::vector<FftPass> fft_passes(uint32_t log_n, uint32_t log_t_final, bool m31)
{
    std::vector<FftPass> v;
    uint32_t lt = std::min(log_t_final, log_n);
    uint32_t S = log_n - lt;
    if (!S) return v;
    // QM31: tiles of 2^12 (64 KB) to 2^13 elements with >= 16 columns (256-byte rows), k <= 10.
    // M31: tiles of 2^13 (32 KB) to 2^14 elements with >= 64 columns (256-byte rows) where the
    // cap allows, else >= 16; k <= 10.
    const uint32_t cap = m31 ? kMaxPassLogTileM31 : kMaxPassLogTile, min_t = m31 ? 13 : 12, want_m = m31 ? 6 : 4;
    const uint32_t kmax = 10;
    uint32_t P = (S + kmax - 1) / kmax, hi = log_n;
    for (uint32_t p = 0; p < P; ++p) {
        uint32_t k = S / P + (p < S % P ? 1 : 0);
        uint32_t lo = hi - k;
        const uint32_t log_t = std::min(cap, std::max(min_t, k + want_m));  // columns from the lo low bits
        v.push_back({lo, k, std::min(log_t - k, lo)});
        hi = lo;
    }
    return v;
}

azk_status check_exec_plan(const azk_circle_exec_plan &x)
{
    if (x.log_tile_commit < 6 || x.log_tile_commit > kDefaultLogTile || x.log_tile_fold < 6 ||
        x.log_tile_fold > kDefaultLogTile || x.log_tile_leafwise < kLeafwiseMinLogTile ||
        x.log_tile_leafwise > kLeafwiseLogTile || x.log_leafwise_below > 31 || x.log_tree_chunk < 3 ||
        x.log_tree_chunk > kUpperLogChunk || (x.log_leafwise_grid && (x.log_leafwise_grid < 4 || x.log_leafwise_grid > 20)) ||
        (x.tree_waves != 4 && x.tree_waves != 8 && x.tree_waves != 16))
        return fail(AZK_ERR_INVALID_KERNEL_CONFIG, "execution plan outside the supported tiles (commit/fold 6..11, "
                                                   "leafwise 3..8, chunk 3..10, leafwise grid 0 or 4..20, tree waves 4/8/16)");
    return AZK_OK;
}

ExecPlan to_exec_plan(const azk_circle_exec_plan &x)
{
    ExecPlan e;
    e.log_tile_commit = x.log_tile_commit;
    e.log_tile_fold = x.log_tile_fold;
    e.log_leafwise_below = x.log_leafwise_below;
    e.log_tile_leafwise = x.log_tile_leafwise;
    e.log_tree_chunk = x.log_tree_chunk;
    e.log_leafwise_grid = x.log_leafwise_grid;
    e.tree_waves = x.tree_waves;
    return e;
}

azk_circle_exec_plan from_exec_plan(const ExecPlan &e)
{
    azk_circle_exec_plan x{};
    x.log_tile_commit = e.log_tile_commit;
    x.log_tile_fold = e.log_tile_fold;
    x.log_leafwise_below = e.log_leafwise_below;
    x.log_tile_leafwise = e.log_tile_leafwise;
    x.log_tree_chunk = e.log_tree_chunk;
    x.log_leafwise_grid = e.log_leafwise_grid;
    x.tree_waves = e.tree_waves;
    return x;
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

std::vector<TreeStep> tree_schedule(uint64_t m, uint32_t level, uint32_t max_log_c)
{
    std::vector<TreeStep> v;
    const uint32_t log_m = ilog2(m);
    while (level < log_m) {
        uint32_t log_c = std::min(max_log_c, log_m - level);
        v.push_back({level, log_c, (m >> level) >> log_c});
        level += log_c;
    }
    return v;
}

void plan_fft(std::vector<Launch<FftPassDesc>> &passes, Launch<FftFinalDesc> &final_, const void *src, void *dst,
              const uint32_t *tw, uint32_t log_n, uint32_t log_t, bool run_fft, uint32_t *tree, bool m31, void *scratch)
{
    const uint32_t lt = std::min(log_t, log_n);
    const void *cur = src;
    void *mid = m31 && scratch ? scratch : dst;  // where the passes leave the codeword for the final pass
    if (run_fft) {
        auto fp = fft_passes(log_n, log_t, m31);
        if (passes.size() < fp.size()) passes.resize(fp.size());
        for (size_t i = 0; i < fp.size(); ++i) {
            FftPassDesc d{(const uint4 *)cur, (uint4 *)mid, tw, log_n, fp[i].lo, fp[i].k, fp[i].m, 0, 0};
            const uint32_t tlog = fp[i].k + fp[i].m;
            passes[i].add(d, (uint32_t)(1ull << (log_n - tlog)), tlog);
            cur = mid;
        }
    }
    FftFinalDesc f{(const uint4 *)cur, (uint4 *)dst, tw, tree, log_n, lt, run_fft ? 1u : 0u, 0};
    final_.add(f, (uint32_t)(1ull << (log_n - lt)), lt);
}

uint32_t plan_fold(Launch<FoldDesc> &big, Launch<FoldDesc> &leaf, const void *in, void *out, const uint32_t *tw,
                   const uint32_t *ch, uint64_t n_out, uint32_t log_t, uint32_t *tree, uint32_t log_leafwise_below,
                   uint32_t log_t_leaf, uint32_t log_grid)
{
    const uint32_t log_out = ilog2(n_out);
    if (tree && log_out < log_leafwise_below) {
        uint32_t lt = log_t_leaf;
        if (log_grid) {  // small tiles for small layers: fewer in-block levels in the one-lane form
            const uint32_t want = log_out > log_grid ? log_out - log_grid : 0;
            lt = std::min(log_t_leaf, std::max(kLeafwiseMinLogTile, want));
        }
        lt = std::min(lt, log_out);
        leaf.add(FoldDesc{(const uint4 *)in, (uint4 *)out, tw, ch, tree, n_out, lt, 0}, (uint32_t)(n_out >> lt), lt);
        return lt;
    }
    const uint32_t lt = std::min(log_t, log_out);
    big.add(FoldDesc{(const uint4 *)in, (uint4 *)out, tw, ch, tree, n_out, lt, 0}, (uint32_t)(n_out >> lt), lt);
    return lt;
}

bool plan_tree(std::vector<Launch<TreeDesc>> &steps, uint32_t *tree, uint64_t m, uint32_t level, uint32_t log_chunk,
               const TranscriptDesc *after)
{
    auto sch = tree_schedule(m, level, log_chunk);
    if (steps.size() < sch.size()) steps.resize(sch.size());
    for (size_t k = 0; k < sch.size(); ++k) {
        TreeDesc d{tree, m, nullptr, nullptr, sch[k].level, sch[k].log_c, 0, 0};
        if (after && k + 1 == sch.size()) {  // the last step is one chunk: one block computes the root
            d.t_ops = after->ops;
            d.t_state = after->state;
            d.t_nops = after->nops;
        }
        steps[k].add(d, (uint32_t)sch[k].chunks, sch[k].log_c);
    }
    return after && !sch.empty();
}

// A one-off program for the single-job plans: tables are built per call, uploaded,
// and launched on the context stream.
struct OneShot {
    TableArena arena;
    std::vector<Launch<FftPassDesc>> passes;
    Launch<FftFinalDesc> final_;
    std::vector<Launch<TreeDesc>> trees;
    std::vector<Launch<FoldDesc>> folds_big, folds_leaf;
    Launch<MerkleRawDesc> raw;
    Launch<CopyDesc> copy;
    const void *key_in = nullptr, *key_out = nullptr;  // the tables were built for these buffers
    bool valid = false;
    bool matches(const void *in, const void *out) const { return valid && key_in == in && key_out == out; }
    void keep(const void *in, const void *out)
    {
        key_in = in;
        key_out = out;
        valid = true;
    }
    void reset()  // keeps the device table buffer
    {
        valid = false;
        arena.host.clear();
        passes.clear();
        final_ = {};
        trees.clear();
        folds_big.clear();
        folds_leaf.clear();
        raw = {};
        copy = {};
    }
};

} // namespace azk::rt

// ====================================================================== context / buffers

extern "C" {

const char *azk_last_error(void) { return g_err.c_str(); }
const char *azk_version(void) { return "azk 0.1 gfx950"; }

azk_status azk_context_create(int device, azk_context **out)
{
    CHECK_ARG(out);
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return fail(AZK_ERR_NO_DEVICE, "no HIP device");
    if (device < 0 || device >= count) return fail(AZK_ERR_NO_DEVICE, "device %d of %d", device, count);
    HIP_TRY(hipSetDevice(device));
    auto *c = new azk_context();
    c->device = device;
    const hipError_t e = hipGetDeviceProperties(&c->props, device);
    if (e != hipSuccess) {
        delete c;
        return fail(AZK_ERR_COMMAND_FAILED, "device %d set-up: %s", device, hipGetErrorString(e));
    }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return fail(AZK_ERR_ALLOCATION, "hipStreamCreate failed");
    }
    *out = c;
    return AZK_OK;
}

void azk_context_destroy(azk_context *ctx)
{
    if (ctx) ctx_release(ctx);
}

azk_status azk_context_capabilities(azk_context *ctx, azk_capabilities *caps)
{
    CHECK_ARG(ctx && caps);
    memset(caps, 0, sizeof *caps);
    // props.name can be empty on gfx950 parts: fall back to the driver's name, then to the
    // architecture and CU count
    if (ctx->props.name[0]) snprintf(caps->name, sizeof caps->name, "%s", ctx->props.name);
    else if (hipDeviceGetName(caps->name, (int)sizeof caps->name, ctx->device) != hipSuccess || !caps->name[0])
        snprintf(caps->name, sizeof caps->name, "AMD %s (%d CUs)", ctx->props.gcnArchName, ctx->props.multiProcessorCount);
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
    ctx_retain(ctx);
    return AZK_OK;
}

azk_status azk_buffer_wrap(azk_context *ctx, void *device_ptr, size_t bytes, azk_buffer **out)
{
    CHECK_ARG(ctx && out && device_ptr);
    *out = new azk_buffer{ctx, device_ptr, bytes, false};
    ctx_retain(ctx);
    return AZK_OK;
}

void azk_buffer_free(azk_buffer *b)
{
    if (!b) return;
    if (b->owned && b->ptr) {
        (void)hipSetDevice(b->ctx->device);
        (void)hipFree(b->ptr);
    }
    azk_context *c = b->ctx;
    delete b;
    ctx_release(c);
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
    ctx_retain(ctx);
    return AZK_OK;
}

extern "C" void azk_keccakf_plan_destroy(azk_keccakf_plan *p) { destroy_on_ctx(p); }

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
    ctx_retain(ctx);
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
extern "C" void azk_hash_plan_destroy(azk_hash_plan *p) { destroy_on_ctx(p); }

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
    DeviceMem tree, staging, path, paths;
    bool committed = false;
    std::mutex mu;
};

// Default treelet: at least 256 blocks (one per CU) until tiles reach 2^11 leaves; 2^16 x
// 64-byte leaves measured 148 us with 2^8-leaf treelets against 243 us with 2^11.
uint32_t azk::rt::merkle_default_log_tile(uint32_t log_m)
{
    return std::min(std::min(kDefaultLogTile, std::max(8u, log_m > 8 ? log_m - 8 : 0u)), log_m);
}

// Descriptors by value: one dependent table read less per launch, and no upload.
void azk::rt::merkle_launch(hipStream_t st, const void *leaves, uint32_t *tree, uint64_t m, uint32_t stride,
                            uint32_t len, uint32_t log_t)
{
    const MerkleRawDesc raw{(const uint8_t *)leaves, tree, m, stride, len, log_t, 0};
    const bool aligned = ((stride & 3u) == 0) && (((uintptr_t)leaves & 3u) == 0);
    launch_merkle_raw(st, nullptr, 1, (uint32_t)(m >> log_t), aligned, log_t, &raw);
    const uint32_t log_m = ilog2(m);
    for (uint32_t level = log_t; level < log_m;) {  // tree_schedule, without a host vector
        const uint32_t log_c = std::min(kUpperLogChunk, log_m - level);
        const TreeDesc d{tree, m, nullptr, nullptr, level, log_c, 0, 0};
        launch_tree(st, nullptr, 1, (uint32_t)((m >> level) >> log_c), log_c, 4, &d);
        level += log_c;
    }
}

extern "C" azk_status azk_merkle_plan_create(azk_context *ctx, const azk_merkle_desc *desc, azk_merkle_plan **out)
{
    CHECK_ARG(ctx && desc && out);
    if (!is_pow2(desc->leaf_count))
        return fail(AZK_ERR_INVALID_LEAF_COUNT, "Leaf count must be a non-zero power of two, got %u", desc->leaf_count);
    if (desc->leaf_length > 136)
        return fail(AZK_ERR_UNSUPPORTED_ONE_BLOCK_LENGTH, "leaves are hashed as one block (<= 136 bytes), got %u",
                    desc->leaf_length);
    if (desc->leaf_stride < desc->leaf_length) return fail(AZK_ERR_INVALID_LAYOUT, "leaf stride shorter than leaf length");
    auto *p = new azk_merkle_plan{ctx, *desc, ilog2(desc->leaf_count), 0, {}, {}, {}, {}, false, {}};
    uint32_t lt = desc->subtree_log ? desc->subtree_log : merkle_default_log_tile(p->log_m);
    if (lt > kDefaultLogTile) lt = kDefaultLogTile;  // 16 x 2^11 bytes of LDS per block
    p->log_t = std::min(lt, p->log_m);
    HIP_TRY(hipSetDevice(ctx->device));
    azk_status s = p->tree.ensure(((size_t)2 * desc->leaf_count - 1) * 32);
    if (!s) s = p->path.ensure(32 * 32 + 32);
    if (s) { delete p; return s; }
    ctx_retain(p->ctx);
    *out = p;
    return AZK_OK;
}

extern "C" void azk_merkle_plan_destroy(azk_merkle_plan *p) { destroy_on_ctx(p); }

static azk_status merkle_enqueue(azk_merkle_plan *p, const void *leaves)
{
    merkle_launch(p->ctx->stream, leaves, (uint32_t *)p->tree.p, p->d.leaf_count, p->d.leaf_stride, p->d.leaf_length,
                  p->log_t);
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

extern "C" azk_status azk_merkle_open_batch(azk_merkle_plan *p, const uint32_t *indices, size_t count,
                                            uint8_t *siblings, size_t siblings_bytes, uint8_t root[32])
{
    CHECK_ARG(p && (indices || !count));
    std::lock_guard<std::mutex> g(p->mu);
    if (!p->committed) return fail(AZK_ERR_INVALID_LAYOUT, "no resident tree: commit before opening");
    if (count > (1u << 24)) return fail(AZK_ERR_INVALID_LAYOUT, "%zu openings in one call (at most 2^24)", count);
    for (size_t k = 0; k < count; ++k)
        if (indices[k] >= p->d.leaf_count)
            return fail(AZK_ERR_INVALID_LAYOUT, "leaf index %u >= %u (opening %zu)", indices[k], p->d.leaf_count, k);
    const size_t per = (size_t)32 * p->log_m, need = per * count;
    if (siblings_bytes < need || (need && !siblings)) return fail(AZK_ERR_INVALID_LAYOUT, "sibling buffer too small");
    HIP_TRY(hipSetDevice(p->ctx->device));
    std::lock_guard<std::mutex> gc(p->ctx->mu);
    // device words: the indices (4 bytes each, 256-byte aligned), then the paths
    const size_t idx_bytes = (4 * count + 255) & ~(size_t)255;
    azk_status s = p->paths.ensure(std::max<size_t>(idx_bytes + need, 256));
    if (s) return s;
    hipStream_t st = p->ctx->stream;
    uint32_t *dev_idx = (uint32_t *)p->paths.p, *dev_out = (uint32_t *)((uint8_t *)p->paths.p + idx_bytes);
    if (count) HIP_TRY(hipMemcpyAsync(dev_idx, indices, 4 * count, hipMemcpyHostToDevice, st));
    launch_merkle_paths(st, (const uint32_t *)p->tree.p, p->d.leaf_count, p->log_m, dev_idx, (uint32_t)count, dev_out);
    HIP_TRY(hipGetLastError());
    if (need) HIP_TRY(hipMemcpyAsync(siblings, dev_out, need, hipMemcpyDeviceToHost, st));
    if (root) HIP_TRY(hipMemcpyAsync(root, (const uint8_t *)p->tree.p + (size_t)32 * (2 * (uint64_t)p->d.leaf_count - 2), 32,
                                     hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
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
    OneShot shot;
    std::mutex mu;
};

extern "C" azk_status azk_prehashed_plan_create(azk_context *ctx, uint32_t leaf_count, azk_prehashed_plan **out)
{
    CHECK_ARG(ctx && out);
    if (!is_pow2(leaf_count))
        return fail(AZK_ERR_INVALID_LEAF_COUNT, "Leaf count must be a non-zero power of two, got %u", leaf_count);
    auto *p = new azk_prehashed_plan{ctx, leaf_count, ilog2(leaf_count), {}, {}, {}};
    HIP_TRY(hipSetDevice(ctx->device));
    azk_status s = p->tree.ensure(((size_t)2 * leaf_count - 1) * 32);
    if (s) { delete p; return s; }
    ctx_retain(p->ctx);
    *out = p;
    return AZK_OK;
}
extern "C" void azk_prehashed_plan_destroy(azk_prehashed_plan *p) { destroy_on_ctx(p); }

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
    OneShot &o = p->shot;
    o.reset();
    const uint64_t words = (uint64_t)p->count * 8;
    o.copy.add(CopyDesc{(const uint32_t *)((uint8_t *)hashes->ptr + off), (uint32_t *)p->tree.p, words, 0, 0},
               (uint32_t)((words + 1023) / 1024));
    plan_tree(o.trees, (uint32_t *)p->tree.p, p->count, 0);
    o.copy.off = o.arena.put(o.copy.v);
    for (auto &t : o.trees) t.off = o.arena.put(t.v);
    s = o.arena.upload();
    if (s) return s;
    s = run_timed(p->ctx, stats, [&]() {
        hipStream_t st = p->ctx->stream;
        launch_copy(st, o.arena.at<CopyDesc>(o.copy.off), o.copy.n(), o.copy.blocks);
        for (auto &t : o.trees) launch_tree(st, o.arena.at<TreeDesc>(t.off), t.n(), t.blocks, t.max_log, 4, t.v.data());
        return AZK_OK;
    });
    if (s) return s;
    HIP_TRY(hipMemcpy(root, (uint8_t *)p->tree.p + ((size_t)2 * p->count - 2) * 32, 32, hipMemcpyDeviceToHost));
    return AZK_OK;
}

// ====================================================================== codeword plan

struct azk_codeword_plan {
    azk_context *ctx;
    uint32_t log_n;
    DeviceMem tw, in, out;
    OneShot shot, shot_m31;
    std::mutex mu;
};

// Final-pass tiles of the M31 codeword: 2^12 words (16 KB plus 16 KB of twiddles).
constexpr uint32_t kFftM31LogTile = 12;

// Circle FFT of one job through the table-driven pass/final kernels (m31: M31 words in and
// out, the passes in place in dst).
static azk_status codeword_enqueue(azk_codeword_plan *p, const void *src, void *dst, bool m31)
{
    OneShot &o = m31 ? p->shot_m31 : p->shot;
    if (!o.matches(src, dst)) {  // tables are reused while the buffers stay the same
        o.reset();
        plan_fft(o.passes, o.final_, src, dst, (const uint32_t *)p->tw.p, p->log_n, m31 ? kFftM31LogTile : kFftOnlyLogTile,
                 true, nullptr, m31, dst);
        for (auto &l : o.passes) l.off = o.arena.put(l.v);
        o.final_.off = o.arena.put(o.final_.v);
        azk_status s = o.arena.upload();
        if (s) return s;
        o.keep(src, dst);
    }
    hipStream_t st = p->ctx->stream;
    for (auto &l : o.passes) launch_fft_pass(st, o.arena.at<FftPassDesc>(l.off), l.n(), l.blocks, l.max_log, m31);
    launch_fft_final(st, o.arena.at<FftFinalDesc>(o.final_.off), o.final_.n(), o.final_.blocks, o.final_.max_log,
                     any_tree(o.final_), m31);
    return AZK_OK;
}

extern "C" azk_status azk_codeword_plan_create(azk_context *ctx, uint32_t log_n, azk_codeword_plan **out)
{
    CHECK_ARG(ctx && out);
    if (log_n < 1 || log_n > 28) return fail(AZK_ERR_INVALID_LAYOUT, "domain log size %u outside 1..28", log_n);
    auto *p = new azk_codeword_plan{ctx, log_n, {}, {}, {}, {}, {}, {}};
    HIP_TRY(hipSetDevice(ctx->device));
    std::lock_guard<std::mutex> gc(ctx->mu);
    azk_status s = materialize_twiddles(ctx, log_n, 1, p->tw, nullptr);
    if (s) { delete p; return s; }
    ctx_retain(p->ctx);
    *out = p;
    return AZK_OK;
}
extern "C" void azk_codeword_plan_destroy(azk_codeword_plan *p) { destroy_on_ctx(p); }

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
        return codeword_enqueue(p, (uint8_t *)coeffs->ptr + coff, (uint8_t *)out->ptr + ooff, false);
    });
}

extern "C" azk_status azk_codeword_execute_m31_resident(azk_codeword_plan *p, const azk_buffer *coeffs, size_t coff,
                                                        azk_buffer *out, size_t ooff, azk_stats *stats)
{
    CHECK_ARG(p && coeffs && out);
    std::lock_guard<std::mutex> g(p->mu);
    const size_t bytes = (size_t)4 << p->log_n;
    azk_status s = check_range(coeffs, coff, bytes);
    if (!s) s = check_range(out, ooff, bytes);
    if (s) return s;
    if (overlaps(coeffs, coff, bytes, out, ooff, bytes)) return fail(AZK_ERR_INVALID_LAYOUT, "coefficient and output ranges alias");
    if (((uintptr_t)coeffs->ptr + coff) % 4 || ((uintptr_t)out->ptr + ooff) % 4)
        return fail(AZK_ERR_INVALID_LAYOUT, "M31 buffers must be 4-byte aligned");
    HIP_TRY(hipSetDevice(p->ctx->device));
    std::lock_guard<std::mutex> gc(p->ctx->mu);
    return run_timed(p->ctx, stats, [&]() {
        return codeword_enqueue(p, (uint8_t *)coeffs->ptr + coff, (uint8_t *)out->ptr + ooff, true);
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
        azk_status e = codeword_enqueue(p, p->in.p, p->out.p, false);
        if (e) return e;
        HIP_TRY(hipMemcpyAsync(evals, p->out.p, bytes, hipMemcpyDeviceToHost, p->ctx->stream));
        return AZK_OK;
    });
}

extern "C" azk_status azk_codeword_execute_m31(azk_codeword_plan *p, const uint32_t *coeffs, size_t in_words,
                                               uint32_t *evals, size_t out_words, azk_stats *stats)
{
    CHECK_ARG(p && coeffs && evals);
    std::lock_guard<std::mutex> g(p->mu);
    const size_t words = (size_t)1 << p->log_n, bytes = 4 * words;
    if (in_words < words || out_words < words) return fail(AZK_ERR_INVALID_LAYOUT, "host arrays shorter than N words");
    HIP_TRY(hipSetDevice(p->ctx->device));
    azk_status s = p->in.ensure(bytes);
    if (!s) s = p->out.ensure(bytes);
    if (s) return s;
    std::lock_guard<std::mutex> gc(p->ctx->mu);
    return run_timed(p->ctx, stats, [&]() -> azk_status {
        HIP_TRY(hipMemcpyAsync(p->in.p, coeffs, bytes, hipMemcpyHostToDevice, p->ctx->stream));
        azk_status e = codeword_enqueue(p, p->in.p, p->out.p, true);
        if (e) return e;
        HIP_TRY(hipMemcpyAsync(evals, p->out.p, bytes, hipMemcpyDeviceToHost, p->ctx->stream));
        return AZK_OK;
    });
}

// executeVerified (CircleCodeword.swift:476-490): the device codeword against the host
// Circle FFT of the same coefficients.
extern "C" azk_status azk_codeword_execute_verified(azk_codeword_plan *p, const void *coeffs, size_t in_bytes,
                                                    void *evals, size_t out_bytes, azk_stats *stats)
{
    azk_status s = azk_codeword_execute(p, coeffs, in_bytes, evals, out_bytes, stats);
    if (s) return s;
    const size_t words = (size_t)4 << p->log_n;
    for (size_t i = 0; i < words; ++i)
        if (((const uint32_t *)coeffs)[i] >= kP) return fail(AZK_ERR_INVALID_LAYOUT, "non-canonical coefficient");
    std::vector<uint32_t> want(words);
    host::circle_fft((const uint32_t *)coeffs, p->log_n, want.data());
    if (memcmp(want.data(), evals, 4 * words))
        return fail(AZK_ERR_CORRECTNESS, "Circle codeword device result did not match the CPU gate");
    return AZK_OK;
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
    DeviceMem fft_tw, fold_tw, a, b, ch, in, out;  // in/out: staging of the host-array variant
    OneShot shot;
    std::mutex mu;
};

extern "C" azk_status azk_fold_plan_create(azk_context *ctx, uint32_t log_n, uint32_t rounds, azk_fold_plan **out)
{
    CHECK_ARG(ctx && out);
    if (log_n < 1 || log_n > 28 || rounds < 1 || rounds > log_n || rounds > (uint32_t)kMaxRounds)
        return fail(AZK_ERR_INVALID_LAYOUT, "fold chain log_n %u rounds %u", log_n, rounds);
    auto *p = new azk_fold_plan{ctx, log_n, rounds, {}, {}, {}, {}, {}, {}, {}, {}, {}};
    HIP_TRY(hipSetDevice(ctx->device));
    std::lock_guard<std::mutex> gc(ctx->mu);
    azk_status s = materialize_twiddles(ctx, log_n, rounds, p->fft_tw, &p->fold_tw);
    if (!s) s = p->a.ensure((size_t)16 << (log_n - 1));
    if (!s) s = p->b.ensure((size_t)16 << (log_n - 1));
    if (!s) s = p->ch.ensure(16 * rounds);
    if (s) { delete p; return s; }
    ctx_retain(p->ctx);
    *out = p;
    return AZK_OK;
}
extern "C" void azk_fold_plan_destroy(azk_fold_plan *p) { destroy_on_ctx(p); }

static azk_status fold_enqueue(azk_fold_plan *p, const void *in, void *out)
{
    OneShot &o = p->shot;
    if (!o.matches(in, out)) {  // tables are reused while the buffers stay the same
        o.reset();
        o.folds_big.resize(p->rounds);
        o.folds_leaf.resize(p->rounds);
        const uint64_t n = 1ull << p->log_n;
        const void *cur = in;
        uint64_t off = 0, m = n;
        for (uint32_t r = 0; r < p->rounds; ++r) {
            void *dst = (r + 1 == p->rounds) ? out : (r % 2 ? p->b.p : p->a.p);
            const uint64_t n_out = m / 2;
            // streaming form (no commitment): 256 outputs per block, the last block partial
            o.folds_big[r].add(FoldDesc{(const uint4 *)cur, (uint4 *)dst, (const uint32_t *)p->fold_tw.p + off,
                                        (const uint32_t *)p->ch.p + 4 * r, nullptr, n_out, 8, 0},
                               (uint32_t)((n_out + 255) >> 8));
            cur = dst;
            off += n_out;
            m = n_out;
        }
        for (auto &l : o.folds_big) l.off = o.arena.put(l.v);
        azk_status s = o.arena.upload();
        if (s) return s;
        o.keep(in, out);
    }
    for (auto &l : o.folds_big)
        launch_fold_stream(p->ctx->stream, o.arena.at<FoldDesc>(l.off), l.n(), l.blocks);
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
    // challenges go up ahead of the timed fold (stream order keeps them in place)
    HIP_TRY(hipMemcpyAsync(p->ch.p, challenges, 16 * p->rounds, hipMemcpyHostToDevice, p->ctx->stream));
    return run_timed(p->ctx, stats, [&]() -> azk_status {
        return fold_enqueue(p, (uint8_t *)evals->ptr + eoff, (uint8_t *)out->ptr + ooff);
    });
}

extern "C" azk_status azk_fold_execute(azk_fold_plan *p, const void *evals, size_t in_bytes, const uint32_t *challenges,
                                       void *out, size_t out_bytes, azk_stats *stats)
{
    CHECK_ARG(p && evals && challenges && out);
    const size_t ib = (size_t)16 << p->log_n, ob = (size_t)16 << (p->log_n - p->rounds);
    if (in_bytes < ib || out_bytes < ob) return fail(AZK_ERR_INVALID_LAYOUT, "host arrays too short");
    azk_buffer bi{}, bo{};
    {
        std::lock_guard<std::mutex> g(p->mu);
        HIP_TRY(hipSetDevice(p->ctx->device));
        azk_status s = p->in.ensure(ib);  // staging kept by the plan between calls
        if (!s) s = p->out.ensure(ob);
        if (s) return s;
        bi = azk_buffer{p->ctx, p->in.p, ib, false};
        bo = azk_buffer{p->ctx, p->out.p, ob, false};
        HIP_TRY(hipMemcpy(p->in.p, evals, ib, hipMemcpyHostToDevice));
    }
    azk_status s = azk_fold_execute_resident(p, &bi, 0, challenges, &bo, 0, stats);
    if (s) return s;
    std::lock_guard<std::mutex> g(p->mu);
    HIP_TRY(hipMemcpy(out, p->out.p, ob, hipMemcpyDeviceToHost));
    return AZK_OK;
}

// CircleFRIFoldChainPlan.executeVerified (CircleFRIFoldPlan.swift:209-225): the device
// chain against the host fold chain.
extern "C" azk_status azk_fold_execute_verified(azk_fold_plan *p, const void *evals, size_t in_bytes,
                                                const uint32_t *challenges, void *out, size_t out_bytes,
                                                azk_stats *stats)
{
    azk_status s = azk_fold_execute(p, evals, in_bytes, challenges, out, out_bytes, stats);
    if (s) return s;
    const size_t words = (size_t)4 << p->log_n, ow = (size_t)4 << (p->log_n - p->rounds);
    for (size_t i = 0; i < words; ++i)
        if (((const uint32_t *)evals)[i] >= kP) return fail(AZK_ERR_INVALID_LAYOUT, "non-canonical evaluation");
    std::vector<uint32_t> want(ow);
    if (!host::fold_chain((const uint32_t *)evals, p->log_n, challenges, p->rounds, want.data()))
        return fail(AZK_ERR_INVALID_LAYOUT, "zero domain coordinate");
    if (memcmp(want.data(), out, 4 * ow)) return fail(AZK_ERR_CORRECTNESS, "fold chain device result did not match the CPU gate");
    return AZK_OK;
}

