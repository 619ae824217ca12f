// Circle domain materialization: the canonical domain's points in storage order, its Circle FFT
// twiddles, and the FRI inverse-domain layers embedded in QM31, materialized on the device once
// per plan and read back or copied into caller buffers.
//
// Reference: CircleDomainMaterializationPlan (Sources/AppleZKProver/CircleDomainMaterializationPlan.swift:
// 28-387, materialize :221-387) with kernels circle_domain_points_materialize,
// circle_codeword_fft_twiddles_materialize, circle_first_fold_twiddles_materialize and
// circle_line_fold_twiddles_materialize (Resources/HashMerkleKernels.metal:1781-1879).  The
// inverse layers are CircleFRILayerOracleV1.inverseDomainLayers (CircleDomain.swift:469-508):
// layer r holds N >> (r + 1) values (t, 0, 0, 0), t = 1/y of the pair's left point for r = 0
// and 1/x doubled r - 1 times for r >= 1 -- the inverses of the first R FFT stages' twiddles,
// which is how they are computed here (k_fft_twiddles, then k_inverse).
#include "device_field.hpp"
#include "kernels.hpp"
#include "runtime_common.hpp"

using namespace azk;
using namespace azk::rt;

namespace {

// Point of storage index s of the canonical 2^log_n domain (bit-reversed natural order; the
// lower half of the natural order is the half coset init + step * i, the upper half its
// conjugates), CircleDomain.swift:273-331.
__device__ cpoint storage_point(uint32_t log_n, uint64_t s)
{
    uint64_t nat = 0, v = s;
    for (uint32_t b = 0; b < log_n; ++b, v >>= 1) nat = (nat << 1) | (v & 1);
    const uint32_t init = 1u << (30 - log_n), step = 1u << (32 - log_n);
    const uint64_t half = 1ull << (log_n - 1);
    uint32_t idx;
    if (nat < half) idx = (uint32_t)((init + (uint64_t)step * nat) & 0x7fffffffu);
    else idx = (uint32_t)((0x80000000u - ((init + (uint64_t)step * (nat - half)) & 0x7fffffffu)) & 0x7fffffffu);
    return cp_from_index(idx);
}

__global__ __launch_bounds__(256) void k_domain_points(uint2 *__restrict__ out, uint32_t log_n)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (1ull << log_n)) return;
    const cpoint p = storage_point(log_n, i);
    out[i] = make_uint2(p.x, p.y);
}

// M31 values as QM31 (t, 0, 0, 0).
__global__ __launch_bounds__(256) void k_embed_qm31(const uint32_t *__restrict__ in, uint4 *__restrict__ out, uint64_t n)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = make_uint4(in[i], 0u, 0u, 0u);
}

uint32_t blocks_of(uint64_t n) { return (uint32_t)((n + 255) / 256); }

} // namespace

struct azk_domain_plan {
    azk_context *ctx = nullptr;
    uint32_t log_n = 0, rounds = 0;
    bool points = false, twiddles = false;
    uint64_t n = 0, inverse_count = 0;
    DeviceMem pts, tw, inv_m31, inv;
    azk_stats stats{};
    std::mutex mu;
};

extern "C" azk_status azk_domain_plan_create(azk_context *ctx, uint32_t log_n, int points, int fft_twiddles,
                                             uint32_t inverse_rounds, azk_domain_plan **out)
{
    CHECK_ARG(ctx && out);
    // size > 1, something to materialize, inverseDomainRoundCount <= logSize (:56-64)
    if (log_n < 1 || log_n > 30 || (!points && !fft_twiddles && !inverse_rounds) || inverse_rounds > log_n)
        return fail(AZK_ERR_INVALID_LAYOUT, "domain materialization: log N %u (1..30), outputs %d/%d/%u", log_n, points,
                    fft_twiddles, inverse_rounds);
    HIP_TRY(hipSetDevice(ctx->device));
    auto *p = new azk_domain_plan();
    p->ctx = ctx;
    p->log_n = log_n;
    p->rounds = inverse_rounds;
    p->points = points != 0;
    p->twiddles = fft_twiddles != 0;
    p->n = 1ull << log_n;
    p->inverse_count = p->n - (p->n >> inverse_rounds);
    azk_status s = AZK_OK;
    if (p->points) s = p->pts.ensure(8 * p->n);
    if (!s && (p->twiddles || p->rounds)) s = p->tw.ensure(4 * p->n);
    if (!s && p->rounds) s = p->inv_m31.ensure(4 * p->inverse_count);
    if (!s && p->rounds) s = p->inv.ensure(16 * p->inverse_count);
    if (!s) {
        std::lock_guard<std::mutex> g(ctx->mu);
        s = run_timed(ctx, &p->stats, [&]() -> azk_status {
            hipStream_t st = ctx->stream;
            if (p->points) hipLaunchKernelGGL(k_domain_points, dim3(blocks_of(p->n)), dim3(256), 0, st, (uint2 *)p->pts.p, log_n);
            if (p->tw.p) launch_fft_twiddles(st, (uint32_t *)p->tw.p, log_n);
            if (p->rounds) {
                launch_inverse(st, (const uint32_t *)p->tw.p, (uint32_t *)p->inv_m31.p, p->inverse_count);
                hipLaunchKernelGGL(k_embed_qm31, dim3(blocks_of(p->inverse_count)), dim3(256), 0, st,
                                   (const uint32_t *)p->inv_m31.p, (uint4 *)p->inv.p, p->inverse_count);
            }
            return AZK_OK;
        });
    }
    if (s) {
        delete p;
        return s;
    }
    ctx_retain(ctx);
    *out = p;
    return AZK_OK;
}

extern "C" void azk_domain_plan_destroy(azk_domain_plan *p) { destroy_on_ctx(p); }

extern "C" azk_status azk_domain_plan_stats(const azk_domain_plan *p, azk_stats *stats)
{
    CHECK_ARG(p && stats);
    *stats = p->stats;
    return AZK_OK;
}

extern "C" azk_status azk_domain_inverse_layout(const azk_domain_plan *p, uint64_t *counts, uint64_t *offsets,
                                                uint32_t rounds)
{
    CHECK_ARG(p && counts && offsets);
    if (rounds != p->rounds) return fail(AZK_ERR_INVALID_LAYOUT, "plan has %u inverse layers, asked for %u", p->rounds, rounds);
    uint64_t off = 0;
    for (uint32_t r = 0; r < rounds; ++r) {  // makeInverseDomainLayout: N >> (r + 1) per round
        counts[r] = p->n >> (r + 1);
        offsets[r] = off;
        off += counts[r];
    }
    return AZK_OK;
}

namespace {

// Device array and byte size of one output; an output the plan does not hold is a layout
// error (require*Buffer / read* throw invalidInputLayout, :134-209).
azk_status output_of(const azk_domain_plan *p, azk_domain_output which, const void **ptr, size_t *bytes)
{
    switch (which) {
    case AZK_DOMAIN_POINTS:
        if (!p->points) break;
        *ptr = p->pts.p;
        *bytes = 8 * p->n;
        return AZK_OK;
    case AZK_DOMAIN_FFT_TWIDDLES:
        if (!p->twiddles) break;
        *ptr = p->tw.p;
        *bytes = 4 * (p->n - 1);
        return AZK_OK;
    case AZK_DOMAIN_INVERSE_LAYERS:
        if (!p->rounds) break;
        *ptr = p->inv.p;
        *bytes = 16 * p->inverse_count;
        return AZK_OK;
    }
    return fail(AZK_ERR_INVALID_LAYOUT, "domain plan did not materialize output %d", (int)which);
}

} // namespace

extern "C" azk_status azk_domain_read(azk_domain_plan *p, azk_domain_output which, void *host, size_t bytes)
{
    CHECK_ARG(p);
    std::lock_guard<std::mutex> g(p->mu);
    const void *src = nullptr;
    size_t need = 0;
    azk_status s = output_of(p, which, &src, &need);
    if (s) return s;
    if (bytes < need || !host) return fail(AZK_ERR_INVALID_LAYOUT, "read of %zu bytes into %zu", need, bytes);
    HIP_TRY(hipSetDevice(p->ctx->device));
    HIP_TRY(hipMemcpy(host, src, need, hipMemcpyDeviceToHost));
    return AZK_OK;
}

extern "C" azk_status azk_domain_copy_resident(azk_domain_plan *p, azk_domain_output which, azk_buffer *dst,
                                               size_t offset)
{
    CHECK_ARG(p && dst);
    std::lock_guard<std::mutex> g(p->mu);
    const void *src = nullptr;
    size_t need = 0;
    azk_status s = output_of(p, which, &src, &need);
    if (!s) s = check_range(dst, offset, need);
    if (s) return s;
    HIP_TRY(hipSetDevice(p->ctx->device));
    std::lock_guard<std::mutex> gc(p->ctx->mu);
    HIP_TRY(hipMemcpyAsync((uint8_t *)dst->ptr + offset, src, need, hipMemcpyDeviceToDevice, p->ctx->stream));
    HIP_TRY(hipStreamSynchronize(p->ctx->stream));
    return AZK_OK;
}
