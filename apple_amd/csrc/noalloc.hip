// Allocation-free SHA3-256 Merkle commitment over raw leaves, for callers that cannot
// allocate device memory (code under hipGraph capture, serving loops): the tree's nodes live
// in a fixed array of this file, one copy per GPU, instead of a plan's buffer.
#include "runtime_common.hpp"

using namespace azk;
using namespace azk::rt;

// The node array of one tree of up to 2^16 leaves ((2^17 - 1) x 32 bytes, 4 MiB), held by
// every GPU the library runs on until the process ends.  Host code never passes `scratch`
// itself to a kernel or a copy: in host code it is a stand-in, not a device address.  The
// address comes from hipGetSymbolAddress, once per device (scratch_of).
constexpr uint32_t kNoAllocMaxLogLeaves = 16;
constexpr size_t kNoAllocScratchBytes = (size_t)32 << (kNoAllocMaxLogLeaves + 1);
static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_scratch[kMaxDevices] = {};

// With `device` current: this device's address of `scratch`, looked up on its first call.
azk_status scratch_of(int device, void **out)
{
    if (device < 0 || device >= kMaxDevices) return fail(AZK_ERR_NO_DEVICE, "device %d out of range", device);
    std::lock_guard<std::mutex> g(g_mu);
    if (!g_scratch[device]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(AZK_ERR_COMMAND_FAILED, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
        g_scratch[device] = ptr;
    }
    *out = g_scratch[device];
    return AZK_OK;
}

} // namespace

extern "C" azk_status azk_merkle_commit_noalloc(azk_context *ctx, const azk_buffer *leaves, size_t leaves_offset,
                                                uint32_t leaf_count, uint32_t leaf_stride, uint32_t leaf_length,
                                                azk_buffer *root, size_t root_offset, void *stream)
{
    CHECK_ARG(ctx && leaves && root);
    if (!is_pow2(leaf_count) || leaf_count > (1u << kNoAllocMaxLogLeaves))
        return fail(AZK_ERR_INVALID_LEAF_COUNT, "allocation-free commit: leaf count must be a power of two of at most %u, got %u",
                    1u << kNoAllocMaxLogLeaves, leaf_count);
    if (leaf_length > 136)
        return fail(AZK_ERR_UNSUPPORTED_ONE_BLOCK_LENGTH, "leaves are hashed as one block (<= 136 bytes), got %u", leaf_length);
    if (leaf_stride < leaf_length) return fail(AZK_ERR_INVALID_LAYOUT, "leaf stride shorter than leaf length");
    azk_status s = check_range(leaves, leaves_offset, (size_t)leaf_count * leaf_stride);
    if (!s) s = check_range(root, root_offset, 32);
    if (s) return s;
    if (leaves->ctx != ctx || root->ctx != ctx) return fail(AZK_ERR_INVALID_ARGUMENT, "buffers of another context");
    HIP_TRY(hipSetDevice(ctx->device));
    void *ptr = nullptr;
    s = scratch_of(ctx->device, &ptr);
    if (s) return s;
    hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
    uint32_t *tree = (uint32_t *)ptr;  // the device address, never `scratch` itself
    const uint64_t m = leaf_count;
    merkle_launch(st, (const uint8_t *)leaves->ptr + leaves_offset, tree, m, leaf_stride, leaf_length,
                  merkle_default_log_tile(ilog2(m)));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync((uint8_t *)root->ptr + root_offset, tree + 8 * (2 * m - 2), 32, hipMemcpyDeviceToDevice, st));
    return AZK_OK;
}
