// gfx950 kernels for the Circle-PCS commit path.
//
// Reference kernels these replace (Sources/AppleZKProver/Resources/HashMerkleKernels.metal):
//   keccak_f1600_permutation_scalar/simdgroup :428-478   -> k_keccakf_batch
//   sha3_256/keccak_256_oneblock_specialized  :791-990   -> k_oneblock_hash
//   sha3_256_merkle_treelet_leaves/parents/fuse_upper :1082-1218 -> k_merkle_raw_tiles,
//                                                        k_tree_reduce
//   circle_codeword_fft_stage(+_leaf_hash)    :1881-1943 -> k_fft_pass, k_fft_final
//   qm31_fri_fold(_challenge_buffer)          :1987-2032 -> k_fold
//   transcript_absorb_keccak / squeeze        :1324-1404 -> k_transcript
//   circle domain / twiddle materialization   :1781-1879 -> k_fft_twiddles, k_inverse
//   extract_sibling + query extraction        :1149-1165 -> k_emit_proof
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <type_traits>

#include "device_field.hpp"
#include "device_keccak.hpp"
#include "kernels.hpp"

namespace azk {

// HBM through global (address space 1) pointers, never flat ones: a flat access also
// counts against lgkmcnt, so every later LDS wait of the wave (one per tree level, one
// per FFT stage pair) would wait for the store to reach memory as well.
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) uint32_t gu32;
typedef __attribute__((address_space(1))) u32x2 gu32x2;
typedef __attribute__((address_space(1))) u32x4 gu32x4;
template <class T>
__device__ __forceinline__ __attribute__((address_space(1))) T *glob(T *p)
{
    return (__attribute__((address_space(1))) T *)p;
}
template <class T>
__device__ __forceinline__ const __attribute__((address_space(1))) T *glob(const T *p)
{
    return (const __attribute__((address_space(1))) T *)p;
}
__device__ __forceinline__ uint4 ld_g4(const uint4 *p)
{
    const u32x4 v = *(const gu32x4 *)glob(p);
    return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void st_g4(uint4 *p, uint4 v) { *(gu32x4 *)glob(p) = u32x4{v.x, v.y, v.z, v.w}; }


// ============================================================== batch hashing

// One Keccak-f[1600] per lane over strided 25 x u64 states; output padding zeroed.
__global__ __launch_bounds__(256) void k_keccakf_batch(const uint8_t *__restrict__ in, uint32_t in_stride,
                                                       uint8_t *__restrict__ out, uint32_t out_stride, uint32_t count)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const uint2 *src = reinterpret_cast<const uint2 *>(in + (size_t)i * in_stride);
    u2 a[25];
#pragma unroll
    for (int k = 0; k < 25; ++k) {
        uint2 v = src[k];
        a[k] = {v.x, v.y};
    }
    kf_v(a);
    uint2 *dst = reinterpret_cast<uint2 *>(out + (size_t)i * out_stride);
#pragma unroll
    for (int k = 0; k < 25; ++k) dst[k] = make_uint2(a[k].lo, a[k].hi);
    for (uint32_t b = 200; b < out_stride; b += 8) *reinterpret_cast<uint2 *>(out + (size_t)i * out_stride + b) = make_uint2(0, 0);
}

// Grinding search (CircleFRITranscriptV1.findGrindingNonce, CircleProofFormat.swift:752-791):
// lane i absorbs nonce base + i (8 LE bytes, SHA3 padding) into the state that follows the
// grinding frame and tests digestMeetsTarget (:663-681) on the first state bytes, most
// significant bit of each byte first.  The smallest hit of the launch is folded into *best
// with a 64-bit atomic minimum.  bits <= 32.
__global__ __launch_bounds__(256) void k_grind(const uint64_t *__restrict__ state, uint32_t bits, uint64_t base,
                                               unsigned long long *best)
{
    const uint64_t nonce = base + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    u2 a[25];
#pragma unroll
    for (int k = 0; k < 25; ++k) {
        const uint64_t v = state[k];
        a[k] = {(uint32_t)v, (uint32_t)(v >> 32)};
    }
    a[0].lo ^= (uint32_t)nonce;
    a[0].hi ^= (uint32_t)(nonce >> 32);
    a[1].lo ^= 0x06u;
    a[16].hi ^= 0x80000000u;
    kf_v(a);
    const uint32_t whole = bits >> 3, rem = bits & 7;
    uint32_t mask = whole >= 4 ? 0xffffffffu : (1u << (8 * whole)) - 1u;  // the zero bytes
    if (rem && whole < 4) mask |= ((0xffu << (8 - rem)) & 0xffu) << (8 * whole);
    if ((a[0].lo & mask) == 0) atomicMin(best, (unsigned long long)nonce);
}

// Loads the first `len` bytes of a message into 34 LE words, zero beyond.
template <bool kAligned>
__device__ __forceinline__ void load_message(const uint8_t *__restrict__ p, uint32_t len, uint32_t (&w)[34])
{
    if constexpr (kAligned) {
        const auto *q = glob(reinterpret_cast<const uint32_t *>(p));
#pragma unroll
        for (int k = 0; k < 34; ++k) {
            uint32_t lo = 4u * k;
            uint32_t v = lo < len ? q[k] : 0u;
            if (lo < len && len - lo < 4) v &= (1u << (8 * (len - lo))) - 1u;
            w[k] = v;
        }
    } else {
#pragma unroll
        for (int k = 0; k < 34; ++k) {
            uint32_t v = 0;
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                uint32_t idx = 4u * k + b;
                if (idx < len) v |= (uint32_t)glob(p)[idx] << (8 * b);
            }
            w[k] = v;
        }
    }
}

// One-block SHA3-256 (suffix 0x06) / Keccak-256 (suffix 0x01) per lane.
template <bool kAligned>
__global__ __launch_bounds__(256) void k_oneblock_hash(const uint8_t *__restrict__ in, uint32_t stride, uint32_t len,
                                                       uint8_t *__restrict__ out, uint32_t out_stride, uint32_t count,
                                                       uint32_t suffix)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    uint32_t w[34];
    load_message<kAligned>(in + (size_t)i * stride, len, w);
    uint32_t d[8];
    sponge_oneblock(w, len, suffix, d);
    uint8_t *o = out + (size_t)i * out_stride;
    if ((reinterpret_cast<uintptr_t>(o) & 15u) == 0) {
        reinterpret_cast<uint4 *>(o)[0] = make_uint4(d[0], d[1], d[2], d[3]);
        reinterpret_cast<uint4 *>(o)[1] = make_uint4(d[4], d[5], d[6], d[7]);
    } else {
#pragma unroll
        for (int k = 0; k < 32; ++k) o[k] = (uint8_t)(d[k / 4] >> (8 * (k % 4)));
    }
    for (uint32_t b = 32; b < out_stride; ++b) o[b] = 0;
}


// Kernels that run the pinned-register Keccak rounds (100 VGPRs of state, rho/pi output and
// temporaries) are held to 128 VGPRs: four waves per SIMD.  Left to itself the compiler
// takes 131-136 for the fold and raw-leaf Merkle kernels (three waves per SIMD).
#define AZK_FOUR_WAVES __attribute__((amdgpu_waves_per_eu(4, 8)))

// ============================================================== descriptor lookup

// Finds the descriptor whose block range holds blockIdx.x (tables are sorted by
// first_block; the search is wave-uniform) and the block's index inside it.
template <class D>
__device__ __forceinline__ const D &find_desc(const D *__restrict__ t, uint32_t n, uint32_t &local)
{
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (t[mid].first_block <= blockIdx.x) lo = mid;
        else hi = mid;
    }
    local = blockIdx.x - t[lo].first_block;
    return t[lo];
}

// ============================================================== Merkle tiles

// Node `idx` of level `l` in a full-tree array of a tree with `m` leaves:
// levels are stored bottom-up, level l starting at node 2m - (2m >> l).
__device__ __forceinline__ gu32 *node_ptr(uint32_t *tree, uint64_t m, int l, uint64_t idx)
{
    return glob(tree) + 8 * ((2 * m - ((2 * m) >> l)) + idx);
}

// A node in LDS (st_node / ld_node on a uint32_t *) or in HBM (on a gu32 *).
__device__ __forceinline__ void st_node(uint32_t *p, const uint32_t (&d)[8])
{
    reinterpret_cast<uint4 *>(p)[0] = make_uint4(d[0], d[1], d[2], d[3]);
    reinterpret_cast<uint4 *>(p)[1] = make_uint4(d[4], d[5], d[6], d[7]);
}
__device__ __forceinline__ void ld_node(const uint32_t *p, uint32_t (&d)[8])
{
    uint4 x = reinterpret_cast<const uint4 *>(p)[0], y = reinterpret_cast<const uint4 *>(p)[1];
    d[0] = x.x; d[1] = x.y; d[2] = x.z; d[3] = x.w; d[4] = y.x; d[5] = y.y; d[6] = y.z; d[7] = y.w;
}
__device__ __forceinline__ void st_node(gu32 *p, const uint32_t (&d)[8])
{
    reinterpret_cast<gu32x4 *>(p)[0] = u32x4{d[0], d[1], d[2], d[3]};
    reinterpret_cast<gu32x4 *>(p)[1] = u32x4{d[4], d[5], d[6], d[7]};
}
__device__ __forceinline__ void ld_node(const gu32 *p, uint32_t (&d)[8])
{
    const u32x4 x = reinterpret_cast<const gu32x4 *>(p)[0], y = reinterpret_cast<const gu32x4 *>(p)[1];
    d[0] = x.x; d[1] = x.y; d[2] = x.z; d[3] = x.w; d[4] = y.x; d[5] = y.y; d[6] = y.z; d[7] = y.w;
}

// Loops over fewer items than threads use the low waves of a block.  Measured on MI355X (HW_ID
// of every wave of 1024 blocks of 4 waves): blocks that share a CU are 256 apart in block
// index and the dispatcher puts wave 0 of each on a different SIMD, so the active low waves of
// co-resident blocks already spread over the CU's SIMDs.

// In-place Merkle levels in LDS: the nodes of the current level sit `s` node slots (8s
// words) apart; parent q of nodes 2q and 2q + 1 overwrites node 2q's slot, so the next
// level sits 2s slots apart.  Each parent reads and writes only its own children's slots,
// so one barrier per level suffices and the reduction needs no second LDS region.

// Parents q < half, one per wave at a time with the one-state-per-wave permutation (lanes
// holding words < 8 load the 64-byte message, the own lanes of words 0..3 write the
// digest).  The loop is wave-uniform.
__device__ __forceinline__ void wave_parents(uint32_t *A, uint32_t half, uint32_t s, int l, uint64_t g,
                                             uint32_t *tree, uint64_t m, const WaveLane &c)
{
    const uint32_t waves = blockDim.x >> 6;
    for (uint32_t q = threadIdx.x >> 6; q < half; q += waves) {
        uint32_t *left = A + 16 * s * q;
        uint32_t lo = 0, hi = 0;
        if (c.word < 8) {  // words 0..3: left child, 4..7: right child (8s words further)
            const uint32_t at = c.word < 4 ? 2 * c.word : 8 * s + 2 * (c.word - 4);
            lo = left[at];
            hi = left[at + 1];
        }
        if (c.word == 8) lo = 0x06u;          // SHA3 padding after 64 bytes
        if (c.word == 16) hi = 0x80000000u;   // final bit of the 136-byte block
        kf_w(lo, hi, c);                      // every load of this wave has returned
        if (c.own && c.word < 4) {
            left[2 * c.word] = lo;
            left[2 * c.word + 1] = hi;
            gu32 *node = node_ptr(tree, m, l + 1, (g >> 1) + q);
            reinterpret_cast<gu32x2 *>(node)[c.word] = u32x2{lo, hi};
        }
    }
}

// Parent q by a lane pair with the two-lanes-per-state permutation: lane `hi` of the pair
// loads and writes the high (odd) 32-bit word of each 64-bit message and digest word.
__device__ __forceinline__ void pair_parent(uint32_t *A, uint32_t q, uint32_t s, bool hi, int l, uint64_t g,
                                            uint32_t *tree, uint64_t m)
{
    uint32_t *left = A + 16 * s * q, *right = left + 8 * s;
    uint32_t a[25];
#pragma unroll
    for (int w = 0; w < 25; ++w) a[w] = 0;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        a[w] = left[2 * w + hi];
        a[4 + w] = right[2 * w + hi];
    }
    a[8] = hi ? 0u : 0x06u;           // SHA3 padding after 64 bytes
    a[16] = hi ? 0x80000000u : 0u;    // final bit of the 136-byte block
    kf_h(a, hi);
    gu32 *node = node_ptr(tree, m, l + 1, (g >> 1) + q);
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        left[2 * w + hi] = a[w];
        node[2 * w + hi] = a[w];
    }
}

// Reduces `cnt` level-`l` nodes held contiguously in LDS (node q at words 8q..8q+7),
// whose first node has global index `g` in that level, up to one node at A[0..7].  Every
// computed level is written to the tree in HBM.  By the number of parents of a level:
// one lane per parent while the lane pairs would take more than 1/kPairDiv of the block
// (~10 us a level alone on a SIMD, but half the issue of the pair form when the SIMDs are
// full); down to kWaveParents per wave, two lanes per parent (5.4 us); the serial top, one
// wave per parent (2.5 us).  kPairDiv 2 in the big-tile folds and in latency-form fold
// launches of two or more blocks per CU (2^21 commit: the 2^17 and 2^18 latency-form folds
// 71 and 108 us instead of 77 and 112); 1 in the final FFT pass's tiles (482 us with 2 against
// 475) and in latency-form launches of one block per CU.
#ifndef AZK_WAVE_PARENTS
#define AZK_WAVE_PARENTS 2
#endif
constexpr uint32_t kWaveParents = AZK_WAVE_PARENTS;
template <uint32_t kPairDiv = 1>
__device__ __forceinline__ void reduce_lds(uint32_t *A, uint32_t cnt, int l, uint64_t g, uint32_t *tree, uint64_t m)
{
    const uint32_t wave_max = kWaveParents * (blockDim.x >> 6);
    const WaveLane c = wave_lane();
    const uint32_t tid = threadIdx.x;
    uint32_t s = 1;
    while (cnt > 1) {
        uint32_t half = cnt >> 1;
        if (half <= wave_max) {
            wave_parents(A, half, s, l, g, tree, m, c);
        } else if (2 * half * kPairDiv <= blockDim.x) {
            if ((tid >> 1) < half) pair_parent(A, tid >> 1, s, tid & 1u, l, g, tree, m);  // pairs are whole
        } else {
            for (uint32_t q = tid; q < half; q += blockDim.x) {
                uint32_t lft[8], rgt[8], o[8];
                ld_node(A + 16 * s * q, lft);
                ld_node(A + 16 * s * q + 8 * s, rgt);
                sha3_parent(lft, rgt, o);
                st_node(A + 16 * s * q, o);
                st_node(node_ptr(tree, m, l + 1, (g >> 1) + q), o);
            }
        }
        __syncthreads();
        cnt = half;
        s *= 2;
        ++l;
        g >>= 1;
    }
}

// Hashes the T QM31 leaves held in LDS (leaf i at words 4i..4i+3, global leaf index
// g0 + i) and builds the tile's subtree up to its root, writing every node.  Leaf pair
// p's 32 bytes are overwritten in place by its parent.  Nothing but the loop index is live
// across a permutation: the second leaf is read after the first is hashed, and the first
// digest waits in the pair's LDS slot (the permutation's own state and temporaries take
// 100 VGPRs, and 128 is the budget for four waves per SIMD).
template <uint32_t kPairDiv = 1>
__device__ __forceinline__ void tile_tree_qm31(uint32_t *A, uint32_t T, uint64_t g0, uint32_t *tree, uint64_t m)
{
    if (T == 1) {
        uint32_t v[4] = {A[0], A[1], A[2], A[3]}, h[8];
        if (threadIdx.x == 0) {
            sha3_leaf16(v, h);
            st_node(node_ptr(tree, m, 0, g0), h);
        }
        return;
    }
    for (uint32_t p = threadIdx.x; p < T / 2; p += blockDim.x) {
        const uint4 e0 = reinterpret_cast<const uint4 *>(A)[2 * p];
        uint32_t v0[4] = {e0.x, e0.y, e0.z, e0.w}, h0[8], h1[8], par[8];
        sha3_leaf16(v0, h0);
        st_node(node_ptr(tree, m, 0, g0 + 2 * p), h0);
        const uint4 e1 = reinterpret_cast<const uint4 *>(A)[2 * p + 1];
        uint32_t v1[4] = {e1.x, e1.y, e1.z, e1.w};
        st_node(A + 8 * p, h0);  // over both leaves of the pair; e1 is in registers
        sha3_leaf16(v1, h1);
        st_node(node_ptr(tree, m, 0, g0 + 2 * p + 1), h1);
        ld_node(A + 8 * p, h0);
        sha3_parent(h0, h1, par);
        st_node(node_ptr(tree, m, 1, (g0 >> 1) + p), par);
        st_node(A + 8 * p, par);
    }
    __syncthreads();
    reduce_lds<kPairDiv>(A, T / 2, 1, g0 >> 1, tree, m);
}

// Raw-leaf Merkle tile: leaves of `len` <= 136 bytes at `stride`, T leaves per
// block.  Reference: SHA3RawLeavesMerkleCommitPlan.encodeCommitmentRoot
// (MerkleCommitter.swift:801-894) with the treelet kernel :1167-1218.
template <bool kAligned>
__device__ __forceinline__ void merkle_raw_tile(const MerkleRawDesc &d, uint32_t tile)
{
    extern __shared__ uint32_t lds[];
    const uint32_t T = 1u << d.log_t;
    const uint32_t stride = d.stride, len = d.len;
    uint32_t *tree = d.tree;
    const uint64_t m = d.m;
    uint32_t *A = lds;
    const uint64_t g0 = (uint64_t)tile * T;
    if (T == 1) {
        if (threadIdx.x == 0) {
            uint32_t w[34], h[8];
            load_message<kAligned>(d.leaves + g0 * stride, len, w);
            sponge_oneblock(w, len, 0x06u, h);
            st_node(node_ptr(tree, m, 0, g0), h);
        }
        return;
    }
    // The first leaf's digest waits in the pair's LDS slot (which the parent overwrites)
    // while the second leaf is hashed, so no digest is live across a permutation.
    for (uint32_t p = threadIdx.x; p < T / 2; p += blockDim.x) {
#pragma unroll 1
        for (int s = 0; s < 2; ++s) {
            uint32_t w[34], h[8];
            load_message<kAligned>(d.leaves + (g0 + 2 * p + s) * (uint64_t)stride, len, w);
            sponge_oneblock(w, len, 0x06u, h);
            st_node(node_ptr(tree, m, 0, g0 + 2 * p + s), h);
            if (s == 0) {
                st_node(A + 8 * p, h);
            } else {
                uint32_t h0[8], par[8];
                ld_node(A + 8 * p, h0);
                sha3_parent(h0, h, par);
                st_node(node_ptr(tree, m, 1, (g0 >> 1) + p), par);
                st_node(A + 8 * p, par);
            }
        }
    }
    __syncthreads();
    reduce_lds(A, T / 2, 1, g0 >> 1, tree, m);
}

template <bool kAligned>
__global__ AZK_FOUR_WAVES __launch_bounds__(256) void k_merkle_raw_tiles(const MerkleRawDesc *__restrict__ tab, uint32_t ntab)
{
    uint32_t tile;
    const MerkleRawDesc &d = find_desc(tab, ntab, tile);
    merkle_raw_tile<kAligned>(d, tile);
}

// One job's tiles with the descriptor as a kernel argument (no table in device memory).
template <bool kAligned>
__global__ AZK_FOUR_WAVES __launch_bounds__(256) void k_merkle_raw_tiles1(const MerkleRawDesc d)
{
    merkle_raw_tile<kAligned>(d, blockIdx.x);
}

__device__ __forceinline__ void run_transcript(const TranscriptOp *ops, uint64_t *state, uint32_t nops);
__device__ __forceinline__ void run_transcript_from(const TranscriptOp *ops, uint64_t *state, uint32_t nops, uint32_t lo,
                                                    uint32_t hi);

// Upper Merkle levels: each block loads 2^log_c nodes of `level` and reduces them
// by log_c levels (MerkleCommitter.swift:853-893: parents per level + fused upper tree).
// A block of a job's last step then runs the job's transcript program for that root.
// Blocks of 4, 8 or 16 waves (the execution plan's tree_waves).
__device__ __forceinline__ void tree_reduce(const TreeDesc &d, uint32_t chunk)
{
    extern __shared__ uint32_t lds[];
    const uint32_t C = 1u << d.log_c;
    uint32_t *tree = d.tree;
    const uint64_t m = d.m;
    const int l = (int)d.level;
    uint32_t *A = lds;
    const uint64_t g = (uint64_t)chunk * C;
    // the root's block loads the transcript state beside its nodes, not after the levels
    uint32_t slo = 0, shi = 0;
    if (d.t_ops && threadIdx.x < 64) {
        const uint32_t w = wave_lane().word;
        if (w < 25) {
            const uint64_t v = glob(d.t_state)[w];
            slo = (uint32_t)v;
            shi = (uint32_t)(v >> 32);
        }
    }
    for (uint32_t q = threadIdx.x; q < C; q += blockDim.x) {
        uint32_t v[8];
        ld_node(node_ptr(tree, m, l, g + q), v);
        st_node(A + 8 * q, v);
    }
    __syncthreads();
    reduce_lds(A, C, l, g, tree, m);
    if (d.t_ops) {  // uniform per block; the barrier makes the root's HBM words visible to wave 0
        __syncthreads();
        if (threadIdx.x < 64) run_transcript_from(d.t_ops, d.t_state, d.t_nops, slo, shi);
    }
}

__global__ AZK_FOUR_WAVES __launch_bounds__(1024) void k_tree_reduce(const TreeDesc *__restrict__ tab, uint32_t ntab)
{
    uint32_t chunk;
    const TreeDesc &d = find_desc(tab, ntab, chunk);
    tree_reduce(d, chunk);
}

// A launch of one tree step takes its descriptor as a kernel argument: the node loads
// then wait on no table read (one dependent HBM round trip less per launch).
__global__ AZK_FOUR_WAVES __launch_bounds__(1024) void k_tree_reduce1(const TreeDesc d)
{
    tree_reduce(d, blockIdx.x);
}

// Word copies, 1024 words per block.
__global__ __launch_bounds__(256) void k_copy(const CopyDesc *__restrict__ tab, uint32_t ntab)
{
    uint32_t blk;
    const CopyDesc &d = find_desc(tab, ntab, blk);
    const uint64_t base = (uint64_t)blk * 1024;
    for (uint32_t i = threadIdx.x; i < 1024; i += blockDim.x)
        if (base + i < d.words) d.dst[base + i] = d.src[base + i];
}

// ============================================================== Circle domain

// FFT twiddles, stage-major (N - 1 words): stage 0 holds y, stage s >= 1 holds x
// doubled s - 1 times, of the domain point at storage index h << (s + 1).
// Reference: CircleCodewordOracle.circleFFTTwiddles (CircleCodeword.swift:129-171),
// canonical domain CircleDomain.swift:225-239,273-285.
__global__ __launch_bounds__(256) void k_fft_twiddles(uint32_t *out, uint32_t log_n)
{
    const uint64_t n = 1ull << log_n;
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    // locate stage s with offset(s) = n - (n >> s) <= i < offset(s + 1)
    uint32_t s = 0;
    while (i >= n - (n >> (s + 1))) ++s;
    uint64_t h = i - (n - (n >> s));
    uint64_t storage = h << (s + 1);
    // bit-reverse storage -> natural index
    uint64_t nat = 0, v = storage;
    for (uint32_t b = 0; b < log_n; ++b) { nat = (nat << 1) | (v & 1); v >>= 1; }
    const uint32_t init = 1u << (31 - (log_n + 1)), step = 1u << (31 - (log_n - 1));
    const uint64_t half = n >> 1;
    uint32_t idx;
    if (nat < half) idx = (uint32_t)((init + (uint64_t)step * nat) & 0x7fffffffu);
    else idx = (uint32_t)((0x80000000u - ((init + (uint64_t)step * (nat - half)) & 0x7fffffffu)) & 0x7fffffffu);
    cpoint p = cp_from_index(idx);
    uint32_t w;
    if (s == 0) w = p.y;
    else {
        w = p.x;
        for (uint32_t k = 1; k < s; ++k) w = double_x(w);
    }
    out[i] = w;
}

// Elementwise M31 inverse (fold twiddles are the inverses of the first FFT stages:
// inverse y for the circle fold, inverse x for the line folds, CircleDomain.swift:399-511).
__global__ __launch_bounds__(256) void k_inverse(const uint32_t *__restrict__ in, uint32_t *__restrict__ out, uint64_t n)
{
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = m31_inv(in[i]);
}

// ============================================================== Circle FFT

__device__ __forceinline__ qm31 ld_qm(const uint4 &v) { return {v.x, v.y, v.z, v.w}; }
__device__ __forceinline__ uint4 st_qm(qm31 q) { return make_uint4(q.a, q.b, q.c, q.d); }

// Element forms of the FFT.  QmElem: a QM31 value (16 bytes).  M31Elem: one M31 word, for
// a codeword of M31 coefficients -- the QM31 lift (a, 0, 0, 0) that the reference transforms
// keeps its three upper components at zero through every butterfly (the twiddles are M31),
// so only the first component is computed and moved: a quarter of the arithmetic and bytes.
struct QmElem {
    typedef uint4 S;
    typedef qm31 V;
    static constexpr uint32_t kMaxPerThread = 8;
    __device__ static V ld(const S &x) { return ld_qm(x); }
    __device__ static S st(V v) { return st_qm(v); }
    __device__ static V scale(V v, uint32_t w) { return qm_scale(v, w); }
    __device__ static V add(V a, V b) { return qm_add(a, b); }
    __device__ static V sub(V a, V b) { return qm_sub(a, b); }
    __device__ static S load_global(const S *p) { return ld_g4(p); }
    __device__ static void store_global(S *p, S v) { st_g4(p, v); }
};
struct M31Elem {
    typedef uint32_t S;
    typedef uint32_t V;
    static constexpr uint32_t kMaxPerThread = 16;
    __device__ static V ld(const S &x) { return x; }
    __device__ static S st(V v) { return v; }
    __device__ static V scale(V v, uint32_t w) { return m31_mul(v, w); }
    __device__ static V add(V a, V b) { return m31_add(a, b); }
    __device__ static V sub(V a, V b) { return m31_sub(a, b); }
    __device__ static S load_global(const S *p) { return *glob(p); }
    __device__ static void store_global(S *p, S v) { *glob(p) = v; }
};

template <class E>
__device__ __forceinline__ void butterfly(typename E::S *s, uint32_t i1, uint32_t i2, uint32_t w)
{
    const typename E::V l = E::ld(s[i1]), r = E::scale(E::ld(s[i2]), w);
    s[i1] = E::st(E::add(l, r));
    s[i2] = E::st(E::sub(l, r));
}

typedef __attribute__((address_space(1))) const u32x4 global_u32x4;
typedef __attribute__((address_space(1))) const uint32_t global_u32;

// The twiddles a tile needs, staged in LDS: stage lo + t uses 2^(k-1-t) consecutive
// words tw[off(lo + t) + (high << (k-1-t)) + j], stored at tws[2^(k-1-t) + j] (slot 0
// is unused, so the 2^k slots load with a uniform trip count).  off(s) = n - (n >> s)
// (CircleCodeword.swift:166-171).
__device__ __forceinline__ void stage_twiddles(uint32_t *tws, const uint32_t *__restrict__ tw, uint64_t n, uint32_t lo,
                                               uint32_t k, uint64_t high)
{
    const global_u32 *g = (const global_u32 *)tw;
    auto src = [&](uint32_t i) -> uint64_t {
        if (i == 0) return 0;
        const uint32_t level = 31 - __builtin_clz(i), sg = lo + k - 1 - level;
        return (sg ? n - (n >> sg) : 0) + (high << level) + (i - (1u << level));
    };
    const uint32_t S = 1u << k, B = blockDim.x;  // S <= 16B
    if (S >= B) {
        const uint32_t J = S / B;
        uint32_t v[16];
#pragma unroll
        for (uint32_t j = 0; j < 16; ++j)
            if (j < J) v[j] = g[src(threadIdx.x + B * j)];
#pragma unroll
        for (uint32_t j = 0; j < 16; ++j)
            if (j < J) tws[threadIdx.x + B * j] = v[j];
    } else if (threadIdx.x < S) {
        tws[threadIdx.x] = g[src(threadIdx.x)];
    }
}

// Stages lo + k - 1 .. lo of the Circle FFT on a tile of T = 2^(k + m) elements held
// in LDS as 2^k strided rows (stride 2^lo) x 2^m contiguous columns.  Butterfly b of
// stage t pairs i1 = b with a zero bit inserted at position t + m and i2 = i1 + 2^(t+m);
// its twiddle is tws[2^(k-1-t) + (b >> (t + m))].  Reference stage loop:
// CircleCodewordOracle.evaluateWithCircleFFT (CircleCodeword.swift:72-98), Metal
// circle_codeword_fft_stage :1881-1908.
//
// Stages run in pairs (t, t - 1): group g of a pair holds the four elements i0 + j 2^sh
// (sh = t - 1 + m, j < 4, i0 = g with two zero bits inserted at sh), so stage t's two
// butterflies (j = 0/2 and 1/3, twiddle index i0 >> (sh + 2)) and stage t - 1's two
// (0/1 and 2/3, twiddle indices 2 hp and 2 hp + 1) run in registers between one LDS load
// and one store: half the LDS traffic and barriers of one stage per round trip.  An odd
// stage count leaves stage 0 as a single radix-2 stage.
template <class E>
__device__ __forceinline__ void fft_tile_stages(typename E::S *s, const uint32_t *tws, uint32_t k, uint32_t m)
{
    int t = (int)k - 1;
    // M31 words: radix-8 rounds (stages t, t - 1, t - 2 between one LDS load and one store) while
    // a block has at least one group of eight per thread: group g holds i0 + j 2^sh (sh = t - 2 + m,
    // j < 8); stage t pairs j with j + 4 (twiddle hp), stage t - 1 pairs j with j + 2 (twiddle
    // 2 hp + j/4) and stage t - 2 pairs j with j + 1 (twiddle 4 hp + j/2), the same twiddle slots
    // as one stage at a time.  Measured on 2^26 codewords: 0.628 ms against 0.645 for M31 words;
    // for QM31 values (eight in 32 VGPRs per thread) 1.60 ms against 1.57-1.60, so QM31 stays
    // radix-4.
    if (std::is_same<E, M31Elem>::value && k + m >= 3 && (1u << (k + m - 3)) >= blockDim.x) {
        for (; t >= 2; t -= 3) {
            const uint32_t groups = 1u << (k + m - 3), sh = (uint32_t)t - 2 + m, d = 1u << sh;
            const uint32_t *tw_t = tws + (1u << (k - 1 - t)), *tw_u = tws + (1u << (k - t)),
                           *tw_v = tws + (1u << (k + 1 - t));
            for (uint32_t g = threadIdx.x; g < groups; g += blockDim.x) {
                const uint32_t hp = g >> sh;
                const uint32_t i0 = (hp << (sh + 3)) | (g & (d - 1));
                typename E::V a[8];
#pragma unroll
                for (uint32_t j = 0; j < 8; ++j) a[j] = E::ld(s[i0 + j * d]);
                const uint32_t w = tw_t[hp];
#pragma unroll
                for (uint32_t j = 0; j < 4; ++j) {
                    const typename E::V r = E::scale(a[j + 4], w);
                    a[j + 4] = E::sub(a[j], r);
                    a[j] = E::add(a[j], r);
                }
#pragma unroll
                for (uint32_t h = 0; h < 2; ++h) {
                    const uint32_t wu = tw_u[2 * hp + h];
#pragma unroll
                    for (uint32_t j = 4 * h; j < 4 * h + 2; ++j) {
                        const typename E::V r = E::scale(a[j + 2], wu);
                        a[j + 2] = E::sub(a[j], r);
                        a[j] = E::add(a[j], r);
                    }
                }
#pragma unroll
                for (uint32_t q = 0; q < 4; ++q) {
                    const uint32_t wv = tw_v[4 * hp + q];
                    const typename E::V r = E::scale(a[2 * q + 1], wv);
                    s[i0 + (2 * q + 1) * d] = E::st(E::sub(a[2 * q], r));
                    s[i0 + 2 * q * d] = E::st(E::add(a[2 * q], r));
                }
            }
            __syncthreads();
        }
    }
    for (; t >= 1; t -= 2) {
        const uint32_t quarter = 1u << (k + m - 2), sh = (uint32_t)t - 1 + m, d = 1u << sh;
        const uint32_t *tw_hi = tws + (1u << (k - 1 - t)), *tw_lo = tws + (1u << (k - t));
        for (uint32_t g = threadIdx.x; g < quarter; g += blockDim.x) {
            const uint32_t hp = g >> sh;
            const uint32_t i0 = (hp << (sh + 2)) | (g & (d - 1));
            typename E::V a0 = E::ld(s[i0]), a1 = E::ld(s[i0 + d]), a2 = E::ld(s[i0 + 2 * d]), a3 = E::ld(s[i0 + 3 * d]);
            const uint32_t w = tw_hi[hp], wa = tw_lo[2 * hp], wb = tw_lo[2 * hp + 1];
            typename E::V r = E::scale(a2, w);
            a2 = E::sub(a0, r);
            a0 = E::add(a0, r);
            r = E::scale(a3, w);
            a3 = E::sub(a1, r);
            a1 = E::add(a1, r);
            r = E::scale(a1, wa);
            s[i0 + d] = E::st(E::sub(a0, r));
            s[i0] = E::st(E::add(a0, r));
            r = E::scale(a3, wb);
            s[i0 + 3 * d] = E::st(E::sub(a2, r));
            s[i0 + 2 * d] = E::st(E::add(a2, r));
        }
        __syncthreads();
    }
    if (t == 0) {
        const uint32_t half = 1u << (k + m - 1), sh = m;
        const uint32_t *tw_t = tws + (1u << (k - 1));
        for (uint32_t b = threadIdx.x; b < half; b += blockDim.x) {
            const uint32_t hp = b >> sh;
            const uint32_t i1 = (hp << (sh + 1)) | (b & ((1u << sh) - 1));
            butterfly<E>(s, i1, i1 + (1u << sh), tw_t[hp]);
        }
        __syncthreads();
    }
}

// Loads T <= E::kMaxPerThread x blockDim elements of a tile into LDS (index(L) maps tile
// slot L to the global element).  For T >= blockDim the trip count is uniform, so a thread's
// loads are all in flight (global, not flat) before the first LDS store.
template <class E, class F>
__device__ __forceinline__ void load_tile(typename E::S *s, const typename E::S *__restrict__ src, uint32_t T, F &&index)
{
    const uint32_t B = blockDim.x;
    constexpr uint32_t JM = E::kMaxPerThread;
    if (T >= B) {
        const uint32_t J = T / B;
        typename E::S v[JM];
#pragma unroll
        for (uint32_t j = 0; j < JM; ++j)
            if (j < J) v[j] = E::load_global(src + index(threadIdx.x + B * j));
#pragma unroll
        for (uint32_t j = 0; j < JM; ++j)
            if (j < J) s[threadIdx.x + B * j] = v[j];
    } else if (threadIdx.x < T) {
        s[threadIdx.x] = E::load_global(src + index(threadIdx.x));
    }
}

// Non-final pass (lo >= 1): tile index -> (high bits, column chunk).  LDS: T elements,
// then the tile's 2^k <= 2^10 twiddle slots.  T / 8 (QM31) or T / 16 (M31) threads per
// block, 256 to 1024.  For M31 the descriptor's src and dst are M31 word arrays.
template <class E>
__global__ __launch_bounds__(1024) void k_fft_pass(const FftPassDesc *__restrict__ tab, uint32_t ntab)
{
    extern __shared__ uint4 sraw[];
    typedef typename E::S S;
    S *s = reinterpret_cast<S *>(sraw);
    uint32_t tile;
    const FftPassDesc &d = find_desc(tab, ntab, tile);
    const uint32_t lo = d.lo, k = d.k, m = d.m;
    const uint32_t T = 1u << (k + m);
    const uint64_t n = 1ull << d.log_n;
    const uint32_t chunk_bits = lo - m;
    const uint64_t high = (uint64_t)tile >> chunk_bits;
    const uint64_t chunk = (uint64_t)tile & ((1ull << chunk_bits) - 1);
    const uint64_t base = (high << (lo + k)) | (chunk << m);
    const uint32_t cmask = (1u << m) - 1u;
    uint32_t *tws = reinterpret_cast<uint32_t *>(s + T);
    stage_twiddles(tws, d.tw, n, lo, k, high);
    auto at = [&](uint32_t L) { return base | ((uint64_t)(L >> m) << lo) | (L & cmask); };
    load_tile<E>(s, reinterpret_cast<const S *>(d.src), T, at);
    __syncthreads();
    fft_tile_stages<E>(s, tws, k, m);
    S *dst = reinterpret_cast<S *>(d.dst);
    for (uint32_t L = threadIdx.x; L < T; L += blockDim.x) E::store_global(dst + at(L), s[L]);
}

// Final pass: stages log_t - 1 .. 0 on a contiguous tile of T = 2^log_t elements,
// then the first-layer leaf hashes and the tile's Merkle subtree, as the reference
// fuses stage 0 with the leaf hash (circle_codeword_fft_stage_leaf_hash :1910-1943)
// and commits the hashes (SHA3PrehashedLeavesMerkleCommitPlan).  With run_fft == 0
// the tile is only hashed (a codeword supplied by the caller).  The twiddles sit in
// the tree's ping-pong region until the subtree is built.
// kTree: some entry of the launch hashes its tile.  The form without the hashing code
// needs a third of the registers (a codeword alone: twice the waves per SIMD).
template <bool kTree>
__global__ __launch_bounds__(256) void k_fft_final(const FftFinalDesc *__restrict__ tab, uint32_t ntab)
{
    extern __shared__ uint32_t lds[];
    uint32_t tile;
    const FftFinalDesc &d = find_desc(tab, ntab, tile);
    const uint32_t log_t = d.log_t;
    const uint32_t T = 1u << log_t;
    uint4 *s = reinterpret_cast<uint4 *>(lds);
    const uint64_t n = 1ull << d.log_n;
    const uint64_t base = (uint64_t)tile * T;
    uint32_t *tree = d.tree;
    if (d.run_fft) stage_twiddles(lds + 4 * T, d.tw, n, 0, log_t, tile);
    load_tile<QmElem>(s, d.src, T, [&](uint32_t L) { return base + L; });
    __syncthreads();
    if (d.run_fft) {
        fft_tile_stages<QmElem>(s, lds + 4 * T, log_t, 0);
        uint4 *dst = d.dst;
        for (uint32_t L = threadIdx.x; L < T; L += blockDim.x) st_g4(dst + base + L, s[L]);
        // every wave has read its elements before the subtree overwrites leaf pairs in place
        __syncthreads();
    }
    if constexpr (kTree) {
        if (tree) tile_tree_qm31(lds, T, base, tree, n);
    }
}

// The final pass over M31 words (a codeword of M31 coefficients; src holds M31 words).
// Without a tree the tile's M31 outputs go to dst as M31 words (LDS: T words, then T twiddle
// words).  With one, dst is the QM31 first layer: the tile is widened to (a, 0, 0, 0) in LDS
// (the M31 tile sits in the top quarter of the 4T-word leaf region of k_fft_final and every
// thread holds its elements in registers across the widening barrier), written out, and
// hashed into the tile's subtree exactly as k_fft_final does.  256 threads, T <= 2^12.
template <bool kTree>
__global__ __launch_bounds__(256) void k_fft_final_m31(const FftFinalDesc *__restrict__ tab, uint32_t ntab)
{
    extern __shared__ uint32_t lds[];
    uint32_t tile;
    const FftFinalDesc &d = find_desc(tab, ntab, tile);
    const uint32_t log_t = d.log_t, T = 1u << log_t;
    const uint64_t n = 1ull << d.log_n;
    const uint64_t base = (uint64_t)tile * T;
    uint32_t *s = kTree ? lds + 3 * T : lds;
    uint32_t *tws = kTree ? lds + 4 * T : lds + T;
    stage_twiddles(tws, d.tw, n, 0, log_t, tile);
    load_tile<M31Elem>(s, reinterpret_cast<const uint32_t *>(d.src), T, [&](uint32_t L) { return base + L; });
    __syncthreads();
    fft_tile_stages<M31Elem>(s, tws, log_t, 0);
    if constexpr (!kTree) {
        uint32_t *dst = reinterpret_cast<uint32_t *>(d.dst);
        for (uint32_t L = threadIdx.x; L < T; L += blockDim.x) glob(dst)[base + L] = s[L];
    } else {
        constexpr uint32_t JM = 16;
        const uint32_t B = blockDim.x;
        uint32_t v[JM];
#pragma unroll
        for (uint32_t j = 0; j < JM; ++j)
            if (threadIdx.x + B * j < T) v[j] = s[threadIdx.x + B * j];
        __syncthreads();  // every element is in a register before the widening overwrites them
        uint4 *leaves = reinterpret_cast<uint4 *>(lds);
#pragma unroll
        for (uint32_t j = 0; j < JM; ++j) {
            const uint32_t L = threadIdx.x + B * j;
            if (L < T) {
                const uint4 e = make_uint4(v[j], 0u, 0u, 0u);
                leaves[L] = e;
                st_g4(d.dst + base + L, e);
            }
        }
        __syncthreads();
        if (d.tree) tile_tree_qm31(lds, T, base, d.tree, n);
    }
}

// ============================================================== FRI fold

// Layer r (2 n_out elements) -> layer r + 1 (n_out elements) with the challenge
// read from device memory, then (if `tree`) the new layer's leaf hashes and tile
// subtrees.  Reference: qm31_fri_fold_challenge_buffer (:2013-2032) and
// QM31FRIFoldOracle.fold (QM31FRIFold.swift:11-44).
// 512 threads per block.  The fold to 2^19 outputs (256 tiles of 2048) runs one block per CU,
// and two waves per SIMD hide latency that one wave leaves exposed.  Measured on MI355X: 175 us
// against 193 us at 256 threads and 188 us at 1024; the fold to 2^20 outputs 302 / 307 / 368 us.
constexpr uint32_t kFoldThreads = 512;
__device__ __forceinline__ void fold_tile(const FoldDesc &d, uint32_t tile)
{
    extern __shared__ uint32_t lds[];
    if (d.t_ops) {  // block-uniform: a transcript block, no barrier below is reached
        if (threadIdx.x < 64) run_transcript(d.t_ops, d.t_state, d.t_nops);
        return;
    }
    const uint32_t T = 1u << d.log_t;
    const uint64_t base = (uint64_t)tile * T;
    const uint32_t *ch = d.ch;
    const qm31 c = {ch[0], ch[1], ch[2], ch[3]};
    const uint4 *in = d.in;
    uint4 *out = d.out;
    const uint32_t *tw = d.tw;
    uint32_t *tree = d.tree;
    uint4 *s = reinterpret_cast<uint4 *>(lds);
    for (uint32_t L = threadIdx.x; L < T; L += blockDim.x) {
        uint64_t g = base + L;
        qm31 o = fri_fold_pair(ld_qm(ld_g4(in + 2 * g)), ld_qm(ld_g4(in + 2 * g + 1)), glob(tw)[g], c);
        uint4 ov = st_qm(o);
        st_g4(out + g, ov);
        if (tree) s[L] = ov;
    }
    if (!tree) return;
    __syncthreads();
    tile_tree_qm31<2>(lds, T, base, tree, d.n_out);
}

__global__ AZK_FOUR_WAVES __launch_bounds__(kFoldThreads) void k_fold(const FoldDesc *__restrict__ tab, uint32_t ntab)
{
    uint32_t tile;
    const FoldDesc &d = find_desc(tab, ntab, tile);
    fold_tile(d, tile);
}

// A fold launch of one or two entries (a job's transcript block and its fold) takes them as
// kernel arguments, so the first loads wait on no table read.
__device__ __forceinline__ FoldDesc pick(const FoldPair &p, uint32_t &local)
{
    const FoldDesc d = (p.n > 1 && blockIdx.x >= p.d[1].first_block) ? p.d[1] : p.d[0];
    local = blockIdx.x - d.first_block;
    return d;
}

__global__ AZK_FOUR_WAVES __launch_bounds__(kFoldThreads) void k_fold_pair(const FoldPair p)
{
    uint32_t tile;
    const FoldDesc d = pick(p, tile);
    fold_tile(d, tile);
}

// Streaming form when no commitment follows (the fold chain plan): one output per
// thread over a full grid of 256-output blocks, measured fastest on MI355X (5.9 TB/s
// at 2^26 -> 2^25 against 4.5-5.1 TB/s for grid-stride loops).
__global__ __launch_bounds__(256) void k_fold_stream(const FoldDesc *__restrict__ tab, uint32_t ntab)
{
    uint32_t tile;
    const FoldDesc &d = find_desc(tab, ntab, tile);
    const uint64_t g = ((uint64_t)tile << 8) + threadIdx.x;
    if (g >= d.n_out) return;
    const uint32_t *ch = d.ch;
    const qm31 c = {ch[0], ch[1], ch[2], ch[3]};
    const uint4 l = d.in[2 * g], r = d.in[2 * g + 1];
    d.out[g] = st_qm(fri_fold_pair(ld_qm(l), ld_qm(r), d.tw[g], c));
}

// Latency form for small layers: one output element per thread, T <= 256 outputs
// per block; each thread folds its pair, hashes the new leaf and the block reduces
// its leaf hashes in LDS (serial chain: one leaf permutation + log2 T levels).
template <uint32_t kPairDiv>
__device__ __forceinline__ void fold_leafwise_tile(const FoldDesc &d, uint32_t tile)
{
    __shared__ uint32_t A[8 * 256];
    if (d.t_ops) {  // block-uniform: a transcript block, no barrier below is reached
        if (threadIdx.x < 64) run_transcript(d.t_ops, d.t_state, d.t_nops);
        return;
    }
    const uint32_t T = 1u << d.log_t;
    const uint64_t base = (uint64_t)tile * T;
    const uint32_t *ch = d.ch;
    const qm31 c = {ch[0], ch[1], ch[2], ch[3]};
    uint32_t *tree = d.tree;
    const uint64_t n_out = d.n_out;
    const uint32_t tid = threadIdx.x;
    if (tree && 2 * T <= blockDim.x) {
        // two lanes per leaf: both fold the pair, the even lane stores the value, and the
        // pair hashes it with the two-lanes-per-state permutation (lo: a, c; hi: b, d)
        const uint32_t L = tid >> 1;
        const bool hi = tid & 1u;
        if (L < T) {
            const uint64_t g = base + L;
            const qm31 o = fri_fold_pair(ld_qm(ld_g4(d.in + 2 * g)), ld_qm(ld_g4(d.in + 2 * g + 1)), glob(d.tw)[g], c);
            if (!hi) st_g4(d.out + g, st_qm(o));
            uint32_t a[25];
#pragma unroll
            for (int w = 0; w < 25; ++w) a[w] = 0;
            a[0] = hi ? o.b : o.a;
            a[1] = hi ? o.d : o.c;
            a[2] = hi ? 0u : 0x06u;
            a[16] = hi ? 0x80000000u : 0u;
            kf_h(a, hi);
            gu32 *node = node_ptr(tree, n_out, 0, g);
#pragma unroll
            for (int w = 0; w < 4; ++w) {
                node[2 * w + hi] = a[w];
                A[8 * L + 2 * w + hi] = a[w];
            }
        }
    } else if (tid < T) {
        const uint32_t L = tid;
        const uint64_t g = base + L;
        qm31 o = fri_fold_pair(ld_qm(ld_g4(d.in + 2 * g)), ld_qm(ld_g4(d.in + 2 * g + 1)), glob(d.tw)[g], c);
        st_g4(d.out + g, st_qm(o));
        if (tree) {
            uint32_t v[4] = {o.a, o.b, o.c, o.d}, h[8];
            sha3_leaf16(v, h);
            st_node(node_ptr(tree, n_out, 0, g), h);
            st_node(A + 8 * L, h);
        }
    }
    if (!tree) return;
    __syncthreads();
    reduce_lds<kPairDiv>(A, T, 0, base, tree, n_out);
}

// kPairDiv 2 when the launch holds two or more blocks per CU (2^21 commit: the
// folds to 2^17 and 2^18 outputs 71 and 107 us instead of 77 and 112), 1 with one block per
// CU, where the pair form's four waves take one SIMD each (to 2^16 outputs: 52 us against 55).
template <uint32_t kPairDiv>
__global__ __launch_bounds__(256) void k_fold_leafwise(const FoldDesc *__restrict__ tab, uint32_t ntab)
{
    uint32_t tile;
    const FoldDesc &d = find_desc(tab, ntab, tile);
    fold_leafwise_tile<kPairDiv>(d, tile);
}

template <uint32_t kPairDiv>
__global__ __launch_bounds__(256) void k_fold_leafwise_pair(const FoldPair p)
{
    uint32_t tile;
    const FoldDesc d = pick(p, tile);
    fold_leafwise_tile<kPairDiv>(d, tile);
}

// ============================================================== transcript

// Transcript sponge on the one-state-per-wave permutation: every lane holding state word w
// (its own lane and any copy) applies the same message word.  Absorbs `len` bytes at `src` (8-byte aligned, readable up to the next
// multiple of 8) as one SHA3-padded message.  Reference: TranscriptState.absorb
// (SHA3Oracle.swift:19-39), Metal transcript_absorb_keccak :1324-1366.
__device__ __forceinline__ void t_absorb(uint32_t &lo, uint32_t &hi, const WaveLane &c, const uint64_t *src,
                                         uint64_t len, uint32_t suffix)
{
    const bool in_rate = c.word < 17;
    uint64_t off = 0;
    while (off + 136 <= len) {
        if (in_rate) {
            const uint64_t w = glob(src)[off / 8 + c.word];
            lo ^= (uint32_t)w;
            hi ^= (uint32_t)(w >> 32);
        }
        kf_w(lo, hi, c);
        off += 136;
    }
    const uint32_t tc = (uint32_t)(len - off);
    if (in_rate) {
        const uint32_t b0 = 8u * c.word;
        uint64_t w = 0;
        if (b0 < tc) {
            w = glob(src)[off / 8 + c.word];
            if (tc - b0 < 8) w &= (1ull << (8 * (tc - b0))) - 1;
        }
        if (b0 <= tc && tc < b0 + 8) w ^= (uint64_t)suffix << (8 * (tc - b0));
        if (c.word == 16) w ^= 0x8000000000000000ull;
        lo ^= (uint32_t)w;
        hi ^= (uint32_t)(w >> 32);
    }
    kf_w(lo, hi, c);
}

// Rejection-sampled u32 words mod `modulus` from a copy of the state, 34 words per
// block.  Reference: TranscriptState.squeezeUInt32 (SHA3Oracle.swift:41-70).
__device__ __forceinline__ void t_squeeze(uint32_t lo, uint32_t hi, const WaveLane &c, uint32_t count,
                                          uint32_t modulus, uint32_t *dst)
{
    const uint64_t limit = (1ull << 32) - ((1ull << 32) % modulus);
    uint32_t got = 0, cand = 0;
    while (got < count) {
        if (cand > 0 && cand % 34 == 0) kf_w(lo, hi, c);
        const uint32_t wi = cand % 34, at = wave_lane_of(wi >> 1);
        const uint32_t wl = __builtin_amdgcn_readlane(lo, at), wh = __builtin_amdgcn_readlane(hi, at);
        const uint64_t w = (wi & 1) ? wh : wl;
        ++cand;
        if (w < limit) {
            if (threadIdx.x == 0) glob(dst)[got] = (uint32_t)(w % modulus);
            ++got;
        }
    }
}

// Runs a transcript program on its job's persistent 200-byte state with one whole wave
// (the one-state-per-wave layout: every lane holding a state word has it in lo, hi; the own
// lanes store it back).
__device__ __forceinline__ void run_transcript_from(const TranscriptOp *ops, uint64_t *state, uint32_t nops, uint32_t lo,
                                                    uint32_t hi)
{
    const WaveLane c = wave_lane();
    const bool holds = c.word < 25;
    for (uint32_t k = 0; k < nops; ++k) {
        const TranscriptOp op = ops[k];
        if (op.kind == TOP_RESET) {
            lo = hi = 0;
        } else if (op.kind == TOP_LOAD_STATE) {
            if (holds) {
                const uint64_t v = glob(reinterpret_cast<const uint64_t *>(op.src))[c.word];
                lo = (uint32_t)v;
                hi = (uint32_t)(v >> 32);
            }
        } else if (op.kind == TOP_ABSORB) {
            t_absorb(lo, hi, c, reinterpret_cast<const uint64_t *>(op.src), op.n, op.suffix);
        } else if (op.kind == TOP_SQUEEZE_U32) {
            t_squeeze(lo, hi, c, (uint32_t)op.n, op.modulus, reinterpret_cast<uint32_t *>(op.dst));
        }
    }
    if (c.own) glob(state)[c.word] = (uint64_t)lo | ((uint64_t)hi << 32);
}

// The same from the state in HBM.
__device__ __forceinline__ void run_transcript(const TranscriptOp *ops, uint64_t *state, uint32_t nops)
{
    const uint32_t w = wave_lane().word;
    uint32_t lo = 0, hi = 0;
    if (w < 25) {
        const uint64_t v = glob(state)[w];
        lo = (uint32_t)v;
        hi = (uint32_t)(v >> 32);
    }
    run_transcript_from(ops, state, nops, lo, hi);
}

__global__ __launch_bounds__(64) void k_transcript(const TranscriptDesc *__restrict__ tab, uint32_t ntab)
{
    uint32_t unused;
    const TranscriptDesc &d = find_desc(tab, ntab, unused);
    run_transcript(d.ops, d.state, d.nops);
}

// ============================================================== proof emission

// Per job: block 0 writes the commitments and the final layer into their slots of
// the proof words; block 1 + q * R + r writes layer r of query q (and the query head
// when r == 0): layer index, pair index, and both openings {leaf index, value,
// sibling count, siblings}.  Layout: CirclePCSFRIProofCodecV1.encode
// (CircleProofFormat.swift:475-535,617-628).
__global__ __launch_bounds__(256) void k_emit_proof(const EmitParams *__restrict__ tab, uint32_t ntab)
{
    uint32_t blk;
    const EmitParams &p = find_desc(tab, ntab, blk);
    uint32_t *w = p.proof;
    const uint32_t rounds = p.rounds;
    if (blk == 0) {
        for (uint32_t i = threadIdx.x; i < 8 * rounds; i += blockDim.x) w[p.roots_off + i] = p.roots[i / 8][i % 8];
        for (uint64_t i = threadIdx.x; i < 4 * p.final_count; i += blockDim.x) w[p.final_off + i] = p.final_layer[i];
        return;
    }
    const uint32_t q = (blk - 1) / rounds, r = (blk - 1) % rounds;
    // indices come from the transcript (mod N/2); the mask keeps every read in bounds
    const uint64_t pair0 = (uint64_t)p.query_pairs[q] & ((1ull << (p.log_n - 1)) - 1);
    const uint64_t pair = pair0 >> r;
    const uint32_t depth = p.log_n - r;  // siblings per opening
    uint32_t *qs = w + p.queries_off + (uint64_t)q * p.query_words;
    uint32_t *ls = qs + p.layer_off[r];
    const uint64_t m = 1ull << depth;
    const uint32_t *tree = p.trees[r];
    const uint4 *layer = p.layers[r];
    const uint32_t open_words = 7 + 8 * depth;
    if (threadIdx.x == 0) {
        if (r == 0) {
            qs[0] = (uint32_t)pair0; qs[1] = (uint32_t)(pair0 >> 32); qs[2] = rounds;
        }
        ls[0] = r; ls[1] = (uint32_t)pair; ls[2] = (uint32_t)(pair >> 32);
        for (int side = 0; side < 2; ++side) {
            uint32_t *o = ls + 3 + side * open_words;
            uint64_t leaf = 2 * pair + side;
            uint4 v = layer[leaf];
            o[0] = (uint32_t)leaf; o[1] = (uint32_t)(leaf >> 32);
            o[2] = v.x; o[3] = v.y; o[4] = v.z; o[5] = v.w;
            o[6] = depth;
        }
    }
    for (uint32_t i = threadIdx.x; i < 2 * 8 * depth; i += blockDim.x) {
        uint32_t side = i / (8 * depth), rem = i % (8 * depth), lv = rem / 8, wd = rem % 8;
        uint64_t leaf = 2 * pair + side;
        uint64_t sib = (leaf >> lv) ^ 1ull;
        ls[3 + side * open_words + 7 + rem] = tree[8 * ((2 * m - ((2 * m) >> lv)) + sib) + wd];
    }
}

// Sibling path of leaf `leaf` (bottom-up, 32 bytes per level) from a resident tree.
// Reference: MerkleOracle.openingSHA3_256 (MerkleOracle.swift:53-110).
__global__ __launch_bounds__(256) void k_merkle_path(const uint32_t *__restrict__ tree, uint64_t m, uint32_t depth,
                                                     uint64_t leaf, uint32_t *__restrict__ out)
{
    uint32_t t = threadIdx.x + blockIdx.x * blockDim.x;
    if (t >= 8 * depth + 8) return;
    if (t >= 8 * depth) {  // root
        out[t] = tree[8 * (2 * m - 2) + (t - 8 * depth)];
        return;
    }
    uint32_t lv = t / 8, wd = t % 8;
    out[t] = tree[8 * ((2 * m - ((2 * m) >> lv)) + ((leaf >> lv) ^ 1ull)) + wd];
}

// Sibling paths of `count` leaves in one launch: word t of the output is word t % (8 depth) of
// leaf t / (8 depth)'s path (the leaf indices are below m, checked by the host).
__global__ __launch_bounds__(256) void k_merkle_paths(const uint32_t *__restrict__ tree, uint64_t m, uint32_t depth,
                                                      const uint32_t *__restrict__ leaves, uint32_t count,
                                                      uint32_t *__restrict__ out)
{
    const uint32_t per = 8 * depth;
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (uint64_t)per * count) return;
    const uint32_t k = (uint32_t)(t / per), r = (uint32_t)(t % per), lv = r / 8, wd = r % 8;
    const uint64_t leaf = leaves[k];
    out[t] = tree[8 * ((2 * m - ((2 * m) >> lv)) + ((leaf >> lv) ^ 1ull)) + wd];
}

// ============================================================== host launchers

static inline uint32_t blocks_for(uint64_t n, uint32_t t) { return (uint32_t)((n + t - 1) / t); }
// a tile of T QM31 values (16T bytes); the final FFT pass also stages T twiddle words after it
static inline size_t tile_lds_bytes(uint32_t log_t) { return (size_t)16 << log_t; }
static inline size_t fft_tile_lds_bytes(uint32_t log_t) { return (size_t)20 << log_t; }  // 40 KB at 2^11

void launch_keccakf_batch(hipStream_t st, const void *in, uint32_t in_stride, void *out, uint32_t out_stride, uint32_t count)
{
    if (!count) return;
    hipLaunchKernelGGL(k_keccakf_batch, dim3(blocks_for(count, 256)), dim3(256), 0, st, (const uint8_t *)in, in_stride,
                       (uint8_t *)out, out_stride, count);
}

void launch_oneblock_hash(hipStream_t st, const void *in, uint32_t stride, uint32_t len, void *out, uint32_t out_stride,
                          uint32_t count, uint32_t suffix)
{
    if (!count) return;
    bool aligned = ((stride & 3u) == 0) && ((reinterpret_cast<uintptr_t>(in) & 3u) == 0);
    if (aligned)
        hipLaunchKernelGGL(k_oneblock_hash<true>, dim3(blocks_for(count, 256)), dim3(256), 0, st, (const uint8_t *)in,
                           stride, len, (uint8_t *)out, out_stride, count, suffix);
    else
        hipLaunchKernelGGL(k_oneblock_hash<false>, dim3(blocks_for(count, 256)), dim3(256), 0, st, (const uint8_t *)in,
                           stride, len, (uint8_t *)out, out_stride, count, suffix);
}

void launch_merkle_path(hipStream_t st, const uint32_t *tree, uint64_t m, uint32_t depth, uint64_t leaf, uint32_t *out)
{
    hipLaunchKernelGGL(k_merkle_path, dim3(blocks_for(8 * depth + 8, 256)), dim3(256), 0, st, tree, m, depth, leaf, out);
}

void launch_merkle_paths(hipStream_t st, const uint32_t *tree, uint64_t m, uint32_t depth, const uint32_t *leaves,
                         uint32_t count, uint32_t *out)
{
    const uint64_t words = (uint64_t)8 * depth * count;
    if (!words) return;
    hipLaunchKernelGGL(k_merkle_paths, dim3(blocks_for(words, 256)), dim3(256), 0, st, tree, m, depth, leaves, count, out);
}

void launch_fft_twiddles(hipStream_t st, uint32_t *out, uint32_t log_n)
{
    uint64_t n = (1ull << log_n) - 1;
    if (!n) return;
    hipLaunchKernelGGL(k_fft_twiddles, dim3(blocks_for(n, 256)), dim3(256), 0, st, out, log_n);
}

void launch_inverse(hipStream_t st, const uint32_t *in, uint32_t *out, uint64_t n)
{
    if (!n) return;
    hipLaunchKernelGGL(k_inverse, dim3(blocks_for(n, 256)), dim3(256), 0, st, in, out, n);
}

void launch_merkle_raw(hipStream_t st, const MerkleRawDesc *t, uint32_t n, uint32_t blocks, bool aligned, uint32_t max_log_t,
                       const MerkleRawDesc *single)
{
    if (!blocks) return;
    const size_t lds = tile_lds_bytes(max_log_t);
    if (single && n == 1) {
        if (aligned) hipLaunchKernelGGL(k_merkle_raw_tiles1<true>, dim3(blocks), dim3(256), lds, st, *single);
        else hipLaunchKernelGGL(k_merkle_raw_tiles1<false>, dim3(blocks), dim3(256), lds, st, *single);
    } else if (aligned) {
        hipLaunchKernelGGL(k_merkle_raw_tiles<true>, dim3(blocks), dim3(256), lds, st, t, n);
    } else {
        hipLaunchKernelGGL(k_merkle_raw_tiles<false>, dim3(blocks), dim3(256), lds, st, t, n);
    }
}

void launch_fft_pass(hipStream_t st, const FftPassDesc *t, uint32_t n, uint32_t blocks, uint32_t max_log_t, bool m31)
{
    if (!blocks) return;
    // T elements + 2^k <= 2^10 twiddle words (passes have at most 10 stages)
    if (m31) {
        const uint32_t threads = std::min(1024u, std::max(256u, (1u << max_log_t) / 16));
        hipLaunchKernelGGL(k_fft_pass<M31Elem>, dim3(blocks), dim3(threads), ((size_t)4 << max_log_t) + 4096, st, t, n);
    } else {
        const uint32_t threads = std::min(1024u, std::max(256u, (1u << max_log_t) / 8));
        hipLaunchKernelGGL(k_fft_pass<QmElem>, dim3(blocks), dim3(threads), ((size_t)16 << max_log_t) + 4096, st, t, n);
    }
}

void launch_fft_final(hipStream_t st, const FftFinalDesc *t, uint32_t n, uint32_t blocks, uint32_t max_log_t, bool tree,
                      bool m31)
{
    if (!blocks) return;
    const size_t lds = fft_tile_lds_bytes(max_log_t);
    if (m31 && tree) hipLaunchKernelGGL(k_fft_final_m31<true>, dim3(blocks), dim3(256), lds, st, t, n);
    else if (m31) hipLaunchKernelGGL(k_fft_final_m31<false>, dim3(blocks), dim3(256), (size_t)8 << max_log_t, st, t, n);
    else if (tree) hipLaunchKernelGGL(k_fft_final<true>, dim3(blocks), dim3(256), lds, st, t, n);
    else hipLaunchKernelGGL(k_fft_final<false>, dim3(blocks), dim3(256), lds, st, t, n);
}

static bool fold_pair(const FoldDesc *host, uint32_t n, FoldPair &p)
{
    if (!host || n == 0 || n > 2) return false;
    p = FoldPair{};
    p.d[0] = host[0];
    if (n > 1) p.d[1] = host[1];
    p.n = n;
    return true;
}

void launch_fold(hipStream_t st, const FoldDesc *t, uint32_t n, uint32_t blocks, uint32_t max_log_t, const FoldDesc *host)
{
    if (!blocks) return;
    FoldPair p;
    if (fold_pair(host, n, p))
        hipLaunchKernelGGL(k_fold_pair, dim3(blocks), dim3(kFoldThreads), tile_lds_bytes(max_log_t), st, p);
    else
        hipLaunchKernelGGL(k_fold, dim3(blocks), dim3(kFoldThreads), tile_lds_bytes(max_log_t), st, t, n);
}

void launch_fold_stream(hipStream_t st, const FoldDesc *t, uint32_t n, uint32_t blocks)
{
    if (!blocks) return;
    hipLaunchKernelGGL(k_fold_stream, dim3(blocks), dim3(256), 0, st, t, n);
}

void launch_fold_leafwise(hipStream_t st, const FoldDesc *t, uint32_t n, uint32_t blocks, uint32_t cus, const FoldDesc *host)
{
    if (!blocks) return;
    FoldPair p;
    const bool many = blocks >= 2 * cus;  // a transcript block beside one tile per CU does not count
    if (fold_pair(host, n, p)) {
        if (many) hipLaunchKernelGGL(k_fold_leafwise_pair<2>, dim3(blocks), dim3(256), 0, st, p);
        else hipLaunchKernelGGL(k_fold_leafwise_pair<1>, dim3(blocks), dim3(256), 0, st, p);
    } else {
        if (many) hipLaunchKernelGGL(k_fold_leafwise<2>, dim3(blocks), dim3(256), 0, st, t, n);
        else hipLaunchKernelGGL(k_fold_leafwise<1>, dim3(blocks), dim3(256), 0, st, t, n);
    }
}

void launch_tree(hipStream_t st, const TreeDesc *t, uint32_t n, uint32_t blocks, uint32_t max_log_c, uint32_t waves,
                 const TreeDesc *single)
{
    if (!blocks) return;
    const size_t lds = (size_t)32 << max_log_c;  // 2^log_c nodes of 32 bytes, reduced in place
    // 4 waves (one per SIMD): 8-wave blocks with 16-node chunks measured no faster
    if (single && n == 1)
        hipLaunchKernelGGL(k_tree_reduce1, dim3(blocks), dim3(64 * waves), lds, st, *single);
    else
        hipLaunchKernelGGL(k_tree_reduce, dim3(blocks), dim3(64 * waves), lds, st, t, n);
}

void launch_grind(hipStream_t st, const uint64_t *state, uint32_t bits, uint64_t base, uint32_t count, unsigned long long *best)
{
    hipLaunchKernelGGL(k_grind, dim3(count / 256), dim3(256), 0, st, state, bits, base, best);
}

void launch_transcript(hipStream_t st, const TranscriptDesc *t, uint32_t n)
{
    if (!n) return;
    hipLaunchKernelGGL(k_transcript, dim3(n), dim3(64), 0, st, t, n);
}

void launch_copy(hipStream_t st, const CopyDesc *t, uint32_t n, uint32_t blocks)
{
    if (!blocks) return;
    hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, st, t, n);
}

void launch_emit_proof(hipStream_t st, const EmitParams *t, uint32_t n, uint32_t blocks)
{
    if (!blocks) return;
    hipLaunchKernelGGL(k_emit_proof, dim3(blocks), dim3(256), 0, st, t, n);
}

} // namespace azk
