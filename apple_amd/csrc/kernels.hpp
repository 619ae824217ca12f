// Host-visible kernel parameter blocks and launchers (implemented in kernels.hip).
//
// Every tiled kernel is driven by a table of descriptors in device memory: one
// descriptor per (job, phase) with the index of its first block; a block finds its
// descriptor by binary search over `first_block`.  A single job is a table of one
// entry; a batch of uneven jobs is one launch over all of their tiles.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace azk {

constexpr int kMaxRounds = 30;
constexpr uint32_t kMaxGrindingBits = 20;  // CircleFRIGrindingV1.maximumLocalSearchBits

enum TranscriptOpKind : uint32_t { TOP_RESET = 0, TOP_ABSORB = 1, TOP_SQUEEZE_U32 = 2, TOP_LOAD_STATE = 3 };

// One step of a device transcript program.  `src` must be 8-byte aligned and
// readable up to the next multiple of 8 bytes past `n`.
struct TranscriptOp {
    uint32_t kind;
    uint32_t suffix;   // domain suffix for ABSORB (0x06 = SHA3 padding)
    uint32_t modulus;  // SQUEEZE_U32
    uint32_t pad;
    uint64_t n;        // bytes (ABSORB) or words (SQUEEZE_U32)
    const void *src;
    void *dst;
};

// Stages lo + k - 1 .. lo of one job's Circle FFT, tiles of 2^(k + m) elements.
struct FftPassDesc {
    const uint4 *src;
    uint4 *dst;
    const uint32_t *tw;
    uint32_t log_n, lo, k, m;
    uint32_t first_block, pad;
};

// Final FFT stages (run_fft) on contiguous tiles of 2^log_t elements, then leaf
// hashes and tile subtrees into `tree` (if set).
struct FftFinalDesc {
    const uint4 *src;
    uint4 *dst;
    const uint32_t *tw;
    uint32_t *tree;
    uint32_t log_n, log_t, run_fft, first_block;
};

// One FRI fold round of one job (2 n_out -> n_out), fused with the new layer's
// leaf hashes and tile subtrees (if `tree`).  A descriptor with `t_ops` set is instead
// one block that runs that transcript program on its first wave, beside the fold.
struct FoldDesc {
    const uint4 *in;
    uint4 *out;
    const uint32_t *tw;
    const uint32_t *ch;
    uint32_t *tree;
    uint64_t n_out;
    uint32_t log_t, first_block;
    const struct TranscriptOp *t_ops;
    uint64_t *t_state;
    uint32_t t_nops, pad;
};

// One or two fold entries passed by value (n of them).
struct FoldPair {
    FoldDesc d[2];
    uint32_t n, pad;
};

// Reduction of chunks of 2^log_c nodes of `level` up by log_c levels.
struct TreeDesc {
    uint32_t *tree;
    uint64_t m;  // leaves of the tree
    // transcript program run by the block that computes the root (final step only)
    const struct TranscriptOp *t_ops;
    uint64_t *t_state;
    uint32_t level, log_c, first_block, t_nops;
};

// A transcript program run by one wave.
struct TranscriptDesc {
    const TranscriptOp *ops;
    uint64_t *state;
    uint32_t nops, first_block;
};

// Word copy of `words` words, 1024 words per block.
struct CopyDesc {
    const uint32_t *src;
    uint32_t *dst;
    uint64_t words;
    uint32_t first_block, pad;
};

// Raw-leaf Merkle tiles.
struct MerkleRawDesc {
    const uint8_t *leaves;
    uint32_t *tree;
    uint64_t m;
    uint32_t stride, len, log_t, first_block;
};

struct EmitParams {
    uint32_t *proof;                 // proof as LE u32 words
    uint64_t roots_off, final_off, queries_off, query_words;
    uint32_t layer_off[kMaxRounds];  // word offset of layer r inside a query section
    uint32_t rounds, log_n;
    uint64_t final_count;
    const uint32_t *roots[kMaxRounds];
    const uint32_t *final_layer;
    const uint32_t *trees[kMaxRounds];
    const uint4 *layers[kMaxRounds];
    const uint32_t *query_pairs;
    uint32_t queries, first_block;   // blocks: 1 + queries * rounds
};

void launch_keccakf_batch(hipStream_t st, const void *in, uint32_t in_stride, void *out, uint32_t out_stride, uint32_t count);
void launch_oneblock_hash(hipStream_t st, const void *in, uint32_t stride, uint32_t len, void *out, uint32_t out_stride,
                          uint32_t count, uint32_t suffix);
void launch_merkle_path(hipStream_t st, const uint32_t *tree, uint64_t m, uint32_t depth, uint64_t leaf, uint32_t *out);
void launch_merkle_paths(hipStream_t st, const uint32_t *tree, uint64_t m, uint32_t depth, const uint32_t *leaves,
                         uint32_t count, uint32_t *out);
void launch_fft_twiddles(hipStream_t st, uint32_t *out, uint32_t log_n);
void launch_inverse(hipStream_t st, const uint32_t *in, uint32_t *out, uint64_t n);

// Table-driven launches: `blocks` = total blocks over all descriptors.
// `single`: host copy of the table's only entry (n == 1), passed by value (`t` may then be null).
void launch_merkle_raw(hipStream_t st, const MerkleRawDesc *t, uint32_t n, uint32_t blocks, bool aligned, uint32_t max_log_t,
                       const MerkleRawDesc *single = nullptr);
void launch_fft_pass(hipStream_t st, const FftPassDesc *t, uint32_t n, uint32_t blocks, uint32_t max_log_t, bool m31);
// tree: some entry hashes its tile (FftFinalDesc::tree set).  m31: the entries transform M31
// words (src and pass buffers hold N M31 words; dst is the QM31 first layer with a tree,
// N M31 words without), tiles of at most 2^12 elements.
void launch_fft_final(hipStream_t st, const FftFinalDesc *t, uint32_t n, uint32_t blocks, uint32_t max_log_t, bool tree,
                      bool m31);
// `host`: host copy of the table; a launch of at most two entries passes them by value.
void launch_fold(hipStream_t st, const FoldDesc *t, uint32_t n, uint32_t blocks, uint32_t max_log_t,
                 const FoldDesc *host = nullptr);
// cus: compute units of the device (the launch's blocks per CU set the Merkle levels' form)
void launch_fold_leafwise(hipStream_t st, const FoldDesc *t, uint32_t n, uint32_t blocks, uint32_t cus,
                          const FoldDesc *host = nullptr);
// No commitment: one output per thread, 256 outputs per block (log_t ignored).
void launch_fold_stream(hipStream_t st, const FoldDesc *t, uint32_t n, uint32_t blocks);
// waves per block: 4, 8 or 16
// `single`: host copy of the table's only entry (n == 1), passed by value to the kernel.
void launch_tree(hipStream_t st, const TreeDesc *t, uint32_t n, uint32_t blocks, uint32_t max_log_c, uint32_t waves = 4,
                 const TreeDesc *single = nullptr);
void launch_transcript(hipStream_t st, const TranscriptDesc *t, uint32_t n);
// Nonces base .. base + count - 1 (count a multiple of 256) against the grinding state.
void launch_grind(hipStream_t st, const uint64_t *state, uint32_t bits, uint64_t base, uint32_t count,
                  unsigned long long *best);
void launch_copy(hipStream_t st, const CopyDesc *t, uint32_t n, uint32_t blocks);
void launch_emit_proof(hipStream_t st, const EmitParams *t, uint32_t n, uint32_t blocks);

} // namespace azk
