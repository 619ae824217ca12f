// Host-visible kernel parameter blocks and launchers (implemented in kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace azk {

constexpr int kMaxRounds = 30;

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
};

void launch_keccakf_batch(hipStream_t st, const void *in, uint32_t in_stride, void *out, uint32_t out_stride, uint32_t count);
void launch_oneblock_hash(hipStream_t st, const void *in, uint32_t stride, uint32_t len, void *out, uint32_t out_stride,
                          uint32_t count, uint32_t suffix);
void launch_merkle_raw(hipStream_t st, const void *leaves, uint32_t stride, uint32_t len, uint64_t m, uint32_t log_t,
                       uint32_t *tree);
void launch_tree_upper(hipStream_t st, uint32_t *tree, uint64_t m, uint32_t level, uint32_t max_log_c);
void launch_copy_words(hipStream_t st, const uint32_t *src, uint32_t *dst, uint64_t n);
void launch_fft_twiddles(hipStream_t st, uint32_t *out, uint32_t log_n);
void launch_inverse(hipStream_t st, const uint32_t *in, uint32_t *out, uint64_t n);
void launch_fft_pass(hipStream_t st, const void *src, void *dst, const uint32_t *tw, uint32_t log_n, uint32_t lo,
                     uint32_t k, uint32_t m);
void launch_fft_final(hipStream_t st, const void *src, void *dst, const uint32_t *tw, uint32_t log_n, uint32_t log_t,
                      bool run_fft, uint32_t *tree);
void launch_fold(hipStream_t st, const void *in, void *out, const uint32_t *tw, const uint32_t *ch, uint64_t n_out,
                 uint32_t log_t, uint32_t *tree);
void launch_merkle_path(hipStream_t st, const uint32_t *tree, uint64_t m, uint32_t depth, uint64_t leaf, uint32_t *out);
void launch_fold_leafwise(hipStream_t st, const void *in, void *out, const uint32_t *tw, const uint32_t *ch,
                         uint64_t n_out, uint32_t log_t, uint32_t *tree);
void launch_transcript(hipStream_t st, const TranscriptOp *ops, uint32_t nops, uint64_t *state);
void launch_emit_proof(hipStream_t st, const EmitParams &p, uint32_t queries);

} // namespace azk
