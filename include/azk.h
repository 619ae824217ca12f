/*
 * azk.h -- C ABI of the MI355X (gfx950) Circle-PCS commit engine.
 *
 * Drop-in boundary for the reference's reusable-plan API (AppleZKProver, Swift over
 * Metal).  Every entry point names the reference interface it replaces
 * (paths relative to the reference repository, Sources/AppleZKProver/).
 * Plain pointers and sizes only; no exceptions cross the ABI; every call returns an
 * azk_status and leaves a thread-local message for azk_last_error().
 *
 * Threading (MerkleCommitter.swift:548,723 executionLock): concurrent calls on one
 * plan are serialized by a per-plan mutex; separate plans may be used concurrently.
 * Calls are synchronous unless named *_async.
 */
#ifndef AZK_H
#define AZK_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* AppleZKProverError, MetalTypes.swift:3-23 (one code per case). */
typedef enum azk_status {
    AZK_OK = 0,
    AZK_ERR_NO_DEVICE = 1,                      /* noMetalDevice */
    AZK_ERR_ALLOCATION = 2,                     /* failedToCreateBuffer / CommandQueue */
    AZK_ERR_UNSUPPORTED_ONE_BLOCK_LENGTH = 3,   /* unsupportedOneBlockLength(n) */
    AZK_ERR_INVALID_LEAF_COUNT = 4,             /* invalidLeafCount(n) */
    AZK_ERR_INVALID_PREHASHED_LEAF_LENGTH = 5,  /* invalidPrehashedLeafLength(n) */
    AZK_ERR_INVALID_LAYOUT = 6,                 /* invalidInputLayout */
    AZK_ERR_CORRECTNESS = 7,                    /* correctnessValidationFailed(msg) */
    AZK_ERR_COMMAND_FAILED = 8,                 /* commandExecutionFailed(msg) */
    AZK_ERR_UNAVAILABLE = 9,                    /* unavailableOnThisPlatform */
    AZK_ERR_INVALID_KERNEL_CONFIG = 10,         /* invalidKernelConfiguration(msg) */
    AZK_ERR_INVALID_ARGUMENT = 11               /* null handle or pointer (no Swift analogue) */
} azk_status;

/* Message for the last failing call on this thread ("" if none). */
const char *azk_last_error(void);
const char *azk_version(void);

typedef struct azk_context azk_context;
typedef struct azk_buffer azk_buffer;

/* GPUCapabilities, MetalContext.swift:11-44. */
typedef struct azk_capabilities {
    char name[64];
    char arch[32];
    int compute_units;
    int wavefront_size;
    size_t lds_bytes_per_workgroup;
    size_t hbm_bytes;
    int clock_khz;
} azk_capabilities;

/* GPUExecutionStats, MetalTypes.swift:71-79 (+ submit time, MerkleCommitter.swift:782-797). */
typedef struct azk_stats {
    uint64_t wall_ns;    /* host wall clock of the call */
    uint64_t gpu_ns;     /* HIP-event time of the device work */
    uint64_t submit_ns;  /* host time spent enqueueing */
} azk_stats;

/* ---- context: MetalContext(device:pipelineCacheConfiguration:), MetalContext.swift:76-112 */
azk_status azk_context_create(int device_ordinal, azk_context **out);
void azk_context_destroy(azk_context *ctx);
azk_status azk_context_capabilities(azk_context *ctx, azk_capabilities *caps);
azk_status azk_context_synchronize(azk_context *ctx);

/* ---- resident buffers: MTLBuffer + offset pairs (MetalBufferFactory.swift:1-159) */
azk_status azk_buffer_alloc(azk_context *ctx, size_t bytes, azk_buffer **out);
/* Wraps caller-owned device memory (not freed by azk_buffer_free). */
azk_status azk_buffer_wrap(azk_context *ctx, void *device_ptr, size_t bytes, azk_buffer **out);
void azk_buffer_free(azk_buffer *buf);
size_t azk_buffer_size(const azk_buffer *buf);
void *azk_buffer_device_ptr(const azk_buffer *buf);
azk_status azk_buffer_upload(azk_buffer *buf, size_t offset, const void *host, size_t bytes);
azk_status azk_buffer_download(const azk_buffer *buf, size_t offset, void *host, size_t bytes);

/* ---- Keccak-f[1600] permutation batch
 * KeccakF1600PermutationBatcher.makePermutationPlan / KeccakF1600PermutationBatchPlan
 * (KeccakF1600PermutationBatcher.swift:31-42,73-294); descriptor MetalTypes.swift:172-188. */
typedef struct azk_keccakf_desc {
    uint32_t count;
    uint32_t input_stride;   /* bytes, multiple of 8, >= 200 */
    uint32_t output_stride;  /* bytes, multiple of 8, >= 200; padding is zeroed */
} azk_keccakf_desc;
typedef struct azk_keccakf_plan azk_keccakf_plan;
azk_status azk_keccakf_plan_create(azk_context *ctx, const azk_keccakf_desc *desc, azk_keccakf_plan **out);
void azk_keccakf_plan_destroy(azk_keccakf_plan *plan);
/* permute(states:) :157-197 -- host arrays in and out. */
azk_status azk_keccakf_permute(azk_keccakf_plan *plan, const void *states, size_t in_bytes, void *out, size_t out_bytes,
                               azk_stats *stats);
/* resident variant: inputs already in HBM, nothing crosses PCIe. */
azk_status azk_keccakf_permute_resident(azk_keccakf_plan *plan, const azk_buffer *in, size_t in_offset, azk_buffer *out,
                                        size_t out_offset, azk_stats *stats);
/* permuteVerified :199-203 -- CPU recomputation gate, AZK_ERR_CORRECTNESS on mismatch. */
azk_status azk_keccakf_permute_verified(azk_keccakf_plan *plan, const void *states, size_t in_bytes, void *out,
                                        size_t out_bytes, azk_stats *stats);
azk_status azk_keccakf_plan_clear(azk_keccakf_plan *plan);

/* ---- one-block SHA3-256 / Keccak-256 batch hashing
 * SHA3FixedOneBlockHashPlan (SHA3BatchHasher.swift:81-264), Keccak256FixedOneBlockHashPlan
 * (Keccak256BatchHasher.swift:73-257). */
typedef struct azk_oneblock_desc {
    uint32_t count;
    uint32_t input_stride;   /* bytes >= input_length */
    uint32_t input_length;   /* 0..136 */
    uint32_t output_stride;  /* bytes >= 32; padding is zeroed */
} azk_oneblock_desc;
typedef struct azk_hash_plan azk_hash_plan;
azk_status azk_sha3_256_plan_create(azk_context *ctx, const azk_oneblock_desc *desc, azk_hash_plan **out);
azk_status azk_keccak_256_plan_create(azk_context *ctx, const azk_oneblock_desc *desc, azk_hash_plan **out);
void azk_hash_plan_destroy(azk_hash_plan *plan);
azk_status azk_hash_execute(azk_hash_plan *plan, const void *messages, size_t in_bytes, void *digests, size_t out_bytes,
                            azk_stats *stats);
azk_status azk_hash_execute_resident(azk_hash_plan *plan, const azk_buffer *in, size_t in_offset, azk_buffer *out,
                                     size_t out_offset, azk_stats *stats);
azk_status azk_hash_execute_verified(azk_hash_plan *plan, const void *messages, size_t in_bytes, void *digests,
                                     size_t out_bytes, azk_stats *stats);
azk_status azk_hash_plan_clear(azk_hash_plan *plan);

/* ---- SHA3-256 Merkle commitment over raw leaves (<= 136 bytes)
 * SHA3MerkleCommitter.makeRawLeavesCommitPlan / SHA3RawLeavesMerkleCommitPlan
 * (MerkleCommitter.swift:152-165,526-1182).  The whole tree stays resident after a
 * commit; openings gather siblings from it (openRawLeaf :896-1063 rebuilt it). */
typedef struct azk_merkle_desc {
    uint32_t leaf_count;     /* power of two */
    uint32_t leaf_stride;
    uint32_t leaf_length;    /* 0..136 */
    uint32_t subtree_log;    /* leaves per LDS treelet = 2^subtree_log; 0 = default: clamp(log2 count - 8, 8, 11) */
} azk_merkle_desc;
typedef struct azk_merkle_plan azk_merkle_plan;
azk_status azk_merkle_plan_create(azk_context *ctx, const azk_merkle_desc *desc, azk_merkle_plan **out);
void azk_merkle_plan_destroy(azk_merkle_plan *plan);
azk_status azk_merkle_commit(azk_merkle_plan *plan, const void *leaves, size_t bytes, uint8_t root[32], azk_stats *stats);
azk_status azk_merkle_commit_resident(azk_merkle_plan *plan, const azk_buffer *leaves, size_t offset, uint8_t root[32],
                                      azk_stats *stats);
azk_status azk_merkle_commit_verified(azk_merkle_plan *plan, const void *leaves, size_t bytes, uint8_t root[32],
                                      azk_stats *stats);
/* Opening of leaf `index` from the last commit: siblings bottom-up, log2(leaf_count) x 32 bytes.
 * MerkleOracle.openingSHA3_256 (MerkleOracle.swift:53-110). */
azk_status azk_merkle_open(azk_merkle_plan *plan, uint32_t index, uint8_t *siblings, size_t siblings_bytes,
                           uint8_t root[32]);
/* Openings of `count` leaves in one call (one gather launch, one copy back): path k occupies
 * siblings[k * 32 log2(leaf_count) ...], each as azk_merkle_open writes it; root as there.  No
 * single counterpart in the reference, whose openRawLeaf (:195-226) opens one leaf per call;
 * any index >= leaf_count rejects the whole call (AZK_ERR_INVALID_LAYOUT). */
azk_status azk_merkle_open_batch(azk_merkle_plan *plan, const uint32_t *indices, size_t count, uint8_t *siblings,
                                 size_t siblings_bytes, uint8_t root[32]);
azk_status azk_merkle_plan_clear(azk_merkle_plan *plan);
/* No counterpart in the reference: an allocation-free variant of azk_merkle_commit_resident
 * for callers that cannot allocate device memory (code under hipGraph capture, serving loops
 * that forbid hipMalloc).  Root of leaf_count = 2^k raw leaves (leaf_count <= 65536, leaf
 * length <= 136 bytes, stride >= length) at (leaves, leaves_offset), written as 32 bytes to
 * (root, root_offset).  The tree lives in a 4 MiB library-owned array of each GPU, held until
 * the process ends; nothing is allocated, copied from the host or waited for: the work is
 * enqueued on `stream` (a hipStream_t; NULL = the context's stream).  Calls that overlap on
 * one GPU race on that array.  Larger leaf counts give AZK_ERR_INVALID_LEAF_COUNT. */
azk_status azk_merkle_commit_noalloc(azk_context *ctx, const azk_buffer *leaves, size_t leaves_offset,
                                     uint32_t leaf_count, uint32_t leaf_stride, uint32_t leaf_length, azk_buffer *root,
                                     size_t root_offset, void *stream);

/* ---- Merkle root over prehashed 32-byte leaves
 * SHA3PrehashedLeavesMerkleCommitPlan(context:leafCount:) (MerkleCommitter.swift:314-524). */
typedef struct azk_prehashed_plan azk_prehashed_plan;
azk_status azk_prehashed_plan_create(azk_context *ctx, uint32_t leaf_count, azk_prehashed_plan **out);
void azk_prehashed_plan_destroy(azk_prehashed_plan *plan);
azk_status azk_prehashed_commit_resident(azk_prehashed_plan *plan, const azk_buffer *leaf_hashes, size_t offset,
                                         uint8_t root[32], azk_stats *stats);

/* ---- Circle codeword (Circle FFT) on the canonical bit-reversed domain
 * CircleCodewordPlan(context:domain:) (CircleCodeword.swift:411-887).  Input: interleaved
 * line-basis coefficients (coeff[2i] = x_i, coeff[2i+1] = y_i; CircleCodeword.swift:100-127),
 * N QM31 = N x 16 bytes.  Output: N QM31 evaluations in storage order. */
typedef struct azk_codeword_plan azk_codeword_plan;
azk_status azk_codeword_plan_create(azk_context *ctx, uint32_t log_n, azk_codeword_plan **out);
void azk_codeword_plan_destroy(azk_codeword_plan *plan);
azk_status azk_codeword_execute_resident(azk_codeword_plan *plan, const azk_buffer *coeffs, size_t coeff_offset,
                                         azk_buffer *out, size_t out_offset, azk_stats *stats);
azk_status azk_codeword_execute(azk_codeword_plan *plan, const void *coeffs, size_t in_bytes, void *evals,
                                size_t out_bytes, azk_stats *stats);
/* executeVerified (:476-490): AZK_ERR_CORRECTNESS unless the evaluations equal the host
 * Circle FFT of the coefficients (azk_host_circle_fft). */
azk_status azk_codeword_execute_verified(azk_codeword_plan *plan, const void *coeffs, size_t in_bytes, void *evals,
                                         size_t out_bytes, azk_stats *stats);
/* The codeword of M31 coefficients: N M31 words in (coefficient i's QM31 lift is (c_i, 0, 0, 0)),
 * N M31 words out (the first components of the QM31 codeword; the other three are zero).  No
 * counterpart in the reference, which lifts M31 coefficients to QM31 and runs the QM31 FFT;
 * for callers whose coefficients are M31 (a quarter of the arithmetic and bytes).  Buffers
 * 4-byte aligned, not aliased. */
azk_status azk_codeword_execute_m31_resident(azk_codeword_plan *plan, const azk_buffer *coeffs, size_t coeff_offset,
                                             azk_buffer *out, size_t out_offset, azk_stats *stats);
azk_status azk_codeword_execute_m31(azk_codeword_plan *plan, const uint32_t *coeffs, size_t in_words, uint32_t *evals,
                                    size_t out_words, azk_stats *stats);
/* FFT twiddles (N - 1 words) as materialized on the device (readCodewordTwiddles :818-820). */
azk_status azk_codeword_read_twiddles(azk_codeword_plan *plan, uint32_t *out, size_t words);

/* ---- Monomial witness columns -> Circle-FFT-basis coefficients
 * CircleWitnessToFFTBasisPlanV1(context:domain:) (CircleWitnessFFTBasis.swift:230-517) and
 * CircleCodewordOracle.circleFFTCoefficients (CircleCodeword.swift:100-127): x / y monomial
 * coefficients (QM31, counts <= N/2 each, not both zero, canonical) -> N interleaved
 * line-basis coefficients (out[2i] = x-line i, out[2i+1] = y-line i), the codeword plan's
 * input.  O(N log^2 N) on the device; the reference's dense matrix stops at N/2 = 4096.
 * Non-canonical input limbs give AZK_ERR_INVALID_LAYOUT (validateResident). */
typedef struct azk_basis_plan azk_basis_plan;
azk_status azk_witness_basis_plan_create(azk_context *ctx, uint32_t log_n, azk_basis_plan **out);
void azk_witness_basis_plan_destroy(azk_basis_plan *plan);
/* executeResident(xWitnessCoefficientBuffer:...:outputCircleCoefficientOffset:) :320-428 */
azk_status azk_witness_basis_execute_resident(azk_basis_plan *plan, const azk_buffer *x, size_t x_offset,
                                              uint32_t x_count, const azk_buffer *y, size_t y_offset,
                                              uint32_t y_count, azk_buffer *out, size_t out_offset, azk_stats *stats);
/* host arrays: x_count / y_count QM31 (4 LE u32 each); out: 4N words */
azk_status azk_witness_basis_execute(azk_basis_plan *plan, const uint32_t *x, uint32_t x_count, const uint32_t *y,
                                     uint32_t y_count, uint32_t *out, size_t out_words, azk_stats *stats);
/* executeVerified (:483-517): AZK_ERR_CORRECTNESS unless azk_witness_basis_check accepts the
 * result (the reference recomputes the basis with its O(n^2) oracle). */
azk_status azk_witness_basis_execute_verified(azk_basis_plan *plan, const uint32_t *x, uint32_t x_count,
                                              const uint32_t *y, uint32_t y_count, uint32_t *out, size_t out_words,
                                              azk_stats *stats);

/* ---- Circle domain materialization
 * CircleDomainMaterializationPlan(context:domain:materializeDomainPoints:materializeCodewordTwiddles:
 * inverseDomainRoundCount:) (CircleDomainMaterializationPlan.swift:54-133): the canonical 2^log_n
 * domain (bit-reversed storage) materialized on the device at creation: points (N x (x, y) u32),
 * Circle FFT twiddles (N - 1 u32, stage-major) and the FRI inverse-domain layers (layer r:
 * N >> (r + 1) QM31 values (t, 0, 0, 0), concatenated).  At least one output; inverse_rounds <= log_n. */
typedef struct azk_domain_plan azk_domain_plan;
typedef enum azk_domain_output {
    AZK_DOMAIN_POINTS = 0,          /* readDomainPoints (:150-168) */
    AZK_DOMAIN_FFT_TWIDDLES = 1,    /* readCodewordTwiddles (:170-183) */
    AZK_DOMAIN_INVERSE_LAYERS = 2   /* readFlatInverseDomain (:202-216) */
} azk_domain_output;
azk_status azk_domain_plan_create(azk_context *ctx, uint32_t log_n, int points, int fft_twiddles, uint32_t inverse_rounds,
                                  azk_domain_plan **out);
void azk_domain_plan_destroy(azk_domain_plan *plan);
/* materializationStats (:41) */
azk_status azk_domain_plan_stats(const azk_domain_plan *plan, azk_stats *stats);
/* inverseDomainLayerCounts / inverseDomainLayerElementOffsets (:37-38), `rounds` entries each. */
azk_status azk_domain_inverse_layout(const azk_domain_plan *plan, uint64_t *counts, uint64_t *offsets, uint32_t rounds);
/* Host readback of one output; an output the plan does not hold gives AZK_ERR_INVALID_LAYOUT. */
azk_status azk_domain_read(azk_domain_plan *plan, azk_domain_output which, void *host, size_t bytes);
/* Resident use (requireDomainPointBuffer and siblings, :134-148): copies the output into a
 * caller buffer at `offset`. */
azk_status azk_domain_copy_resident(azk_domain_plan *plan, azk_domain_output which, azk_buffer *dst, size_t offset);

/* ---- Circle FRI fold chain with explicit challenges
 * CircleFRIFoldPlan / CircleFRIFoldChainPlan (CircleFRIFoldPlan.swift:5-280). */
typedef struct azk_fold_plan azk_fold_plan;
azk_status azk_fold_plan_create(azk_context *ctx, uint32_t log_n, uint32_t rounds, azk_fold_plan **out);
void azk_fold_plan_destroy(azk_fold_plan *plan);
/* challenges: rounds x 4 LE u32; output: N >> rounds QM31. */
azk_status azk_fold_execute_resident(azk_fold_plan *plan, const azk_buffer *evals, size_t evals_offset,
                                     const uint32_t *challenges, azk_buffer *out, size_t out_offset, azk_stats *stats);
azk_status azk_fold_execute(azk_fold_plan *plan, const void *evals, size_t in_bytes, const uint32_t *challenges,
                            void *out, size_t out_bytes, azk_stats *stats);
/* executeVerified (CircleFRIFoldPlan.swift:209-225): against azk_host_fold_chain. */
azk_status azk_fold_execute_verified(azk_fold_plan *plan, const void *evals, size_t in_bytes,
                                     const uint32_t *challenges, void *out, size_t out_bytes, azk_stats *stats);

/* ---- Circle PCS/FRI commit superstep (the prover)
 * CircleCodewordPCSFRIProverV1 (CircleCodeword.swift:920-1419) and
 * CirclePCSFRIResidentProverV1 (CircleFRIFoldPlan.swift:693-1025), security parameters
 * CircleFRISecurityParametersV1 (CircleProofFormat.swift:129-158).  Output: the proof
 * in CirclePCSFRIProofCodecV1 bytes (CircleProofFormat.swift:475-535).  Only roots, the
 * final layer and the opened leaves/paths leave HBM. */
typedef struct azk_circle_prover_desc {
    uint32_t log_n;            /* codeword log size, 2..28 */
    uint32_t rounds;           /* FRI rounds, 1..min(log_n, 28) */
    uint32_t log_blowup;
    uint32_t query_count;
    uint32_t folding_step;     /* must be 1 */
    uint32_t grinding_bits;    /* 0..20 (maximumLocalSearchBits); the nonce search runs on host threads */
    uint8_t public_input_digest[32];
} azk_circle_prover_desc;
typedef struct azk_circle_prover azk_circle_prover;
azk_status azk_circle_prover_create(azk_context *ctx, const azk_circle_prover_desc *desc, azk_circle_prover **out);
void azk_circle_prover_destroy(azk_circle_prover *prover);
size_t azk_circle_prover_proof_size(const azk_circle_prover *prover);
/* proveCircleFFTCoefficientsResident (CircleCodeword.swift:1007-1018). */
azk_status azk_circle_prove_fft_coefficients_resident(azk_circle_prover *prover, const azk_buffer *coeffs,
                                                      size_t offset, uint8_t *proof, size_t capacity,
                                                      size_t *proof_len, azk_stats *stats);
/* proveCircleFFTCoefficientsResident (CircleCodeword.swift:1007-1018) for M31 coefficients:
 * N M31 words, coefficient i lifted to (coeffs[i], 0, 0, 0) -- the same proof bytes as
 * azk_circle_prove_fft_coefficients_resident on the lifted array, with the FFT on M31 words.
 * The reference has no M31-input entry; it is for callers whose coefficients are M31 (the
 * benchmark configuration's are).  Host-array form: `words` >= N. */
azk_status azk_circle_prove_fft_coefficients_m31_resident(azk_circle_prover *prover, const azk_buffer *coeffs,
                                                          size_t offset, uint8_t *proof, size_t capacity,
                                                          size_t *proof_len, azk_stats *stats);
azk_status azk_circle_prove_fft_coefficients_m31(azk_circle_prover *prover, const uint32_t *coeffs, size_t words,
                                                 uint8_t *proof, size_t capacity, size_t *proof_len, azk_stats *stats);
/* prove(evaluationsBuffer:offset:) (CircleFRIFoldPlan.swift:782-848). */
azk_status azk_circle_prove_evaluations_resident(azk_circle_prover *prover, const azk_buffer *evals, size_t offset,
                                                 uint8_t *proof, size_t capacity, size_t *proof_len,
                                                 azk_stats *stats);
/* Host-array variant: uploads the line-basis coefficients, then proves. */
azk_status azk_circle_prove_fft_coefficients(azk_circle_prover *prover, const void *coeffs, size_t bytes,
                                             uint8_t *proof, size_t capacity, size_t *proof_len, azk_stats *stats);
/* proveCircleFFTCoefficientsResidentVerified (CircleCodeword.swift:1082-1113) and the
 * host-array form: AZK_ERR_CORRECTNESS unless azk_circle_verify_with_coefficients accepts
 * the proof (the resident form downloads the coefficients for the gate). */
azk_status azk_circle_prove_fft_coefficients_verified(azk_circle_prover *prover, const void *coeffs, size_t bytes,
                                                      uint8_t *proof, size_t capacity, size_t *proof_len,
                                                      azk_stats *stats);
azk_status azk_circle_prove_fft_coefficients_resident_verified(azk_circle_prover *prover, const azk_buffer *coeffs,
                                                               size_t offset, uint8_t *proof, size_t capacity,
                                                               size_t *proof_len, azk_stats *stats);
/* clearReusableBuffers (CircleCodeword.swift:1238-1253). */
azk_status azk_circle_prover_clear(azk_circle_prover *prover);
/* 1 = replay the superstep as a captured hipGraph (default), 0 = eager stream launches. */
azk_status azk_circle_prover_set_graph(azk_circle_prover *prover, int enable);
/* Per-phase device times of the last prove when timing is on (HIP events). */
typedef struct azk_phase_times {
    uint64_t fft_ns;        /* Circle FFT passes (non-final) */
    uint64_t commit0_ns;    /* final FFT pass fused with leaf hashes + first-layer tree */
    uint64_t fold_ns;       /* fold kernels fused with next-layer leaf hashes + tiles */
    uint64_t tree_ns;       /* upper Merkle levels */
    uint64_t transcript_ns; /* device Fiat-Shamir rounds (critical-path part) */
    uint64_t emit_ns;       /* readback of the tail, index upload, proof gather, proof readback */
    uint64_t host_ns;       /* host final-layer absorb + query squeeze */
    uint64_t total_ns;      /* first to last event, including the host step */
    uint32_t tree_launches; /* upper-tree launches of the program */
    uint32_t tree_levels;   /* upper Merkle levels they reduce */
} azk_phase_times;
azk_status azk_circle_prover_set_timing(azk_circle_prover *prover, int enable);
azk_status azk_circle_prover_phase_times(const azk_circle_prover *prover, azk_phase_times *out);

/* ---- Execution plan and autotuner.  The tile shapes of a prover's program: the
 * reference's plan configurations (MerkleCommitter treelet depth, threads per
 * threadgroup) chosen by MetalProofPlanner.tuneMerkleCommitExecutionPlan
 * (Planner/MetalProofPlanner.swift:20-117) and persisted by PlanDatabase
 * (Planner/PlanDatabase.swift:43-320).  Every plan produces the same proof bytes. */
typedef struct azk_circle_exec_plan {
    uint32_t log_tile_commit;     /* final FFT pass + first-layer treelets: 2^6..2^11 leaves per block (11) */
    uint32_t log_tile_fold;       /* committed folds, big-tile form: 2^6..2^11 outputs per block (11) */
    uint32_t log_leafwise_below;  /* committed layers below 2^this fold one leaf per thread (19) */
    uint32_t log_tile_leafwise;   /* at most 2^3..2^8 leaves per block in that form (8) */
    uint32_t log_tree_chunk;      /* upper Merkle levels reduced per block, 3..10 (3) */
    uint32_t log_leafwise_grid;   /* a leafwise layer of 2^k outputs takes tiles of 2^(k - this)
                                     leaves (2^(k - this + ceil(log2 J)) in a batch of J jobs),
                                     clamped to 2^3..2^log_tile_leafwise; 0 = fixed tile, else
                                     4..20 (8) */
    uint32_t tree_waves;          /* waves per upper-tree block: 4, 8 or 16 (4) */
} azk_circle_exec_plan;
azk_status azk_circle_prover_set_exec_plan(azk_circle_prover *prover, const azk_circle_exec_plan *plan);
azk_status azk_circle_prover_get_exec_plan(const azk_circle_prover *prover, azk_circle_exec_plan *plan);
/* TuningConfiguration (MetalProofPlanner.swift:617-636); zero fields take the defaults. */
typedef struct azk_tuning_config {
    uint32_t warmup_runs;       /* 2 */
    uint32_t measured_runs;     /* 5 */
    uint32_t max_candidates;    /* 28 plans raced (coordinate descent from the default) */
    uint32_t validation_runs;   /* 3 random inputs on which the winner must equal the default plan */
    const char *database_path;  /* JSON-lines plan records; NULL = no persistence */
    int force;                  /* 1 = race even when the database holds a winner */
} azk_tuning_config;
typedef struct azk_tuning_result {
    azk_circle_exec_plan winner;   /* also set on the prover */
    double median_gpu_ns, p95_gpu_ns, median_submit_ns;
    double default_median_gpu_ns;  /* the default plan in the same race */
    uint32_t candidates_measured, candidates_rejected;
    int from_database;             /* 1 = winner read from the database, no race run */
} azk_tuning_result;
/* Races plans on the caller's coefficients: each candidate's first proof must equal the
 * default plan's proof (else it is rejected), the objective is median GPU time + median
 * submit time with p95 as tie break, and the winner must match the default plan on
 * `validation_runs` random inputs before it is persisted and set. */
azk_status azk_circle_prover_tune(azk_circle_prover *prover, const azk_buffer *coeffs, size_t offset,
                                  const azk_tuning_config *config, azk_tuning_result *result);
/* PlanDriftPolicy / PlanDriftStatus (PlanDatabase.swift:15-36): EMA (alpha 0.2) of live
 * prove times against the tuned median; stale after >= 8 samples and > 25% drift.
 * A stale plan is reported, not changed. */
typedef struct azk_drift_status {
    uint32_t samples;
    int stale;
    double ema_gpu_ns, baseline_gpu_ns, relative_drift;
} azk_drift_status;
azk_status azk_circle_prover_drift(const azk_circle_prover *prover, azk_drift_status *out);

/* CirclePCSFRIProofCodecV1.decode + CirclePCSFRIProofVerifierV1.verify
 * (CircleProofFormat.swift:537-645,1397-1546) on the host: *accepted = 1 when the proof
 * verifies against the 32-byte public-input digest, 0 when it does not; malformed bytes
 * give AZK_ERR_INVALID_LAYOUT (the reference's decode throws). */
azk_status azk_circle_verify(const uint8_t *proof, size_t len, const uint8_t *public_input_digest, int *accepted);
/* The same, and every opened first-layer value (query pairs and claimed openings) equals
 * the host Circle FFT of `coeffs` (4 x 2^logN words, the proof's domain) at its leaf. */
azk_status azk_circle_verify_with_coefficients(const uint8_t *proof, size_t len, const uint8_t *public_input_digest,
                                               const uint32_t *coeffs, size_t coeff_words, int *accepted);
/* CircleFRIGrindingV1.digestMeetsTarget (CircleProofFormat.swift:663-681): 1 when the first
 * `bits` bits of the digest (most significant bit of byte 0 first) are zero; 0 when they are
 * not or when bits exceeds 8 x len. */
int azk_circle_grinding_digest_meets_target(const uint8_t *digest, size_t len, uint32_t bits);

/* CirclePCSFRIPolynomialClaimV1 (CircleProofFormat.swift:1018-1093) as plain arrays: the
 * canonical bit-reversed domain of 2^log_n points, a polynomial in monomial x / y
 * coefficients (QM31, 4 LE u32 each; p(x, y) = X(x) + y Y(x), CircleCodeword.swift:53-70)
 * and claim_count evaluation claims at strictly increasing storage indices, each with its
 * circle point (x, y) and QM31 value. */
typedef struct azk_circle_polynomial_claim {
    uint32_t log_n;
    uint32_t x_count, y_count, claim_count;
    const uint32_t *x_coeffs;        /* x_count x 4 words */
    const uint32_t *y_coeffs;        /* y_count x 4 words */
    const uint64_t *storage_indices; /* claim_count */
    const uint32_t *points;          /* claim_count x 2 words (x, y) */
    const uint32_t *values;          /* claim_count x 4 words */
} azk_circle_polynomial_claim;
/* CirclePCSFRIPolynomialClaimV1.make (:1048-1073): fills claim->points and claim->values
 * (caller-owned arrays passed as `points` / `values`) for the claim's storage indices. */
azk_status azk_circle_polynomial_claim_make(const azk_circle_polynomial_claim *claim, uint32_t *points,
                                            uint32_t *values);
/* CirclePCSFRIPolynomialClaimDigestV1.digest (:1095-1151): the public-input digest a prover
 * of this claim commits to. */
azk_status azk_circle_polynomial_claim_digest(const azk_circle_polynomial_claim *claim, uint8_t digest[32]);
/* CirclePCSFRIPolynomialVerifierV1.verify (:1548-1659): the proof verifies against the
 * claim's digest, its domain is the claim's, the coefficient counts fit the blowup, and
 * every claimed opening matches its claim (index, point, polynomial value) and opens against
 * the first commitment. */
azk_status azk_circle_verify_polynomial(const uint8_t *proof, size_t len, const azk_circle_polynomial_claim *claim,
                                        int *accepted);
/* CirclePCSFRIParameterSetV1 (:207-286): security profile of a statement. */
typedef struct azk_circle_parameter_set {
    uint32_t log_blowup, queries, grinding_bits, target_soundness_bits;
} azk_circle_parameter_set;
/* CirclePCSFRIContractVerifierV1.verify(encodedProof:statement:) (:1726-1752) with the
 * statement CirclePCSFRIStatementV1(parameterSet:polynomialClaim:) (:1660-1700): the proof's
 * security parameters, round count (log_n - log_blowup) and constant final layer of
 * 2^log_blowup values match the profile, then azk_circle_verify_polynomial.  An invalid
 * statement (domain not above the blowup, coefficients over capacity, target above the
 * nominal security) gives AZK_ERR_INVALID_LAYOUT, as its initializer throws. */
azk_status azk_circle_verify_contract(const uint8_t *proof, size_t len, const azk_circle_parameter_set *params,
                                      const azk_circle_polynomial_claim *claim, int *accepted);

/* ---- CPU gates (host code; what the *_verified entry points compare against)
 * CircleCodewordOracle.evaluateWithCircleFFT (CircleCodeword.swift:72-98): N = 2^log_n
 * interleaved line-basis coefficients -> N evaluations, canonical domain, storage order. */
azk_status azk_host_circle_fft(const uint32_t *coeffs, uint32_t log_n, uint32_t *evals);
/* CircleFRIProofBuilderV1.prove (CircleProofFormat.swift:1172-1395) of the Circle FFT of the
 * coefficients, on the host (log N <= 24): the proof the *_verified prover entry points compare
 * the device's with, byte for byte (proveVerified, CircleCodeword.swift:1152-1177). */
azk_status azk_host_circle_prove(const uint32_t *coeffs, const azk_circle_prover_desc *desc, uint8_t *proof,
                                 size_t capacity, size_t *len);
/* CircleFRILayerOracleV1.fold (CircleDomain.swift:513-544): rounds x 4 challenge words. */
azk_status azk_host_fold_chain(const uint32_t *evals, uint32_t log_n, const uint32_t *challenges, uint32_t rounds,
                               uint32_t *out);
/* No reference counterpart (its gate is the O(n^2) oracle): *consistent = 1 when the basis
 * and the monomial columns agree at two pseudo-random QM31 points per column. */
azk_status azk_witness_basis_check(const uint32_t *x, uint32_t x_count, const uint32_t *y, uint32_t y_count,
                                   uint32_t log_n, const uint32_t *basis, int *consistent);

/* ---- Merkle-transcript FRI fold chain
 * CircleFRIMerkleTranscriptFoldChainPlan (CircleFRIFoldPlan.swift:301-649): per round commit
 * the layer, absorb the root, squeeze the challenge, fold.  Outputs rounds x 32-byte roots,
 * rounds x 4 challenge words and the final layer (2^(log_n - rounds) QM31).  The descriptor
 * is the prover's (its security parameters and digest enter the transcript prefix). */
typedef struct azk_chain_plan azk_chain_plan;
azk_status azk_chain_plan_create(azk_context *ctx, const azk_circle_prover_desc *desc, azk_chain_plan **out);
void azk_chain_plan_destroy(azk_chain_plan *plan);
/* executeResident (:433-455) */
azk_status azk_chain_execute_resident(azk_chain_plan *plan, const azk_buffer *evals, size_t evals_offset,
                                      uint8_t *commitments, uint32_t *challenges, azk_buffer *out, size_t out_offset,
                                      azk_stats *stats);
/* execute(evaluations:) (:390-408) */
azk_status azk_chain_execute(azk_chain_plan *plan, const uint32_t *evals, size_t eval_words, uint8_t *commitments,
                             uint32_t *challenges, uint32_t *final_layer, size_t final_words, azk_stats *stats);
/* executeVerified (:410-431): host Merkle roots, host transcript and host folds. */
azk_status azk_chain_execute_verified(azk_chain_plan *plan, const uint32_t *evals, size_t eval_words,
                                      uint8_t *commitments, uint32_t *challenges, uint32_t *final_layer,
                                      size_t final_words, azk_stats *stats);

/* ---- Batched form: many independent commits of uneven sizes in one captured graph.
 * No reference counterpart (the reference runs one command buffer per job); each
 * job's proof is byte-identical to its plain single-job prove.  A batch plan captures
 * the jobs once (provers, coefficient buffers and offsets are fixed) and replays. */
typedef struct azk_circle_batch azk_circle_batch;
azk_status azk_circle_batch_create(azk_circle_prover *const *provers, const azk_buffer *const *coeffs,
                                   const size_t *offsets, uint32_t jobs, azk_circle_batch **out);
/* The same batch from M31 coefficients (N words per job, as azk_circle_prove_fft_coefficients_m31_resident:
 * each job's proof equals its plain prove; 4-byte aligned buffers).  No reference counterpart. */
azk_status azk_circle_batch_create_m31(azk_circle_prover *const *provers, const azk_buffer *const *coeffs,
                                       const size_t *offsets, uint32_t jobs, azk_circle_batch **out);
void azk_circle_batch_destroy(azk_circle_batch *batch);
azk_status azk_circle_batch_prove(azk_circle_batch *batch, uint8_t *const *proofs, const size_t *capacities,
                                  size_t *proof_lens, azk_stats *stats);
/* The autotuner on the batch program as a whole (azk_circle_prover_tune's race, gate and store,
 * MetalProofPlanner.swift:20-117): one execution plan for every job, each candidate's proofs
 * equal to the default plan's for every job, the batch's graph time as the objective; records
 * keyed by the batch's job shapes.  The winner is left set on every prover of the batch. */
azk_status azk_circle_batch_tune(azk_circle_batch *batch, const azk_tuning_config *config, azk_tuning_result *result);
/* One-shot form: create, prove, destroy. */
azk_status azk_circle_prove_batch(azk_circle_prover *const *provers, const azk_buffer *const *coeffs,
                                  const size_t *offsets, uint32_t jobs, uint8_t *const *proofs,
                                  const size_t *capacities, size_t *proof_lens, azk_stats *stats);

/* ---- M31 multilinear sum-check (the serving job)
 * M31MultilinearSumcheckProofBuilderV1.prove (SumcheckOracle.swift:817-885); the
 * reference runs it on the CPU only.  Rounds, transcript and folds run on the device;
 * the initial digest M31SumcheckEncodingV1.digestWords (:1288-1299), a serial SHA3 over
 * 4n bytes, runs on a host thread. */
typedef struct azk_sumcheck_plan azk_sumcheck_plan;
typedef struct azk_sumcheck_result {
    uint32_t claimed_sum;        /* statement.claimedHypercubeSum */
    uint32_t final_evaluation;
    uint8_t initial_digest[32];  /* statement.initialEvaluationDigest */
    uint32_t *round_polys;       /* caller memory, 2 * rounds words (e0, e1) or NULL */
    uint32_t *challenges;        /* caller memory, rounds words or NULL */
} azk_sumcheck_result;
/* Plan for 2^log_n lanes, 1 <= log_n <= 30 (M31MultilinearSumcheckStatementV1 :711-739). */
azk_status azk_sumcheck_plan_create(azk_context *ctx, uint32_t log_n, azk_sumcheck_plan **out);
void azk_sumcheck_plan_destroy(azk_sumcheck_plan *plan);
/* Bytes of M31MultilinearSumcheckProofCodecV1.encode for 2^log_n lanes (:1176-1194). */
size_t azk_sumcheck_proof_size(uint32_t log_n);
/* prove(evaluations:claimedHypercubeSum:) on a host vector; writes the codec bytes.
 * claimed_sum may be NULL; a wrong one gives AZK_ERR_CORRECTNESS, non-canonical words
 * AZK_ERR_INVALID_LAYOUT. */
azk_status azk_sumcheck_prove(azk_sumcheck_plan *plan, const uint32_t *evals, size_t count, const uint32_t *claimed_sum,
                              uint8_t *proof, size_t capacity, size_t *proof_len, azk_stats *stats);
/* The same with the initial digest supplied by a caller that already holds digestWords of
 * the words (SumcheckOracle.swift:1288-1299; e.g. an AIR reduction that digests its evaluation
 * vector, ApplicationTheorem.swift:2357-2362): the serial host hash of the words is skipped.  A
 * digest that is not the words' gives a proof the verifier refuses. */
azk_status azk_sumcheck_prove_with_digest(azk_sumcheck_plan *plan, const uint32_t *evals, size_t count,
                                          const uint8_t *initial_digest, const uint32_t *claimed_sum, uint8_t *proof,
                                          size_t capacity, size_t *proof_len, azk_stats *stats);
/* Resident form: words already in device memory.  `initial_digest` is digestWords of
 * the same words when the caller has it, else NULL (the words are read back and hashed). */
azk_status azk_sumcheck_prove_resident(azk_sumcheck_plan *plan, const azk_buffer *evals, size_t offset,
                                       const uint8_t *initial_digest, const uint32_t *claimed_sum,
                                       azk_sumcheck_result *result, azk_stats *stats);
/* ---- M31 sum-check chunk (the reference's GPU sum-check)
 * MetalSumcheckChunkPlan(context:descriptor:) with SumcheckChunkDescriptor(laneCount:roundsPerSuperstep:)
 * (Planner/MetalSumcheckChunkPlan.swift:20-28,72-210): lane_count a power of two in 2..2^30,
 * 1 <= rounds <= log2(lane_count).  Each round absorbs the whole active vector, squeezes one
 * challenge and folds v[2i] + r v[2i + 1] (SumcheckOracle.m31Chunk, SumcheckOracle.swift:212-276).
 * Outputs (SumcheckChunkOracleResult): final vector (lane_count >> rounds words), the coefficient
 * log (every round's vector, azk_sumcheck_chunk_coefficient_words in all) and the challenges. */
typedef struct azk_sumcheck_chunk_plan azk_sumcheck_chunk_plan;
azk_status azk_sumcheck_chunk_plan_create(azk_context *ctx, uint64_t lane_count, uint32_t rounds,
                                          azk_sumcheck_chunk_plan **out);
void azk_sumcheck_chunk_plan_destroy(azk_sumcheck_chunk_plan *plan);
size_t azk_sumcheck_chunk_coefficient_words(uint64_t lane_count, uint32_t rounds);
/* execute(evaluations:) (:212-221): canonical host words. */
azk_status azk_sumcheck_chunk_execute(azk_sumcheck_chunk_plan *plan, const uint32_t *evals, size_t count,
                                      uint32_t *final_vector, uint32_t *coefficients, uint32_t *challenges,
                                      azk_stats *stats);
/* executeUploadedVector(buffer:offset:) (:235-243): words already in device memory (8-byte aligned). */
azk_status azk_sumcheck_chunk_execute_resident(azk_sumcheck_chunk_plan *plan, const azk_buffer *evals, size_t offset,
                                               uint32_t *final_vector, uint32_t *coefficients, uint32_t *challenges,
                                               azk_stats *stats);
/* executeVerified (:223-233): AZK_ERR_CORRECTNESS unless a host recomputation agrees. */
azk_status azk_sumcheck_chunk_execute_verified(azk_sumcheck_chunk_plan *plan, const uint32_t *evals, size_t count,
                                               uint32_t *final_vector, uint32_t *coefficients, uint32_t *challenges,
                                               azk_stats *stats);
/* clearReusableBuffers (:245-258). */
azk_status azk_sumcheck_chunk_plan_clear(azk_sumcheck_chunk_plan *plan);
/* M31SumcheckProofBuilderV1.prove + M31SumcheckProofCodecV1.encode (SumcheckOracle.swift:540-556,
 * 1108-1124): the "AZKSCV1" proof bytes (0 size for an invalid shape). */
size_t azk_sumcheck_chunk_proof_size(uint64_t lane_count, uint32_t rounds);
azk_status azk_sumcheck_chunk_prove(azk_sumcheck_chunk_plan *plan, const uint32_t *evals, size_t count, uint8_t *proof,
                                    size_t capacity, size_t *len, azk_stats *stats);
/* M31SumcheckProofCodecV1.decode + M31SumcheckVerifierV1.verify(proof:statement:) against the
 * proof's own statement, scope revealedEvaluationVectorFoldingTrace (:558-700): *accepted = 1 when
 * both digests, every challenge and every fold agree; malformed bytes give AZK_ERR_INVALID_LAYOUT. */
azk_status azk_sumcheck_chunk_verify(const uint8_t *proof, size_t len, int *accepted);

/* M31MultilinearSumcheckProofCodecV1.decode + M31MultilinearSumcheckVerifierV1.verify
 * (:941-1064,1196-1233) against the statement the proof carries; *accepted = 1 or 0.
 * Malformed bytes give AZK_ERR_INVALID_LAYOUT (the reference's decode throws). */
azk_status azk_sumcheck_verify(const uint8_t *proof, size_t len, int *accepted);

/* ---- Serving form: jobs arrive at any time; no reference counterpart (the reference runs
 * one synchronous command buffer per call).  Commit jobs are run by continuous batching
 * (each device program covers every admitted commit, up to max_batch); sum-check and
 * fold-chain jobs run on their own lanes.  A job is admitted when the device state it needs fits the HBM
 * budget next to what is live; its result is byte-identical to the plain call's.
 * Inputs are read in place: the caller keeps them unchanged until the job is done
 * (azk_server_wait).  Pinned inputs (azk_host_alloc) upload asynchronously. */
typedef struct azk_server azk_server;
typedef struct azk_server_desc {
    int device;
    uint32_t max_batch;       /* commits per device program; 0 = 64 */
    uint32_t aux_lanes;       /* concurrent sum-check / fold-chain jobs, 1..16; 0 = 4 */
    uint32_t reserved;
    uint64_t hbm_budget;      /* bytes of live job state; 0 = half the free device memory */
} azk_server_desc;
typedef struct azk_job_info {
    azk_status status;
    uint32_t kind;            /* 0 = Circle PCS commit, 1 = multilinear sum-check, 2 = fold chain, 3 = sum-check chunk */
    uint32_t batch_jobs;      /* commits in the device program that ran the job */
    size_t result_bytes;      /* proof bytes (CirclePCSFRIProofCodecV1 / M31MultilinearSumcheckProofCodecV1) */
    uint64_t submit_ns, admit_ns, start_ns, done_ns;  /* azk_now_ns clock */
} azk_job_info;
azk_status azk_server_create(const azk_server_desc *desc, azk_server **out);
/* Runs every submitted job to completion, then frees the server. */
void azk_server_destroy(azk_server *server);
azk_status azk_server_submit_commit(azk_server *server, const azk_circle_prover_desc *desc, const void *coeffs,
                                    size_t bytes, uint64_t *ticket);
/* A commit from M31 line-basis coefficients (at least 2^log_n words; coefficient i's QM31 lift is
 * (coeffs[i], 0, 0, 0)), as azk_circle_prove_fft_coefficients_m31: the proof bytes equal those of
 * azk_server_submit_commit on the lifted array, from a quarter of the upload. */
azk_status azk_server_submit_commit_m31(azk_server *server, const azk_circle_prover_desc *desc, const uint32_t *coeffs,
                                        size_t words, uint64_t *ticket);
azk_status azk_server_submit_sumcheck(azk_server *server, const uint32_t *evals, size_t count, uint64_t *ticket);
/* A multilinear sum-check whose initial digest the caller supplies (azk_sumcheck_prove_with_digest). */
azk_status azk_server_submit_sumcheck_digest(azk_server *server, const uint32_t *evals, size_t count,
                                             const uint8_t digest[32], uint64_t *ticket);
/* A sum-check chunk of `rounds` rounds (azk_sumcheck_chunk_prove): the result is the AZKSCV1 proof. */
azk_status azk_server_submit_sumcheck_chunk(azk_server *server, const uint32_t *evals, size_t count, uint32_t rounds,
                                            uint64_t *ticket);
/* A CircleFRIFoldChainPlan job (CircleFRIFoldPlan.swift:132-280): 2^log_n QM31 evaluations
 * folded `rounds` times with the given challenges (4 words per round, copied at submit);
 * the result is the 2^(log_n - rounds) folded QM31 values. */
azk_status azk_server_submit_fold(azk_server *server, uint32_t log_n, uint32_t rounds, const void *evals, size_t bytes,
                                  const uint32_t *challenges, uint64_t *ticket);
/* Blocks until the job is done; returns the job's status. */
azk_status azk_server_wait(azk_server *server, uint64_t ticket, azk_job_info *info);
/* Copies the result (out may be NULL) and forgets the ticket; waits if needed. */
azk_status azk_server_take(azk_server *server, uint64_t ticket, uint8_t *out, size_t capacity, size_t *len);
uint64_t azk_now_ns(void);
/* Pinned host memory for inputs (the buffer kind "pinned_host"). */
azk_status azk_host_alloc(size_t bytes, void **out);
void azk_host_free(void *ptr);

/* ---- Measurement support (no reference counterpart): the roofline's VALU figures measured on
 * this chip.  cycles_per_instr: SIMD cycles per wave64 instruction of v_xor_b32, v_bitop3_b32,
 * v_alignbit_b32, v_add_u32, v_min_u32, v_mad_u64_u32 (8 independent chains per lane,
 * registers only), at clock_mhz (s_memtime against the 100 MHz s_memrealtime); the Keccak
 * figures are the one-lane Keccak-f[1600] on register-resident states (no memory traffic). */
typedef struct azk_valu_calibration {
    double keccak_perms_per_s, keccak_clock_mhz, keccak_cycles_per_wave_perm;
    double cycles_per_instr[6], clock_mhz[6];
} azk_valu_calibration;
azk_status azk_calibrate_valu(azk_context *ctx, azk_valu_calibration *out);
/* The serial chain's latency figures on this chip (no reference counterpart): microseconds per
 * dependent Keccak-f[1600] on one wave in the one-state-per-wave form (241 in a row, the 100 MHz
 * clock read in the kernel) and in the two-lanes-per-state form (32 states on the wave), and
 * microseconds per kernel of a captured graph of 64 back-to-back empty kernels. */
typedef struct azk_latency_calibration {
    double wave_perm_us, pair_perm_us, graph_kernel_us;
} azk_latency_calibration;
azk_status azk_calibrate_latency(azk_context *ctx, azk_latency_calibration *out);

#ifdef __cplusplus
}
#endif
#endif /* AZK_H */
