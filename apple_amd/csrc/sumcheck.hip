// M31 multilinear sum-check on gfx950: the serving job's device path.
//
// Reference: M31MultilinearSumcheckProofBuilderV1.prove (SumcheckOracle.swift:817-885),
// round polynomial (:887-902), fold (1 - r) a + r b (:904-924), transcript framing
// M31MultilinearSumcheckTranscriptV1 (:1066-1106), initial digest
// M31SumcheckEncodingV1.digestWords (:1288-1299), codec M31MultilinearSumcheckProofCodecV1
// (:1173-1233) and verifier M31MultilinearSumcheckVerifierV1 (:941-1064).  The reference
// runs all of it on the CPU.
//
// Device program for n = 2^k words (k rounds):
//   k_sc_reduce0   even/odd sums of the input (+ count of words >= p), per block
//   k_sc_round     one wave: sums the partials, absorbs the header (round 0), the round
//                  frame and the challenge frame, squeezes the challenge (one-state-per-wave
//                  Keccak)
//   k_sc_fold      v'[i] = v[2i] + c (v[2i+1] - v[2i]) fused with the next round's sums
//   k_sc_tail      the last rounds (<= 2^kTailLog words) in one block: LDS-resident
//                  vector, block reductions, wave-0 transcript, LDS folds
// (1 - c) a + c b and a + c (b - a) are the same element of F_p, so the canonical
// outputs equal the reference's.
//
// The initial digest is one serial SHA3 sponge over 4n + 50 bytes (up to 123k
// permutations at n = 2^22); it runs on a host thread while the device uploads and
// reduces, and only the transcript waits for it.
#include <thread>

#include "device_field.hpp"
#include "device_keccak.hpp"
#include "host_sha3.hpp"
#include "runtime_common.hpp"

using namespace azk;
using namespace azk::rt;

namespace {

constexpr uint32_t kScThreads = 256;
constexpr uint32_t kTailLog = 12;          // vectors of <= 4096 words finish in one block
constexpr uint32_t kMaxPartials = 1024;    // blocks of the reduce / fold kernels
constexpr uint32_t kFrameLanes = 17;       // one padded 136-byte block
constexpr uint32_t kMaxLogN = 30;

// Byte offsets inside the frames (base frame = u32 len || 36-byte domain || u32 ver || u8 type).
constexpr int kBase = 4 + 36 + 4 + 1;
constexpr int kHdrClaimed = kBase + 8 + 4;  // header: u64 n, u32 rounds, u32 claimed, digest
constexpr int kRoundE0 = kBase + 4;         // round: u32 index, u32 e0, u32 e1
constexpr int kRoundE1 = kBase + 8;

// out words: [0, 2R) round polynomials, [2R, 3R) challenges, 3R final, 3R+1 claimed sum,
// 3R+2 count of non-canonical input words (clamped).
struct ScOut {
    uint32_t *w;
    uint32_t rounds;
};

__device__ __forceinline__ uint64_t wave_sum(uint64_t v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        uint32_t lo = __shfl_xor((uint32_t)v, o), hi = __shfl_xor((uint32_t)(v >> 32), o);
        v += (uint64_t)lo | ((uint64_t)hi << 32);
    }
    return v;
}

// Block sums of three u64 values; thread 0 gets the totals.
__device__ __forceinline__ void block_sum3(uint64_t &a, uint64_t &b, uint64_t &c)
{
    __shared__ uint64_t sh[3][kScThreads / 64];
    a = wave_sum(a);
    b = wave_sum(b);
    c = wave_sum(c);
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) { sh[0][w] = a; sh[1][w] = b; sh[2][w] = c; }
    __syncthreads();
    if (threadIdx.x == 0) {
        a = b = c = 0;
#pragma unroll
        for (uint32_t i = 0; i < kScThreads / 64; ++i) { a += sh[0][i]; b += sh[1][i]; c += sh[2][i]; }
    }
}

__device__ __forceinline__ uint32_t uni(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// The transcript runs on one wave in the one-state-per-wave layout: every lane that holds
// state word w (its own lane and any copy) applies the same block word, so a padded frame
// block is read as one u64 per lane.

// This lane's part of a little-endian u32 at byte offset OFF of a padded block.
template <int OFF>
__device__ __forceinline__ uint64_t put32(const WaveLane &c, uint32_t v)
{
    constexpr uint32_t lane = OFF / 8, sh = (OFF % 8) * 8;
    uint64_t x = c.word == lane ? (uint64_t)v << sh : 0;
    if constexpr (sh > 32) {
        if (c.word == lane + 1) x ^= (uint64_t)v >> (64 - sh);
    }
    return x;
}

__device__ __forceinline__ uint64_t block_word(const WaveLane &c, const uint64_t *src)
{
    return c.word < kFrameLanes ? src[c.word] : 0;
}

__device__ __forceinline__ void absorb_block(uint32_t &lo, uint32_t &hi, const WaveLane &c, uint64_t w)
{
    lo ^= (uint32_t)w;
    hi ^= (uint32_t)(w >> 32);
    kf_w(lo, hi, c);
}

// First u32 below 2^32 - (2^32 mod p) read from a copy of the state, reduced mod p
// (TranscriptState.squeezeUInt32 with count 1, SHA3Oracle.swift:41-70).  Only the two
// words 2^32 - 2 and 2^32 - 1 are rejected; the loop bound (2^16 candidates) is never
// reached by a permutation output and only keeps the loop finite.
__device__ __forceinline__ uint32_t squeeze_m31(uint32_t lo, uint32_t hi, const WaveLane &c)
{
    const uint64_t limit = (1ull << 32) - ((1ull << 32) % M31_P);
    for (uint32_t cand = 0; cand < (1u << 16); ++cand) {
        if (cand > 0 && cand % 34 == 0) kf_w(lo, hi, c);
        const uint32_t wi = cand % 34, at = wave_lane_of(wi >> 1);
        // readlane returns a signed int: keep the word unsigned before widening
        const uint32_t w = (uint32_t)__builtin_amdgcn_readlane((wi & 1) ? hi : lo, at);
        if (w < limit) return (uint32_t)(w % M31_P);
    }
    return 0;
}

// One round of the transcript on the wave's state: header first when r == 0, then the
// round frame (r, e0, e1) and the challenge frame (r, p); returns the challenge.  `tmpl`
// holds padded blocks: [header][round 0][challenge 0][round 1]...
__device__ __forceinline__ uint32_t sc_transcript_round(uint32_t &lo, uint32_t &hi, const WaveLane &c,
                                                        const uint64_t *tmpl, uint32_t r, uint32_t e0, uint32_t e1)
{
    const uint64_t *rt = tmpl + (size_t)kFrameLanes * (1 + 2 * r);
    const uint64_t round_w = block_word(c, rt), chal_w = block_word(c, rt + kFrameLanes);  // both loads in flight
    if (r == 0) absorb_block(lo, hi, c, block_word(c, tmpl) ^ put32<kHdrClaimed>(c, m31_add(e0, e1)));
    absorb_block(lo, hi, c, round_w ^ put32<kRoundE0>(c, e0) ^ put32<kRoundE1>(c, e1));
    absorb_block(lo, hi, c, chal_w);
    return squeeze_m31(lo, hi, c);
}

// The state words of this lane from memory (zero state when r == 0), and back (own lanes).
__device__ __forceinline__ void load_state(uint32_t &lo, uint32_t &hi, const WaveLane &c, const uint64_t *state,
                                           bool zero)
{
    const uint64_t v = (!zero && c.word < 25) ? state[c.word] : 0;
    lo = (uint32_t)v;
    hi = (uint32_t)(v >> 32);
}

// Per-block partial sums of the input: even-index words, odd-index words, words >= p.
__global__ __launch_bounds__(kScThreads) void k_sc_reduce0(const uint32_t *__restrict__ v, uint64_t m,
                                                           uint64_t *__restrict__ partials)
{
    uint64_t e = 0, o = 0, bad = 0;
    const uint64_t stride = (uint64_t)gridDim.x * kScThreads;  // even
    for (uint64_t i = (uint64_t)blockIdx.x * kScThreads + threadIdx.x; i < m; i += stride) {
        const uint32_t x = v[i];
        bad += x >= M31_P;
        if (i & 1) o += x; else e += x;
    }
    block_sum3(e, o, bad);
    if (threadIdx.x == 0) {
        partials[3 * blockIdx.x] = e;
        partials[3 * blockIdx.x + 1] = o;
        partials[3 * blockIdx.x + 2] = bad;
    }
}

// One wave: round r of the transcript from `nparts` partial sums.
__global__ __launch_bounds__(64) void k_sc_round(const uint64_t *__restrict__ partials, uint32_t nparts, uint32_t r,
                                                 const uint64_t *__restrict__ tmpl, uint64_t *__restrict__ state,
                                                 ScOut out)
{
    uint64_t e = 0, o = 0, bad = 0;
    for (uint32_t i = threadIdx.x; i < nparts; i += 64) {
        e += partials[3 * i];
        o += partials[3 * i + 1];
        bad += partials[3 * i + 2];
    }
    e = wave_sum(e);
    o = wave_sum(o);
    bad = wave_sum(bad);
    const uint32_t e0 = uni((uint32_t)(e % M31_P)), e1 = uni((uint32_t)(o % M31_P));
    const WaveLane wl = wave_lane();
    uint32_t lo, hi;
    load_state(lo, hi, wl, state, r == 0);
    const uint32_t c = sc_transcript_round(lo, hi, wl, tmpl, r, e0, e1);
    if (threadIdx.x == 0) {
        out.w[2 * r] = e0;
        out.w[2 * r + 1] = e1;
        out.w[2 * out.rounds + r] = c;
        if (r == 0) {
            out.w[3 * out.rounds + 1] = m31_add(e0, e1);
            out.w[3 * out.rounds + 2] = bad > 0xffffffffull ? 0xffffffffu : (uint32_t)bad;
        }
    }
    if (wl.own) state[wl.word] = (uint64_t)lo | ((uint64_t)hi << 32);
}

// Fold of round r (challenge read from out) fused with the partial sums of round r + 1.
__global__ __launch_bounds__(kScThreads) void k_sc_fold(const uint32_t *__restrict__ in, uint32_t *__restrict__ dst,
                                                        uint64_t m_out, const uint32_t *__restrict__ ch,
                                                        uint64_t *__restrict__ partials)
{
    const uint32_t c = *ch;
    uint64_t e = 0, o = 0, zero = 0;
    const uint64_t stride = (uint64_t)gridDim.x * kScThreads;
    for (uint64_t i = (uint64_t)blockIdx.x * kScThreads + threadIdx.x; i < m_out; i += stride) {
        const uint32_t x0 = in[2 * i], x1 = in[2 * i + 1];
        const uint32_t y = m31_add(x0, m31_mul(c, m31_sub(x1, x0)));
        dst[i] = y;
        if (i & 1) o += y; else e += y;
    }
    block_sum3(e, o, zero);
    if (threadIdx.x == 0) {
        partials[3 * blockIdx.x] = e;
        partials[3 * blockIdx.x + 1] = o;
        partials[3 * blockIdx.x + 2] = 0;
    }
}

// Rounds r0 .. rounds-1 of a vector of 2^(rounds - r0) <= 2^kTailLog words in one block.
__global__ __launch_bounds__(kScThreads) void k_sc_tail(const uint32_t *__restrict__ in, uint32_t r0, uint32_t rounds,
                                                        const uint64_t *__restrict__ tmpl,
                                                        uint64_t *__restrict__ state, ScOut out)
{
    __shared__ uint32_t A[1u << kTailLog], B[1u << (kTailLog - 1)];
    __shared__ uint32_t sh_c;
    __shared__ uint64_t sh_sum[3];
    uint32_t m = 1u << (rounds - r0);
    uint64_t bad = 0;
    for (uint32_t i = threadIdx.x; i < m; i += kScThreads) {
        const uint32_t x = in[i];
        bad += x >= M31_P;
        A[i] = x;
    }
    __syncthreads();
    const bool w0 = threadIdx.x < 64;  // wave 0 holds the transcript state
    const WaveLane wl = wave_lane();
    uint32_t lo = 0, hi = 0;
    if (w0) load_state(lo, hi, wl, state, r0 == 0);
    uint32_t *cur = A, *nxt = B;
    for (uint32_t r = r0; r < rounds; ++r) {
        uint64_t e = 0, o = 0;
        for (uint32_t i = threadIdx.x; i < m; i += kScThreads) {  // kScThreads even: parity fixed
            if (i & 1) o += cur[i]; else e += cur[i];
        }
        uint64_t b = (r == r0) ? bad : 0;
        block_sum3(e, o, b);
        if (threadIdx.x == 0) { sh_sum[0] = e; sh_sum[1] = o; sh_sum[2] = b; }
        __syncthreads();
        if (w0) {
            const uint32_t e0 = uni((uint32_t)(sh_sum[0] % M31_P)), e1 = uni((uint32_t)(sh_sum[1] % M31_P));
            const uint32_t c = sc_transcript_round(lo, hi, wl, tmpl, r, e0, e1);
            if (threadIdx.x == 0) {
                sh_c = c;
                out.w[2 * r] = e0;
                out.w[2 * r + 1] = e1;
                out.w[2 * out.rounds + r] = c;
                if (r == 0) {
                    out.w[3 * out.rounds + 1] = m31_add(e0, e1);
                    out.w[3 * out.rounds + 2] = sh_sum[2] > 0xffffffffull ? 0xffffffffu : (uint32_t)sh_sum[2];
                }
            }
        }
        __syncthreads();
        const uint32_t c = sh_c;
        const uint32_t half = m >> 1;
        for (uint32_t i = threadIdx.x; i < half; i += kScThreads) {
            const uint32_t x0 = cur[2 * i], x1 = cur[2 * i + 1];
            nxt[i] = m31_add(x0, m31_mul(c, m31_sub(x1, x0)));
        }
        __syncthreads();
        uint32_t *t = cur;
        cur = nxt;
        nxt = t;
        m = half;
    }
    if (threadIdx.x == 0) out.w[3 * out.rounds] = cur[0];
    if (w0 && wl.own) state[wl.word] = (uint64_t)lo | ((uint64_t)hi << 32);
}

// ------------------------------------------------------------------ host side

void put_u32(std::vector<uint8_t> &f, uint32_t v) { for (int i = 0; i < 4; ++i) f.push_back((uint8_t)(v >> (8 * i))); }
void put_u64(std::vector<uint8_t> &f, uint64_t v) { for (int i = 0; i < 8; ++i) f.push_back((uint8_t)(v >> (8 * i))); }

const char kDomain[] = "AppleZKProver.M31MultilinearSumcheck";
const char kWordsDomain[] = "AppleZKProver.M31Sumcheck.Words.V1";
const uint8_t kMagic[8] = {0x41, 0x5a, 0x4b, 0x4d, 0x53, 0x56, 0x31, 0x00};  // "AZKMSV1\0"

// baseFrame (SumcheckOracle.swift:1096-1103)
std::vector<uint8_t> base_frame(uint8_t type)
{
    std::vector<uint8_t> f;
    put_u32(f, (uint32_t)(sizeof kDomain - 1));
    f.insert(f.end(), kDomain, kDomain + sizeof kDomain - 1);
    put_u32(f, 1);
    f.push_back(type);
    return f;
}

std::vector<uint8_t> header_frame(uint64_t n, uint32_t rounds, uint32_t claimed, const uint8_t digest[32])
{
    std::vector<uint8_t> f = base_frame(0);
    put_u64(f, n);
    put_u32(f, rounds);
    put_u32(f, claimed);
    f.insert(f.end(), digest, digest + 32);
    return f;
}
std::vector<uint8_t> round_frame(uint32_t r, uint32_t e0, uint32_t e1)
{
    std::vector<uint8_t> f = base_frame(1);
    put_u32(f, r);
    put_u32(f, e0);
    put_u32(f, e1);
    return f;
}
std::vector<uint8_t> challenge_frame(uint32_t r)
{
    std::vector<uint8_t> f = base_frame(2);
    put_u32(f, r);
    put_u32(f, kP);
    return f;
}

// A frame as a padded one-block message (SHA3 suffix 0x06), as lanes.
void pad_block(const std::vector<uint8_t> &f, uint64_t *lanes)
{
    uint8_t blk[136] = {0};
    memcpy(blk, f.data(), f.size());
    blk[f.size()] ^= 0x06;
    blk[135] ^= 0x80;
    memcpy(lanes, blk, 136);
}

// M31SumcheckEncodingV1.digestWords (SumcheckOracle.swift:1288-1299): SHA3-256 of
// u32 len || domain || u32 1 || u64 n || words LE, streamed without a copy of the words.
void words_digest(const uint32_t *w, uint64_t n, uint8_t out[32])
{
    std::vector<uint8_t> pre;
    put_u32(pre, (uint32_t)(sizeof kWordsDomain - 1));
    pre.insert(pre.end(), kWordsDomain, kWordsDomain + sizeof kWordsDomain - 1);
    put_u32(pre, 1);
    put_u64(pre, n);
    uint64_t s[25] = {0};
    uint8_t blk[136];
    size_t fill = pre.size();
    memcpy(blk, pre.data(), fill);
    const uint8_t *src = reinterpret_cast<const uint8_t *>(w);
    uint64_t left = n * 4;
    while (left) {
        const size_t take = (size_t)std::min<uint64_t>(left, 136 - fill);
        memcpy(blk + fill, src, take);
        fill += take;
        src += take;
        left -= take;
        if (fill == 136) {
            host::xor_block(s, blk, 136);
            host::keccak_f(s);
            fill = 0;
        }
    }
    memset(blk + fill, 0, 136 - fill);
    blk[fill] ^= 0x06;
    blk[135] ^= 0x80;
    host::xor_block(s, blk, 136);
    host::keccak_f(s);
    memcpy(out, s, 32);
}

uint32_t hadd(uint32_t a, uint32_t b) { uint32_t s = a + b; return s >= kP ? s - kP : s; }
uint32_t hsub(uint32_t a, uint32_t b) { return a >= b ? a - b : a + (kP - b); }
uint32_t hmul(uint32_t a, uint32_t b) { return (uint32_t)(((uint64_t)a * b) % kP); }

} // namespace

void azk::host::sumcheck_words_digest(const uint32_t *words, uint64_t n, uint8_t out[32]) { words_digest(words, n, out); }

struct azk_sumcheck_plan {
    azk_context *ctx;
    uint32_t log_n;
    DeviceMem buf, partials, tmpl, state, out, staging;
    PinnedMem hdr_pin, out_pin;
    std::vector<uint64_t> tmpl_host;
    std::mutex mu;
};

namespace {

uint32_t grid_for(uint64_t items) { return (uint32_t)std::min<uint64_t>(kMaxPartials, (items + 2047) / 2048); }

// Everything after the header template is in place.  `hdr_ready` is recorded when the
// header block is uploaded; it is filled in by the caller.
azk_status sc_enqueue_rounds(azk_sumcheck_plan *p, hipStream_t st, const uint32_t *src, bool reduced0)
{
    const uint32_t R = p->log_n;
    ScOut out{(uint32_t *)p->out.p, R};
    const uint64_t *tmpl = (const uint64_t *)p->tmpl.p;
    uint64_t *state = (uint64_t *)p->state.p, *parts = (uint64_t *)p->partials.p;
    if (R <= kTailLog) {
        hipLaunchKernelGGL(k_sc_tail, dim3(1), dim3(kScThreads), 0, st, src, 0u, R, tmpl, state, out);
        HIP_TRY(hipGetLastError());
        return AZK_OK;
    }
    const uint64_t n = 1ull << R;
    uint32_t *bufs[2] = {(uint32_t *)p->buf.p, (uint32_t *)p->buf.p + (n >> 1)};
    const uint32_t *cur = src;
    uint32_t nparts = grid_for(n);
    if (!reduced0) hipLaunchKernelGGL(k_sc_reduce0, dim3(nparts), dim3(kScThreads), 0, st, src, n, parts);
    const uint32_t big = R - kTailLog;
    for (uint32_t r = 0; r < big; ++r) {
        hipLaunchKernelGGL(k_sc_round, dim3(1), dim3(64), 0, st, parts, nparts, r, tmpl, state, out);
        const uint64_t m_out = n >> (r + 1);
        uint32_t *dst = bufs[r & 1];
        const uint32_t g = grid_for(m_out);
        hipLaunchKernelGGL(k_sc_fold, dim3(g), dim3(kScThreads), 0, st, cur, dst, m_out,
                           (const uint32_t *)out.w + 2 * R + r, parts);
        nparts = g;
        cur = dst;
    }
    hipLaunchKernelGGL(k_sc_tail, dim3(1), dim3(kScThreads), 0, st, cur, big, R, tmpl, state, out);
    HIP_TRY(hipGetLastError());
    return AZK_OK;
}

void fill_templates(azk_sumcheck_plan *p)
{
    const uint32_t R = p->log_n;
    p->tmpl_host.assign((size_t)kFrameLanes * (1 + 2 * R), 0);
    for (uint32_t r = 0; r < R; ++r) {
        pad_block(round_frame(r, 0, 0), &p->tmpl_host[(size_t)kFrameLanes * (1 + 2 * r)]);
        pad_block(challenge_frame(r), &p->tmpl_host[(size_t)kFrameLanes * (2 + 2 * r)]);
    }
}

// Runs the program on device words at `src` with the digest from `digest_fn` (called on
// the host after the input reduction is enqueued); fills `res`.
template <class DigestFn>
azk_status sc_run(azk_sumcheck_plan *p, const uint32_t *src, const uint32_t *claimed, DigestFn &&digest_fn,
                  azk_sumcheck_result *res, azk_stats *stats)
{
    const uint32_t R = p->log_n;
    const uint64_t n = 1ull << R;
    hipStream_t st = p->ctx->stream;
    Timer t;
    azk_status s = t.init();
    if (s) return s;
    const uint64_t w0 = now_ns();
    HIP_TRY(hipEventRecord(t.a, st));
    const bool pre = R > kTailLog;
    if (pre) hipLaunchKernelGGL(k_sc_reduce0, dim3(grid_for(n)), dim3(kScThreads), 0, st, src, n, (uint64_t *)p->partials.p);
    uint8_t digest[32];
    s = digest_fn(digest);
    if (s) return s;
    // Header block without the claimed sum (the device inserts it).
    pad_block(header_frame(n, R, 0, digest), (uint64_t *)p->hdr_pin.p);
    HIP_TRY(hipMemcpyAsync(p->tmpl.p, p->hdr_pin.p, 136, hipMemcpyHostToDevice, st));
    s = sc_enqueue_rounds(p, st, src, pre);
    if (s) return s;
    const size_t ob = (size_t)(3 * R + 3) * 4;
    HIP_TRY(hipMemcpyAsync(p->out_pin.p, p->out.p, ob, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipEventRecord(t.b, st));
    const uint64_t w1 = now_ns();
    HIP_TRY(hipStreamSynchronize(st));
    const uint32_t *o = (const uint32_t *)p->out_pin.p;
    if (o[3 * R + 2]) return fail(AZK_ERR_INVALID_LAYOUT, "evaluations are not canonical M31 words (%u >= p)", o[3 * R + 2]);
    if (claimed && *claimed != o[3 * R + 1])
        return fail(AZK_ERR_CORRECTNESS, "Claimed multilinear hypercube sum does not match the revealed evaluation vector.");
    if (res) {
        res->claimed_sum = o[3 * R + 1];
        res->final_evaluation = o[3 * R];
        memcpy(res->initial_digest, digest, 32);
        if (res->round_polys) memcpy(res->round_polys, o, (size_t)2 * R * 4);
        if (res->challenges) memcpy(res->challenges, o + 2 * R, (size_t)R * 4);
    }
    if (stats) {
        stats->submit_ns = w1 - w0;
        stats->wall_ns = now_ns() - w0;
        stats->gpu_ns = t.elapsed_ns();
    }
    return AZK_OK;
}

size_t proof_size(uint32_t log_n)
{
    const uint64_t n = 1ull << log_n;
    return 8 + 4 + 4 + 8 + 4 + 4 + 32 + (8 + 4 * n) + (8 + 8 * (uint64_t)log_n) + (8 + 4 * (uint64_t)log_n) + 4;
}

struct Writer {
    uint8_t *p;
    void u32(uint32_t v) { for (int i = 0; i < 4; ++i) *p++ = (uint8_t)(v >> (8 * i)); }
    void u64(uint64_t v) { for (int i = 0; i < 8; ++i) *p++ = (uint8_t)(v >> (8 * i)); }
    void bytes(const void *s, size_t n) { memcpy(p, s, n); p += n; }
};

// M31MultilinearSumcheckProofCodecV1.encode (SumcheckOracle.swift:1176-1194).
void encode_proof(const uint32_t *ev, uint32_t log_n, const azk_sumcheck_result &r, uint8_t *dst)
{
    const uint64_t n = 1ull << log_n;
    Writer w{dst};
    w.bytes(kMagic, 8);
    w.u32(1);
    w.u32(1);
    w.u64(n);
    w.u32(log_n);
    w.u32(r.claimed_sum);
    w.bytes(r.initial_digest, 32);
    w.u64(n);
    w.bytes(ev, (size_t)n * 4);  // little-endian host
    w.u64(log_n);
    w.bytes(r.round_polys, (size_t)log_n * 8);
    w.u64(log_n);
    w.bytes(r.challenges, (size_t)log_n * 4);
    w.u32(r.final_evaluation);
}

} // namespace

extern "C" {

size_t azk_sumcheck_proof_size(uint32_t log_n) { return (log_n >= 1 && log_n <= kMaxLogN) ? proof_size(log_n) : 0; }

azk_status azk_sumcheck_plan_create(azk_context *ctx, uint32_t log_n, azk_sumcheck_plan **out)
{
    CHECK_ARG(ctx && out);
    if (log_n < 1 || log_n > kMaxLogN) return fail(AZK_ERR_INVALID_LAYOUT, "sum-check lane count 2^%u out of range", log_n);
    auto *p = new azk_sumcheck_plan{ctx, log_n, {}, {}, {}, {}, {}, {}, {}, {}, {}, {}};
    HIP_TRY(hipSetDevice(ctx->device));
    fill_templates(p);
    azk_status s = p->buf.ensure(log_n > kTailLog ? ((size_t)4 << log_n) * 3 / 4 : 64);
    if (!s) s = p->partials.ensure((size_t)kMaxPartials * 24);
    if (!s) s = p->tmpl.ensure(p->tmpl_host.size() * 8);
    if (!s) s = p->state.ensure(200);
    if (!s) s = p->out.ensure((size_t)(3 * log_n + 3) * 4);
    if (!s) s = p->hdr_pin.ensure(136);
    if (!s) s = p->out_pin.ensure((size_t)(3 * log_n + 3) * 4);
    if (!s && hipMemcpy(p->tmpl.p, p->tmpl_host.data(), p->tmpl_host.size() * 8, hipMemcpyHostToDevice) != hipSuccess)
        s = fail(AZK_ERR_COMMAND_FAILED, "template upload failed");
    if (s) { delete p; return s; }
    ctx_retain(p->ctx);
    *out = p;
    return AZK_OK;
}

void azk_sumcheck_plan_destroy(azk_sumcheck_plan *p) { destroy_on_ctx(p); }

azk_status azk_sumcheck_prove_resident(azk_sumcheck_plan *p, const azk_buffer *evals, size_t offset,
                                       const uint8_t *initial_digest, const uint32_t *claimed_sum,
                                       azk_sumcheck_result *res, azk_stats *stats)
{
    CHECK_ARG(p && evals && res);
    std::lock_guard<std::mutex> g(p->mu);
    const size_t bytes = (size_t)4 << p->log_n;
    azk_status s = check_range(evals, offset, bytes);
    if (s) return s;
    if (offset % 4) return fail(AZK_ERR_INVALID_LAYOUT, "evaluation offset must be a multiple of 4");
    HIP_TRY(hipSetDevice(p->ctx->device));
    std::lock_guard<std::mutex> gc(p->ctx->mu);
    const uint32_t *src = (const uint32_t *)((const uint8_t *)evals->ptr + offset);
    return sc_run(p, src, claimed_sum, [&](uint8_t d[32]) -> azk_status {
        if (initial_digest) { memcpy(d, initial_digest, 32); return AZK_OK; }
        // The caller holds no copy: read the words back and hash them here.
        std::vector<uint32_t> host((size_t)1 << p->log_n);
        HIP_TRY(hipMemcpyAsync(host.data(), src, bytes, hipMemcpyDeviceToHost, p->ctx->stream));
        HIP_TRY(hipStreamSynchronize(p->ctx->stream));
        words_digest(host.data(), host.size(), d);
        return AZK_OK;
    }, res, stats);
}

azk_status azk_sumcheck_prove(azk_sumcheck_plan *p, const uint32_t *evals, size_t count, const uint32_t *claimed_sum,
                              uint8_t *proof, size_t capacity, size_t *proof_len, azk_stats *stats)
{
    return azk_sumcheck_prove_with_digest(p, evals, count, nullptr, claimed_sum, proof, capacity, proof_len, stats);
}

azk_status azk_sumcheck_prove_with_digest(azk_sumcheck_plan *p, const uint32_t *evals, size_t count,
                                          const uint8_t *initial_digest, const uint32_t *claimed_sum, uint8_t *proof,
                                          size_t capacity, size_t *proof_len, azk_stats *stats)
{
    CHECK_ARG(p && evals && proof && proof_len);
    std::lock_guard<std::mutex> g(p->mu);
    const uint32_t R = p->log_n;
    if (count != ((size_t)1 << R)) return fail(AZK_ERR_INVALID_LAYOUT, "plan is for 2^%u words, got %zu", R, count);
    const size_t need = proof_size(R);
    if (capacity < need) return fail(AZK_ERR_INVALID_LAYOUT, "proof capacity %zu < %zu", capacity, need);
    HIP_TRY(hipSetDevice(p->ctx->device));
    std::lock_guard<std::mutex> gc(p->ctx->mu);
    azk_status s = p->staging.ensure(count * 4);  // device copy of the host words
    if (s) return s;
    // The digest (a serial sponge over the words) runs on a second host thread while
    // this one uploads and enqueues the input reduction.
    uint8_t digest[32];
    std::thread hasher;
    const bool threaded = !initial_digest && R >= 14;
    if (initial_digest) memcpy(digest, initial_digest, 32);
    else if (threaded) hasher = std::thread([&] { words_digest(evals, count, digest); });
    else words_digest(evals, count, digest);
    hipError_t e = hipMemcpyAsync(p->staging.p, evals, count * 4, hipMemcpyHostToDevice, p->ctx->stream);
    std::vector<uint32_t> polys((size_t)2 * R), ch(R);
    azk_sumcheck_result r{};
    r.round_polys = polys.data();
    r.challenges = ch.data();
    if (e == hipSuccess)
        s = sc_run(p, (const uint32_t *)p->staging.p, claimed_sum, [&](uint8_t d[32]) -> azk_status {
            if (threaded) hasher.join();
            memcpy(d, digest, 32);
            return AZK_OK;
        }, &r, stats);
    if (hasher.joinable()) hasher.join();
    if (e != hipSuccess) return fail(AZK_ERR_COMMAND_FAILED, "upload: %s", hipGetErrorString(e));
    if (s) return s;
    encode_proof(evals, R, r, proof);
    *proof_len = need;
    return AZK_OK;
}

// M31MultilinearSumcheckProofCodecV1.decode + M31MultilinearSumcheckVerifierV1.verify
// (SumcheckOracle.swift:941-1064,1196-1233) against the statement carried in the proof.
azk_status azk_sumcheck_verify(const uint8_t *proof, size_t len, int *accepted)
{
    CHECK_ARG(proof && accepted);
    *accepted = 0;
    size_t at = 0;
    auto need = [&](size_t k) { return at + k <= len; };
    auto rd32 = [&]() { uint32_t v = 0; for (int i = 0; i < 4; ++i) v |= (uint32_t)proof[at++] << (8 * i); return v; };
    auto rd64 = [&]() { uint64_t v = 0; for (int i = 0; i < 8; ++i) v |= (uint64_t)proof[at++] << (8 * i); return v; };
    auto bad = [&](const char *what) { return fail(AZK_ERR_INVALID_LAYOUT, "sum-check proof: %s", what); };
    if (!need(8 + 4 + 4 + 8 + 4 + 4 + 32 + 8) || memcmp(proof, kMagic, 8)) return bad("magic");
    at = 8;
    const uint32_t pver = rd32(), sver = rd32();
    const uint64_t n = rd64();
    const uint32_t rounds = rd32(), claimed = rd32();
    uint8_t digest[32];
    memcpy(digest, proof + at, 32);
    at += 32;
    if (pver != 1 || sver != 1 || n < 2 || (n & (n - 1)) || n > (1ull << kMaxLogN) || rounds != ilog2(n) || claimed >= kP)
        return bad("statement");
    if (rd64() != n || !need(n * 4)) return bad("evaluation count");
    std::vector<uint32_t> cur(n);
    memcpy(cur.data(), proof + at, n * 4);
    at += n * 4;
    for (uint32_t v : cur) if (v >= kP) return bad("non-canonical evaluation");
    if (!need(8) || rd64() != rounds || !need((size_t)8 * rounds)) return bad("round count");
    std::vector<uint32_t> polys((size_t)2 * rounds);
    for (auto &v : polys) v = rd32();
    if (!need(8) || rd64() != rounds || !need((size_t)4 * rounds + 4)) return bad("challenge count");
    std::vector<uint32_t> ch(rounds);
    for (auto &v : ch) v = rd32();
    const uint32_t fin = rd32();
    if (at != len) return bad("trailing bytes");
    for (uint32_t v : polys) if (v >= kP) return bad("non-canonical round polynomial");
    for (uint32_t v : ch) if (v >= kP) return bad("non-canonical challenge");
    if (fin >= kP) return bad("non-canonical final evaluation");

    uint8_t d2[32];
    words_digest(cur.data(), n, d2);
    bool ok = memcmp(d2, digest, 32) == 0;
    host::Transcript t;
    auto hf = header_frame(n, rounds, claimed, digest);
    t.absorb(hf.data(), hf.size());
    uint32_t prev = claimed;
    for (uint32_t r = 0; r < rounds && ok; ++r) {
        uint32_t e0 = 0, e1 = 0;
        for (size_t i = 0; i < cur.size() / 2; ++i) { e0 = hadd(e0, cur[2 * i]); e1 = hadd(e1, cur[2 * i + 1]); }
        if (e0 != polys[2 * r] || e1 != polys[2 * r + 1] || hadd(e0, e1) != prev) { ok = false; break; }
        auto rf = round_frame(r, e0, e1);
        t.absorb(rf.data(), rf.size());
        auto cf = challenge_frame(r);
        t.absorb(cf.data(), cf.size());
        uint32_t c;
        t.squeeze_u32(1, kP, &c);
        if (c != ch[r]) ok = false;
        prev = hadd(hmul(hsub(1, ch[r]), e0), hmul(ch[r], e1));
        for (size_t i = 0; i < cur.size() / 2; ++i)
            cur[i] = hadd(hmul(hsub(1, ch[r]), cur[2 * i]), hmul(ch[r], cur[2 * i + 1]));
        cur.resize(cur.size() / 2);
    }
    ok = ok && cur.size() == 1 && fin == cur[0] && fin == prev;
    *accepted = ok ? 1 : 0;
    return AZK_OK;
}

} // extern "C"
