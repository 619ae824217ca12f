// Monomial -> Circle-FFT-basis conversion on gfx950: the witness transform that feeds the
// commit path's FFT-basis input.
//
// Reference: CircleCodewordOracle.circleFFTCoefficients (CircleCodeword.swift:100-127),
// lineFFTBasisCoefficients (:187-217), multiplyChebyshevByX (:219-236) and
// chebyshevToLineBasis (:238-262), O(n^2) on the CPU; the resident plan
// CircleWitnessToFFTBasisPlanV1 (CircleWitnessFFTBasis.swift:230-517) multiplies by a dense
// M31 matrix and is capped at 4096 coefficients (:91).
//
// Algorithm, O(n log^2 n).  The line basis of capacity N = 2^L is
//   b_i(x) = prod over the set bits k of i of T_{2^k}(x)
// (chebyshevToLineBasis writes c = low + T_{N/2} high and recurses).  In the doubled
// variable y = 2x, p(x) = P(y) with P_k = m_k 2^-k, and y^2 = 2 T_2(x) + 2 = y' + 2 where
// y' = 2 T_2(x) is the doubled variable one level down.  Hence
//   P(y) = E(y' + 2) + y O(y' + 2)        (E, O: even / odd coefficients of P)
// and the even (odd) line-basis coefficients of p are those of E(t + 2) (O(t + 2)) one
// level down.  In the doubled basis D_i = prod (2 T_{2^k})^{i_k} = 2^popcount(i) b_i the
// recursion is exact, it ends at length 1, and
//   line[i] = 2^popcount(i) * (value at storage position bitrev_L(i)).
// The Taylor shift g(t) = f(t + 2) of a length-H f is one correlation,
//   g_j = (1/j!) sum_m F_{j+m} R_m,   F_k = f_k k!,   R_m = 2^m / m!,
// done as a cyclic convolution of size S = 2H over CM31 = F_p[i]: the circle group's
// elements x + iy have every order 2^k <= 2^31, so CM31 has the roots of unity.  Forward
// DIF of the zero-padded F (natural in, bit-reversed out), times K = DIF(reversed R) / S
// (precomputed per size), inverse DIT (bit-reversed in, natural out).  A QM31 value is two
// CM31 lanes and K is the transform of a real sequence, so both lanes share it.  Short
// shifts (H <= 32) use the sum directly.  All of it is exact arithmetic in F_p, so the
// canonical outputs equal the reference's.
//
// Device program for one conversion (x and y families together):
//   levels with block B > 2^kLogChunk: k_bs_fwd passes (the first gathers the level's
//     even/odd halves from the level array, scales them, and does the first DIF stage),
//     k_bs_chunk (last DIF stages, times K, first DIT stages), k_bs_inv passes (the last
//     writes g_j / j! into the next level array);
//   the level with B = 2^kLogChunk: k_bs_chunk alone (gather ... scatter in one kernel);
//   blocks of B <= 2^kLogTail: k_bs_tail runs every remaining level in LDS and writes
//     out[2i + family] = 2^popcount(i) D[bitrev(i)].
#include "device_field.hpp"
#include "host_gates.hpp"
#include "runtime_common.hpp"

using namespace azk;
using namespace azk::rt;

namespace {

constexpr uint32_t kBsThreads = 256;
constexpr uint32_t kLogChunk = 11;     // LDS chunk of one transform (32 KB)
constexpr uint32_t kLogTail = 10;      // tail block (data + two transforms: 48 KB)
constexpr uint32_t kLogPassTile = 11;  // global pass tile (rows x columns)
constexpr uint32_t kLogMinCols = 3;    // >= 8 contiguous QM31 (128 B) per tile row
constexpr uint32_t kDirectMax = 32;    // shifts of length <= this are direct sums
constexpr uint32_t kMaxLogN = 26;
// the fused bottom pass (stages 2..0) sits below the gathered first stage: chunks of >= 16
// and tail transforms of >= 16 (H > kDirectMax >= 4)
static_assert(kLogChunk >= 4 && kDirectMax >= 4 && kLogTail < kLogChunk, "stage layout");

enum : uint32_t { kPrepare = 1, kFinalize = 2, kForwardOnly = 4 };

struct BsArgs {
    const uint4 *src[2];    // level input per family (the caller's x / y at level 0)
    uint32_t count[2];      // level 0: valid input entries per family
    uint4 *dst;             // level output: family f at dst + f * N
    uint4 *w;               // transforms of size S, transform t at w + t * S
    const uint2 *tw;        // tw[h + j] = w_{2h}^j (CM31), h = 2^r
    const uint2 *kh;        // kh[S + k] = K_S[k]
    const uint32_t *fact;   // k!
    const uint32_t *ifact;  // 1 / k!
    const uint32_t *rt;     // 2^m / m!
    uint32_t *flag;         // set when an input limb is not canonical
    uint32_t log_n;         // L (capacity N = 2^L per family)
    uint32_t log_s;         // block / transform size of the level
    uint32_t lvl;           // level (2^lvl blocks per family)
    uint32_t lt, lb, log_c; // global pass: stages h = 2^lt .. 2^lb, 2^log_c columns per tile
    uint32_t mode;
};

__device__ __forceinline__ uint4 q_add(uint4 x, uint4 y)
{
    return {m31_add(x.x, y.x), m31_add(x.y, y.y), m31_add(x.z, y.z), m31_add(x.w, y.w)};
}
__device__ __forceinline__ uint4 q_sub(uint4 x, uint4 y)
{
    return {m31_sub(x.x, y.x), m31_sub(x.y, y.y), m31_sub(x.z, y.z), m31_sub(x.w, y.w)};
}
__device__ __forceinline__ uint4 q_scale(uint4 x, uint32_t s)
{
    return {m31_mul(x.x, s), m31_mul(x.y, s), m31_mul(x.z, s), m31_mul(x.w, s)};
}
// both CM31 lanes of x times w
__device__ __forceinline__ uint4 q_cmul(uint4 x, uint2 w)
{
    const uint32_t w1n = m31_neg(w.y);
    uint4 r;
    cm_mul_lazy(x.x, x.y, w.x, w.y, w1n, r.x, r.y);
    cm_mul_lazy(x.z, x.w, w.x, w.y, w1n, r.z, r.w);
    return r;
}
// ... times conj(w) = w^-1 for a root of unity: conj(w) = w0 + (p - w1) i
__device__ __forceinline__ uint4 q_cmul_conj(uint4 x, uint2 w)
{
    const uint32_t w1n = m31_neg(w.y);
    uint4 r;
    cm_mul_lazy(x.x, x.y, w.x, w1n, w.y, r.x, r.y);
    cm_mul_lazy(x.z, x.w, w.x, w1n, w.y, r.z, r.w);
    return r;
}

// x * 2^e for canonical x, 0 <= e < 31: a rotation of the 31-bit value (2^31 = 1 mod p).
__device__ __forceinline__ uint32_t m31_rotl(uint32_t x, uint32_t e)
{
    return e ? (((x << e) & M31_P) | (x >> (31 - e))) : x;
}
__device__ __forceinline__ uint4 q_rotl(uint4 v, uint32_t e)
{
    return {m31_rotl(v.x, e), m31_rotl(v.y, e), m31_rotl(v.z, e), m31_rotl(v.w, e)};
}
// acc < 2^40 -> acc mod p
__device__ __forceinline__ uint32_t fold40(uint64_t acc)
{
    const uint32_t r = (uint32_t)(acc & M31_P) + (uint32_t)(acc >> 31);  // < 2^31 + 2^9
    return r >= M31_P ? r - M31_P : r;
}
__device__ __forceinline__ uint32_t inv_pow2_exp(uint32_t n) { return (31u - n % 31u) % 31u; } // 2^-n = 2^this

// Entry n of block b (of 2^log_s entries) of family f at the level's source.  Level 0
// reads the caller's arrays: zero past the count, canonical check, times 2^-n.
__device__ __forceinline__ uint4 load_entry(const BsArgs &a, uint32_t f, uint32_t b, uint32_t n)
{
    if (a.lvl == 0) {
        if (n >= a.count[f]) return make_uint4(0, 0, 0, 0);
        uint4 v = a.src[f][n];
        if ((v.x >= M31_P) | (v.y >= M31_P) | (v.z >= M31_P) | (v.w >= M31_P)) {
            atomicOr(a.flag, 1u);
            return make_uint4(0, 0, 0, 0);
        }
        return q_rotl(v, inv_pow2_exp(n));
    }
    return a.src[f][((size_t)b << a.log_s) + n];
}

// F_k of transform t = (f, b, sigma): entry 2k + sigma of block b, times k!.
__device__ __forceinline__ uint4 gather_half(const BsArgs &a, uint32_t t, uint32_t k)
{
    const uint32_t f = t >> (a.lvl + 1), b = (t >> 1) & ((1u << a.lvl) - 1), sigma = t & 1;
    return q_scale(load_entry(a, f, b, 2 * k + sigma), a.fact[k]);
}

// g_j / j! of transform t into the next level array: block 2b + sigma, entry j.
__device__ __forceinline__ void scatter_half(const BsArgs &a, uint32_t t, uint32_t j, uint4 g)
{
    const uint32_t f = t >> (a.lvl + 1), bs = t & ((2u << a.lvl) - 1);
    a.dst[((size_t)f << a.log_n) + ((size_t)bs << (a.log_s - 1)) + j] = q_scale(g, a.ifact[j]);
}

// ---------------------------------------------------------------- radix-2^K stage groups
// A DIF group covers K consecutive stages, half-sizes (R/2) q, ..., 2q, q (R = 2^K), on
// blocks of R q positions: for j < q it takes x_m = X[j + m q], runs the R-point DFT with
// its internal factors (+-1, +-i and powers of w8; for the generator (2, 1268011823)
// w4 = -i and w8 = 2^15 (1 - i)), and multiplies slot m by w_{Rq}^(j rev_K(m)).  That is
// the same map as K radix-2 DIF stages, with R - 1 general products per R elements instead
// of K R / 2 and one LDS round trip instead of K.  The DIT group is its inverse: conjugate
// external factors first, then the internal stages in reverse order.

__device__ __forceinline__ void cm_negi(uint32_t &a, uint32_t &b) { const uint32_t t = a; a = b; b = m31_neg(t); }  // * -i
__device__ __forceinline__ void cm_posi(uint32_t &a, uint32_t &b) { const uint32_t t = a; a = m31_neg(b); b = t; }  // * i
// * w8 = 2^15 (1 - i):        2^15 ((a + b) + (b - a) i)
__device__ __forceinline__ void cm_w8(uint32_t &a, uint32_t &b)
{
    const uint32_t s = m31_add(a, b), d = m31_sub(b, a);
    a = m31_rotl(s, 15);
    b = m31_rotl(d, 15);
}
// * w8^3 = 2^15 (-1 - i):     2^15 ((b - a) - (a + b) i)
__device__ __forceinline__ void cm_w8_3(uint32_t &a, uint32_t &b)
{
    const uint32_t s = m31_add(a, b), d = m31_sub(b, a);
    a = m31_rotl(d, 15);
    b = m31_rotl(m31_neg(s), 15);
}
// * conj(w8) = 2^15 (1 + i):  2^15 ((a - b) + (a + b) i)
__device__ __forceinline__ void cm_w8c(uint32_t &a, uint32_t &b)
{
    const uint32_t s = m31_add(a, b), d = m31_sub(a, b);
    a = m31_rotl(d, 15);
    b = m31_rotl(s, 15);
}
// * conj(w8^3) = 2^15 (-1 + i): 2^15 (-(a + b) + (a - b) i)
__device__ __forceinline__ void cm_w8c_3(uint32_t &a, uint32_t &b)
{
    const uint32_t s = m31_add(a, b), d = m31_sub(a, b);
    a = m31_rotl(m31_neg(s), 15);
    b = m31_rotl(d, 15);
}
#define AZK_LANES(op, v) do { op((v).x, (v).y); op((v).z, (v).w); } while (0)

// x * w_{2h}^e (or its conjugate) for e < 2h; the table holds w_{2h}^t for t < h and
// w_{2h}^(t + h) = -w_{2h}^t.
template <bool CONJ>
__device__ __forceinline__ uint4 q_mul_pow(uint4 x, const uint2 *tw, uint32_t h, uint32_t e)
{
    const bool neg = e >= h;
    const uint2 w = tw[h + (neg ? e - h : e)];
    const uint32_t w0 = neg ? m31_neg(w.x) : w.x, w1 = neg ? m31_neg(w.y) : w.y;
    return CONJ ? q_cmul_conj(x, make_uint2(w0, w1)) : q_cmul(x, make_uint2(w0, w1));
}

__device__ __forceinline__ void bfly(uint4 &a, uint4 &c)
{
    const uint4 s = q_add(a, c), d = q_sub(a, c);
    a = s;
    c = d;
}

template <int K>
__device__ __forceinline__ uint32_t rev_slot(int m)
{
    return K == 3 ? ((m & 1) << 2 | (m & 2) | (m >> 2)) : K == 2 ? ((m & 1) << 1 | (m >> 1)) : (uint32_t)m;
}

// K DIF stages on the R slots; EXT = false when q = 1 (every external factor is 1).
template <int K, bool EXT>
__device__ __forceinline__ void dif_core(uint4 (&x)[1 << K], uint32_t j, uint32_t log_q, const uint2 *tw)
{
    constexpr int R = 1 << K;
    if constexpr (K == 3) {
#pragma unroll
        for (int m = 0; m < 4; ++m) bfly(x[m], x[m + 4]);
        AZK_LANES(cm_w8, x[5]);
        AZK_LANES(cm_negi, x[6]);
        AZK_LANES(cm_w8_3, x[7]);
    }
    if constexpr (K >= 2) {
#pragma unroll
        for (int m = 0; m < R; ++m)
            if (!(m & 2)) {
                bfly(x[m], x[m + 2]);
                if (m & 1) AZK_LANES(cm_negi, x[m + 2]);
            }
    }
#pragma unroll
    for (int m = 0; m < R; m += 2) bfly(x[m], x[m + 1]);
    if constexpr (EXT) {
        const uint32_t h = 1u << (log_q + K - 1);
#pragma unroll
        for (int m = 1; m < R; ++m) x[m] = q_mul_pow<false>(x[m], tw, h, j * rev_slot<K>(m));
    }
}

template <int K, bool EXT>
__device__ __forceinline__ void dit_core(uint4 (&x)[1 << K], uint32_t j, uint32_t log_q, const uint2 *tw)
{
    constexpr int R = 1 << K;
    if constexpr (EXT) {
        const uint32_t h = 1u << (log_q + K - 1);
#pragma unroll
        for (int m = 1; m < R; ++m) x[m] = q_mul_pow<true>(x[m], tw, h, j * rev_slot<K>(m));
    }
#pragma unroll
    for (int m = 0; m < R; m += 2) bfly(x[m], x[m + 1]);
    if constexpr (K >= 2) {
#pragma unroll
        for (int m = 0; m < R; ++m)
            if (!(m & 2)) {
                if (m & 1) AZK_LANES(cm_posi, x[m + 2]);
                bfly(x[m], x[m + 2]);
            }
    }
    if constexpr (K == 3) {
        AZK_LANES(cm_w8c, x[5]);
        AZK_LANES(cm_posi, x[6]);
        AZK_LANES(cm_w8c_3, x[7]);
#pragma unroll
        for (int m = 0; m < 4; ++m) bfly(x[m], x[m + 4]);
    }
}

// One group of K stages over n LDS elements: slots at LDS distance 2^l_lds, position q =
// 2^log_q; jpos(i0) = (position of LDS element i0) mod q.
template <int K, bool INV, class J>
__device__ __forceinline__ void lds_group(uint4 *sh, uint32_t n, uint32_t l_lds, uint32_t log_q, const uint2 *tw, J &&jpos)
{
    for (uint32_t g = threadIdx.x; g < (n >> K); g += kBsThreads) {
        const uint32_t i0 = ((g >> l_lds) << (l_lds + K)) | (g & ((1u << l_lds) - 1));
        uint4 x[1 << K];
#pragma unroll
        for (int m = 0; m < (1 << K); ++m) x[m] = sh[i0 + ((uint32_t)m << l_lds)];
        const uint32_t j = jpos(i0);
        if constexpr (INV) dit_core<K, true>(x, j, log_q, tw);
        else dif_core<K, true>(x, j, log_q, tw);
#pragma unroll
        for (int m = 0; m < (1 << K); ++m) sh[i0 + ((uint32_t)m << l_lds)] = x[m];
    }
    __syncthreads();
}

template <bool INV, class J>
__device__ __forceinline__ void lds_group_k(uint32_t k, uint4 *sh, uint32_t n, uint32_t l_lds, uint32_t log_q,
                                            const uint2 *tw, J &&jpos)
{
    if (k == 3) lds_group<3, INV>(sh, n, l_lds, log_q, tw, jpos);
    else if (k == 2) lds_group<2, INV>(sh, n, l_lds, log_q, tw, jpos);
    else lds_group<1, INV>(sh, n, l_lds, log_q, tw, jpos);
}

// DIF stages (position logs) hi .. lo over n LDS elements, top group takes the remainder;
// LDS log distance of position log l is l + d_off.
template <class J>
__device__ __forceinline__ void dif_stages(uint4 *sh, uint32_t n, int hi, int lo, int d_off, const uint2 *tw, J &&jpos_of)
{
    if (hi < lo) return;
    const uint32_t count = hi - lo + 1, rem = count % 3;
    int top = hi;
    if (rem) {
        const uint32_t lq = top - rem + 1;
        lds_group_k<false>(rem, sh, n, lq + d_off, lq, tw, jpos_of(lq));
        top -= rem;
    }
    for (; top >= lo + 2; top -= 3) lds_group<3, false>(sh, n, top - 2 + d_off, top - 2, tw, jpos_of(top - 2));
}

// DIT stages lo .. hi, bottom-up, the remainder last.
template <class J>
__device__ __forceinline__ void dit_stages(uint4 *sh, uint32_t n, int lo, int hi, int d_off, const uint2 *tw, J &&jpos_of)
{
    if (hi < lo) return;
    const uint32_t count = hi - lo + 1, rem = count % 3;
    int bot = lo;
    for (; bot + 2 <= hi - (int)rem; bot += 3) lds_group<3, true>(sh, n, bot + d_off, bot, tw, jpos_of(bot));
    if (rem) lds_group_k<true>(rem, sh, n, bot + d_off, bot, tw, jpos_of(bot));
}

// Stages 2, 1, 0 (DIF) of transforms laid out contiguously, the pointwise product with
// K (kv(i) = K at LDS index i), and DIT stages 0, 1, 2: one register pass.
template <class KV>
__device__ __forceinline__ void fused_bottom(uint4 *sh, uint32_t n, KV &&kv)
{
    for (uint32_t g = threadIdx.x; g < (n >> 3); g += kBsThreads) {
        const uint32_t i0 = g << 3;
        uint4 x[8];
#pragma unroll
        for (int m = 0; m < 8; ++m) x[m] = sh[i0 + m];
        dif_core<3, false>(x, 0, 0, nullptr);
#pragma unroll
        for (int m = 0; m < 8; ++m) x[m] = q_cmul(x[m], kv(i0 + m));
        dit_core<3, false>(x, 0, 0, nullptr);
#pragma unroll
        for (int m = 0; m < 8; ++m) sh[i0 + m] = x[m];
    }
    __syncthreads();
}

// Forward DIF passes over stages h = 2^lt .. 2^lb (lb >= kLogChunk) of transforms of size
// 2^log_s: a tile is 2^(lt-lb+1) rows at stride 2^lb times 2^log_c contiguous columns.
// kPrepare (lt = log_s - 1): rows of the upper half are the zero padding, so the first
// stage is (F, F w) and the lower rows come from the level array.
__global__ __launch_bounds__(kBsThreads) void k_bs_fwd(BsArgs a)
{
    extern __shared__ uint4 sh[];
    const uint32_t log_r = a.lt - a.lb + 1, log_tile = log_r + a.log_c;
    const uint32_t tile = 1u << log_tile, C = 1u << a.log_c;
    const uint32_t log_tiles = a.log_s - log_tile;
    const uint32_t t = blockIdx.x >> log_tiles, g = blockIdx.x & ((1u << log_tiles) - 1);
    const uint32_t log_groups = a.lb - a.log_c;  // column groups per 2^(lt+1) block
    const size_t base = ((size_t)(g >> log_groups) << (a.lt + 1)) + ((size_t)(g & ((1u << log_groups) - 1)) << a.log_c);
    const uint32_t j0 = (uint32_t)(base & ((1u << a.lb) - 1));
    uint4 *x = a.w + ((size_t)t << a.log_s);
    int l_first = (int)a.lt;
    if (a.mode & kPrepare) {
        // h = 2^lt = S/2: base < h (one block of 2h per transform), p = r * 2^lb + j0 + c < H
        const uint32_t half = tile >> 1, H = 1u << a.lt;
        for (uint32_t e = threadIdx.x; e < half; e += kBsThreads) {
            const uint32_t r = e >> a.log_c, c = e & (C - 1);
            const uint32_t p = (r << a.lb) + j0 + c;
            uint4 v = gather_half(a, t, p);
            sh[e] = v;
            sh[e + half] = q_cmul(v, a.tw[H + p]);
        }
        l_first = (int)a.lt - 1;
    } else {
        for (uint32_t e = threadIdx.x; e < tile; e += kBsThreads) {
            const uint32_t r = e >> a.log_c, c = e & (C - 1);
            sh[e] = x[base + ((size_t)r << a.lb) + c];
        }
    }
    __syncthreads();
    const uint32_t lb = a.lb, log_c = a.log_c;
    dif_stages(sh, tile, l_first, (int)lb, (int)log_c - (int)lb, a.tw, [&](uint32_t lq) {
        return [=](uint32_t i0) { return (((i0 >> log_c) & ((1u << (lq - lb)) - 1)) << lb) + j0 + (i0 & (C - 1)); };
    });
    for (uint32_t e = threadIdx.x; e < tile; e += kBsThreads) {
        const uint32_t r = e >> a.log_c, c = e & (C - 1);
        x[base + ((size_t)r << a.lb) + c] = sh[e];
    }
}

// Inverse DIT passes over stages h = 2^lb .. 2^lt; kFinalize (lt = log_s - 1): the last
// stage only forms the lower half, g_j for j < H, and writes g_j / j! to the next level.
__global__ __launch_bounds__(kBsThreads) void k_bs_inv(BsArgs a)
{
    extern __shared__ uint4 sh[];
    const uint32_t log_r = a.lt - a.lb + 1, log_tile = log_r + a.log_c;
    const uint32_t tile = 1u << log_tile, C = 1u << a.log_c;
    const uint32_t log_tiles = a.log_s - log_tile;
    const uint32_t t = blockIdx.x >> log_tiles, g = blockIdx.x & ((1u << log_tiles) - 1);
    const uint32_t log_groups = a.lb - a.log_c;
    const size_t base = ((size_t)(g >> log_groups) << (a.lt + 1)) + ((size_t)(g & ((1u << log_groups) - 1)) << a.log_c);
    const uint32_t j0 = (uint32_t)(base & ((1u << a.lb) - 1));
    uint4 *x = a.w + ((size_t)t << a.log_s);
    for (uint32_t e = threadIdx.x; e < tile; e += kBsThreads) {
        const uint32_t r = e >> a.log_c, c = e & (C - 1);
        sh[e] = x[base + ((size_t)r << a.lb) + c];
    }
    __syncthreads();
    const int l_last = (a.mode & kFinalize) ? (int)a.lt - 1 : (int)a.lt;
    const uint32_t lb = a.lb, log_c = a.log_c;
    dit_stages(sh, tile, (int)lb, l_last, (int)log_c - (int)lb, a.tw, [&](uint32_t lq) {
        return [=](uint32_t i0) { return (((i0 >> log_c) & ((1u << (lq - lb)) - 1)) << lb) + j0 + (i0 & (C - 1)); };
    });
    if (a.mode & kFinalize) {
        const uint32_t half = tile >> 1, H = 1u << a.lt;
        for (uint32_t e = threadIdx.x; e < half; e += kBsThreads) {
            const uint32_t r = e >> a.log_c, c = e & (C - 1);
            const uint32_t p = (r << a.lb) + j0 + c;
            scatter_half(a, t, p, q_add(sh[e], q_cmul_conj(sh[e + half], a.tw[H + p])));
        }
        return;
    }
    for (uint32_t e = threadIdx.x; e < tile; e += kBsThreads) {
        const uint32_t r = e >> a.log_c, c = e & (C - 1);
        x[base + ((size_t)r << a.lb) + c] = sh[e];
    }
}

// One chunk of 2^min(log_s, kLogChunk) contiguous elements of one transform: the DIF
// stages below the chunk size, times K, the DIT stages below the chunk size.  When the
// chunk is the whole transform (log_s <= kLogChunk), kPrepare / kFinalize gather from and
// scatter to the level arrays.  kForwardOnly stops after the DIF (used to build K).
__global__ __launch_bounds__(kBsThreads) void k_bs_chunk(BsArgs a)
{
    extern __shared__ uint4 sh[];
    const uint32_t log_tc = a.log_s < kLogChunk ? a.log_s : kLogChunk;
    const uint32_t tc = 1u << log_tc, log_chunks = a.log_s - log_tc;
    const uint32_t t = blockIdx.x >> log_chunks, c = blockIdx.x & ((1u << log_chunks) - 1);
    const size_t pbase = (size_t)c << log_tc;
    uint4 *x = a.w + ((size_t)t << a.log_s) + pbase;
    int l_first = (int)log_tc - 1;
    if (a.mode & kPrepare) {
        const uint32_t H = tc >> 1;
        for (uint32_t k = threadIdx.x; k < H; k += kBsThreads) {
            uint4 v = gather_half(a, t, k);
            sh[k] = v;
            sh[k + H] = q_cmul(v, a.tw[H + k]);
        }
        l_first = (int)log_tc - 2;
    } else {
        for (uint32_t e = threadIdx.x; e < tc; e += kBsThreads) sh[e] = x[e];
    }
    __syncthreads();
    auto jpos = [](uint32_t lq) { return [=](uint32_t i0) { return i0 & ((1u << lq) - 1); }; };
    if (a.mode & kForwardOnly) {
        dif_stages(sh, tc, l_first, 0, 0, a.tw, jpos);
        for (uint32_t e = threadIdx.x; e < tc; e += kBsThreads) x[e] = sh[e];
        return;
    }
    dif_stages(sh, tc, l_first, 3, 0, a.tw, jpos);
    const uint2 *k = a.kh + ((size_t)1 << a.log_s) + pbase;
    fused_bottom(sh, tc, [&](uint32_t i) { return k[i]; });
    const int l_last = (a.mode & kFinalize) ? (int)log_tc - 2 : (int)log_tc - 1;
    dit_stages(sh, tc, 3, l_last, 0, a.tw, jpos);
    if (a.mode & kFinalize) {
        const uint32_t H = tc >> 1;
        for (uint32_t j = threadIdx.x; j < H; j += kBsThreads)
            scatter_half(a, t, j, q_add(sh[j], q_cmul_conj(sh[j + H], a.tw[H + j])));
        return;
    }
    for (uint32_t e = threadIdx.x; e < tc; e += kBsThreads) x[e] = sh[e];
}

// Every remaining level of one block of B = 2^log_s <= 2^kLogTail entries (family f,
// block b of 2^lvl), in LDS: D holds the block, W two transforms per sub-block.  Output
// out[2i + f] = 2^popcount(i) D[s] with i = bitrev_L(b B + s).
__global__ __launch_bounds__(kBsThreads) void k_bs_tail(BsArgs a, uint4 *__restrict__ out)
{
    extern __shared__ uint4 sh[];
    const uint32_t B = 1u << a.log_s;
    uint4 *D = sh, *W = sh + B;
    const uint32_t f = blockIdx.x >> a.lvl, b = blockIdx.x & ((1u << a.lvl) - 1);
    for (uint32_t n = threadIdx.x; n < B; n += kBsThreads) D[n] = load_entry(a, f, b, n);
    __syncthreads();
    auto jpos = [](uint32_t lq) { return [=](uint32_t i0) { return i0 & ((1u << lq) - 1); }; };
    for (uint32_t lb = a.log_s; lb >= 2; --lb) {  // sub-blocks of 2^lb, halves of H
        const uint32_t Bc = 1u << lb, H = Bc >> 1;
        if (H <= kDirectMax) {
            // g_j = (1/j!) sum_m (f_{j+m} (j+m)!) 2^m / m!, f_k = D[sub + 2k + sigma]
            for (uint32_t n = threadIdx.x; n < B; n += kBsThreads) D[n] = q_scale(D[n], a.fact[(n & (Bc - 1)) >> 1]);
            __syncthreads();
            for (uint32_t o = threadIdx.x; o < B; o += kBsThreads) {
                const uint32_t sub = o & ~(Bc - 1), sigma = (o >> (lb - 1)) & 1, j = o & (H - 1);
                uint64_t acc0 = 0, acc1 = 0, acc2 = 0, acc3 = 0;
                for (uint32_t m = 0; m + j < H; ++m) {
                    const uint4 v = D[sub + 2 * (j + m) + sigma];
                    const uint32_t r = a.rt[m];
                    // each folded product is < 2^32, so the <= 32-term sums stay below 2^37
                    uint64_t t0 = (uint64_t)v.x * r, t1 = (uint64_t)v.y * r, t2 = (uint64_t)v.z * r,
                             t3 = (uint64_t)v.w * r;
                    acc0 += (t0 & M31_P) + (t0 >> 31);
                    acc1 += (t1 & M31_P) + (t1 >> 31);
                    acc2 += (t2 & M31_P) + (t2 >> 31);
                    acc3 += (t3 & M31_P) + (t3 >> 31);
                }
                const uint4 g = {fold40(acc0), fold40(acc1), fold40(acc2), fold40(acc3)};
                W[o] = q_scale(g, a.ifact[j]);
            }
            __syncthreads();
            for (uint32_t n = threadIdx.x; n < B; n += kBsThreads) D[n] = W[n];
            __syncthreads();
            continue;
        }
        // transforms t' = 2 * sub-block + sigma of size Bc at W + t' Bc (2B elements)
        for (uint32_t e = threadIdx.x; e < B; e += kBsThreads) {
            const uint32_t tp = e >> (lb - 1), k = e & (H - 1);
            const uint32_t sub = (tp >> 1) << lb, sigma = tp & 1;
            uint4 v = q_scale(D[sub + 2 * k + sigma], a.fact[k]);
            W[(tp << lb) + k] = v;
            W[(tp << lb) + H + k] = q_cmul(v, a.tw[H + k]);
        }
        __syncthreads();
        dif_stages(W, 2 * B, (int)lb - 2, 3, 0, a.tw, jpos);
        const uint2 *k = a.kh + Bc;
        fused_bottom(W, 2 * B, [&](uint32_t i) { return k[i & (Bc - 1)]; });
        dit_stages(W, 2 * B, 3, (int)lb - 2, 0, a.tw, jpos);
        for (uint32_t e = threadIdx.x; e < B; e += kBsThreads) {
            const uint32_t tp = e >> (lb - 1), j = e & (H - 1);
            const uint4 g = q_add(W[(tp << lb) + j], q_cmul_conj(W[(tp << lb) + H + j], a.tw[H + j]));
            D[((tp >> 1) << lb) + ((tp & 1) << (lb - 1)) + j] = q_scale(g, a.ifact[j]);
        }
        __syncthreads();
    }
    for (uint32_t s = threadIdx.x; s < B; s += kBsThreads) {
        const uint32_t gs = (b << a.log_s) + s;
        const uint32_t i = a.log_n ? __brev(gs) >> (32 - a.log_n) : 0;
        out[2 * (size_t)i + f] = q_rotl(D[s], __popc(i));
    }
}

// K_S input: (reversed R) / S as real values in lane 0 of w[0, S), R_m = 2^m / m! for m < S/2.
__global__ __launch_bounds__(kBsThreads) void k_bs_kernel_init(uint4 *w, const uint32_t *rt, uint32_t log_s)
{
    const uint32_t S = 1u << log_s, H = S >> 1;
    const uint32_t inv_s = 1u << ((31u - log_s % 31u) % 31u);  // 2^-log_s
    for (uint32_t k = blockIdx.x * kBsThreads + threadIdx.x; k < S; k += gridDim.x * kBsThreads) {
        uint32_t v = 0;
        if (k == 0) v = inv_s;
        else if (k > H) v = m31_mul(rt[S - k], inv_s);
        w[k] = make_uint4(v, 0, 0, 0);
    }
}
// kh[S + k] = lane 0 of the transformed w[k].
__global__ __launch_bounds__(kBsThreads) void k_bs_kernel_pack(uint2 *kh, const uint4 *w, uint32_t log_s)
{
    const uint32_t S = 1u << log_s;
    for (uint32_t k = blockIdx.x * kBsThreads + threadIdx.x; k < S; k += gridDim.x * kBsThreads) {
        const uint4 v = w[k];
        kh[(size_t)S + k] = make_uint2(v.x, v.y);
    }
}

// ---------------------------------------------------------------- host arithmetic

uint32_t hm_add(uint32_t a, uint32_t b) { uint32_t s = a + b; return s >= kP ? s - kP : s; }
uint32_t hm_sub(uint32_t a, uint32_t b) { return a >= b ? a - b : a + (kP - b); }
uint32_t hm_mul(uint32_t a, uint32_t b) { return (uint32_t)(((uint64_t)a * b) % kP); }
uint32_t hm_pow(uint32_t b, uint64_t e)
{
    uint32_t r = 1;
    while (e) {
        if (e & 1) r = hm_mul(r, b);
        b = hm_mul(b, b);
        e >>= 1;
    }
    return r;
}
struct HCm { uint32_t r, i; };
HCm hc_mul(HCm x, HCm y) { return {hm_sub(hm_mul(x.r, y.r), hm_mul(x.i, y.i)), hm_add(hm_mul(x.r, y.i), hm_mul(x.i, y.r))}; }
HCm hc_pow(HCm b, uint64_t e)
{
    HCm r = {1, 0};
    while (e) {
        if (e & 1) r = hc_mul(r, b);
        b = hc_mul(b, b);
        e >>= 1;
    }
    return r;
}
// 2 + 1268011823 i: the circle group generator (order 2^31) as a unit of CM31.
constexpr HCm kCircleGen = {2u, 1268011823u};

struct GlobalPass { uint32_t lt, lb, log_c; };

// DIF stages log_s - 1 .. kLogChunk in balanced passes of <= kLogPassTile - kLogMinCols.
std::vector<GlobalPass> global_passes(uint32_t log_s)
{
    std::vector<GlobalPass> v;
    if (log_s <= kLogChunk) return v;
    const uint32_t stages = log_s - kLogChunk, per = kLogPassTile - kLogMinCols;
    const uint32_t np = (stages + per - 1) / per;
    uint32_t top = log_s - 1;
    for (uint32_t i = 0; i < np; ++i) {
        const uint32_t k = stages / np + (i < stages % np ? 1 : 0);
        const uint32_t lb = top + 1 - k;
        v.push_back({top, lb, kLogPassTile - k});
        top = lb - 1;
    }
    return v;
}

} // namespace

struct azk_basis_plan {
    azk_context *ctx;
    uint32_t log_n;  // domain log size; capacity N = 2^(log_n - 1) per family
    DeviceMem lv[2], w, tw, kh, fact, ifact, rt, flag;
    DeviceMem in_x, in_y, out;  // host-variant staging
    std::mutex mu;
};

namespace {

BsArgs base_args(azk_basis_plan *p)
{
    BsArgs a{};
    a.tw = (const uint2 *)p->tw.p;
    a.kh = (const uint2 *)p->kh.p;
    a.fact = (const uint32_t *)p->fact.p;
    a.ifact = (const uint32_t *)p->ifact.p;
    a.rt = (const uint32_t *)p->rt.p;
    a.flag = (uint32_t *)p->flag.p;
    a.log_n = p->log_n - 1;
    return a;
}

size_t lds_bytes(uint32_t log_elems) { return (size_t)16 << log_elems; }

void launch_fwd(hipStream_t st, BsArgs a, uint32_t transforms)
{
    const uint32_t log_tile = a.lt - a.lb + 1 + a.log_c;
    const uint32_t blocks = transforms << (a.log_s - log_tile);
    hipLaunchKernelGGL(k_bs_fwd, dim3(blocks), dim3(kBsThreads), lds_bytes(log_tile), st, a);
}
void launch_inv(hipStream_t st, BsArgs a, uint32_t transforms)
{
    const uint32_t log_tile = a.lt - a.lb + 1 + a.log_c;
    const uint32_t blocks = transforms << (a.log_s - log_tile);
    hipLaunchKernelGGL(k_bs_inv, dim3(blocks), dim3(kBsThreads), lds_bytes(log_tile), st, a);
}
void launch_chunk(hipStream_t st, BsArgs a, uint32_t transforms)
{
    const uint32_t log_tc = std::min(a.log_s, kLogChunk);
    const uint32_t blocks = transforms << (a.log_s - log_tc);
    hipLaunchKernelGGL(k_bs_chunk, dim3(blocks), dim3(kBsThreads), lds_bytes(log_tc), st, a);
}

bool uses_kernel(uint32_t s) { return (1u << (s - 1)) > kDirectMax || s > kLogTail; }

void launch_tail(hipStream_t st, BsArgs a, uint4 *out)
{
    hipLaunchKernelGGL(k_bs_tail, dim3(2u << a.lvl), dim3(kBsThreads), 3 * lds_bytes(a.log_s), st, a, out);
}

// Forward DIF of one transform of size 2^s held in w[0, 2^s).
void enqueue_forward(hipStream_t st, BsArgs a, uint32_t s)
{
    a.log_s = s;
    a.mode = 0;
    for (const GlobalPass &g : global_passes(s)) {
        a.lt = g.lt;
        a.lb = g.lb;
        a.log_c = g.log_c;
        launch_fwd(st, a, 1);
    }
    a.mode = kForwardOnly;
    launch_chunk(st, a, 1);
}

// K_S = DIF(reversed R_S) / S for every transform size with an NTT shift.
azk_status build_kernels(azk_basis_plan *p)
{
    hipStream_t st = p->ctx->stream;
    const uint32_t L = p->log_n - 1;
    BsArgs a = base_args(p);
    a.w = (uint4 *)p->w.p;
    for (uint32_t s = 2; s <= L; ++s) {
        if (!uses_kernel(s)) continue;
        const uint32_t S = 1u << s, g = std::min<uint32_t>((S + kBsThreads - 1) / kBsThreads, 1024);
        hipLaunchKernelGGL(k_bs_kernel_init, dim3(g), dim3(kBsThreads), 0, st, (uint4 *)p->w.p,
                           (const uint32_t *)p->rt.p, s);
        enqueue_forward(st, a, s);
        hipLaunchKernelGGL(k_bs_kernel_pack, dim3(g), dim3(kBsThreads), 0, st, (uint2 *)p->kh.p,
                           (const uint4 *)p->w.p, s);
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(st));
    return AZK_OK;
}

// The whole conversion: x / y monomial columns (QM31, counts <= N) -> 2N interleaved
// line-basis coefficients.
azk_status basis_enqueue(azk_basis_plan *p, const uint4 *x, uint32_t nx, const uint4 *y, uint32_t ny, uint4 *out)
{
    hipStream_t st = p->ctx->stream;
    const uint32_t L = p->log_n - 1;
    const size_t N = (size_t)1 << L;
    HIP_TRY(hipMemsetAsync(p->flag.p, 0, 4, st));
    BsArgs a = base_args(p);
    a.src[0] = x;
    a.src[1] = y;
    a.count[0] = nx;
    a.count[1] = ny;
    a.w = (uint4 *)p->w.p;
    uint32_t lvl = 0, cur = 0;
    for (; L - lvl > kLogTail; ++lvl) {
        uint4 *dst = (uint4 *)p->lv[cur].p;
        a.lvl = lvl;
        a.log_s = L - lvl;
        a.dst = dst;
        const uint32_t transforms = 4u << lvl;
        if (a.log_s <= kLogChunk) {
            a.mode = kPrepare | kFinalize;
            launch_chunk(st, a, transforms);
        } else {
            const std::vector<GlobalPass> ps = global_passes(a.log_s);
            for (size_t i = 0; i < ps.size(); ++i) {
                a.lt = ps[i].lt;
                a.lb = ps[i].lb;
                a.log_c = ps[i].log_c;
                a.mode = i == 0 ? kPrepare : 0;
                launch_fwd(st, a, transforms);
            }
            a.mode = 0;
            launch_chunk(st, a, transforms);
            for (size_t i = ps.size(); i-- > 0;) {
                a.lt = ps[i].lt;
                a.lb = ps[i].lb;
                a.log_c = ps[i].log_c;
                a.mode = i == 0 ? kFinalize : 0;
                launch_inv(st, a, transforms);
            }
        }
        a.src[0] = dst;
        a.src[1] = dst + N;
        cur ^= 1;
    }
    a.lvl = lvl;
    a.log_s = L - lvl;
    a.mode = 0;
    launch_tail(st, a, out);
    return AZK_OK;
}

azk_status read_flag(azk_basis_plan *p)
{
    uint32_t f = 0;
    HIP_TRY(hipMemcpy(&f, p->flag.p, 4, hipMemcpyDeviceToHost));
    if (f) return fail(AZK_ERR_INVALID_LAYOUT, "non-canonical QM31 limb in the monomial coefficients");
    return AZK_OK;
}

} // namespace

extern "C" {

azk_status azk_witness_basis_plan_create(azk_context *ctx, uint32_t log_n, azk_basis_plan **out)
{
    CHECK_ARG(ctx && out);
    if (log_n < 1 || log_n > kMaxLogN) return fail(AZK_ERR_INVALID_LAYOUT, "domain log size %u outside 1..%u", log_n, kMaxLogN);
    auto *p = new azk_basis_plan{};
    p->ctx = ctx;
    p->log_n = log_n;
    const uint32_t L = log_n - 1;
    const size_t N = (size_t)1 << L, H = std::max<size_t>(N / 2, 1);
    auto bad = [&](azk_status s) { delete p; return s; };
    if (hipSetDevice(ctx->device) != hipSuccess) return bad(fail(AZK_ERR_COMMAND_FAILED, "hipSetDevice failed"));
    azk_status s = AZK_OK;
    const bool levels = L > kLogTail;
    if (levels) {
        if (!s) s = p->lv[0].ensure(32 * N);
        if (!s) s = p->lv[1].ensure(32 * N);
    }
    if (!s) s = p->w.ensure((levels ? 64 : 16) * N);
    if (!s) s = p->tw.ensure(8 * N);
    if (!s) s = p->kh.ensure(16 * N);
    if (!s) s = p->fact.ensure(4 * H);
    if (!s) s = p->ifact.ensure(4 * H);
    if (!s) s = p->rt.ensure(4 * H);
    if (!s) s = p->flag.ensure(4);
    if (s) return bad(s);
    // k!, 1/k!, 2^m/m! for k < N/2 (all below p, so every factorial is a unit)
    std::vector<uint32_t> fact(H), ifact(H), rt(H);
    fact[0] = 1;
    for (size_t k = 1; k < H; ++k) fact[k] = hm_mul(fact[k - 1], (uint32_t)k);
    ifact[H - 1] = hm_pow(fact[H - 1], kP - 2);
    for (size_t k = H - 1; k > 0; --k) ifact[k - 1] = hm_mul(ifact[k], (uint32_t)k);
    uint32_t pw = 1;
    for (size_t m = 0; m < H; ++m) {
        rt[m] = hm_mul(pw, ifact[m]);
        pw = hm_add(pw, pw);
    }
    // tw[h + j] = w_{2h}^j, w_{2h} = g^(2^31 / 2h)
    std::vector<HCm> tw(N, HCm{1, 0});
    for (uint32_t r = 0; r < L; ++r) {
        const size_t h = (size_t)1 << r;
        const HCm w = hc_pow(kCircleGen, (uint64_t)1 << (30 - r));
        HCm v = {1, 0};
        for (size_t j = 0; j < h; ++j) {
            tw[h + j] = v;
            v = hc_mul(v, w);
        }
    }
    if (hipMemcpy(p->fact.p, fact.data(), 4 * H, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->ifact.p, ifact.data(), 4 * H, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->rt.p, rt.data(), 4 * H, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->tw.p, tw.data(), 8 * N, hipMemcpyHostToDevice) != hipSuccess)
        return bad(fail(AZK_ERR_COMMAND_FAILED, "table upload failed"));
    {
        std::lock_guard<std::mutex> gc(ctx->mu);
        s = build_kernels(p);
    }
    if (s) return bad(s);
    ctx_retain(ctx);
    *out = p;
    return AZK_OK;
}

void azk_witness_basis_plan_destroy(azk_basis_plan *p) { destroy_on_ctx(p); }

azk_status azk_witness_basis_execute_resident(azk_basis_plan *p, const azk_buffer *x, size_t x_offset,
                                              uint32_t x_count, const azk_buffer *y, size_t y_offset,
                                              uint32_t y_count, azk_buffer *out, size_t out_offset, azk_stats *stats)
{
    CHECK_ARG(p && out && (x || !x_count) && (y || !y_count));
    std::lock_guard<std::mutex> g(p->mu);
    const size_t N = (size_t)1 << (p->log_n - 1), ob = 32 * N;
    if (x_count > N || y_count > N || x_count + (uint64_t)y_count == 0)
        return fail(AZK_ERR_INVALID_LAYOUT, "coefficient counts %u, %u: each at most %zu, not both zero", x_count,
                    y_count, N);
    azk_status s = check_range(out, out_offset, ob);
    if (!s && x_count) s = check_range(x, x_offset, (size_t)16 * x_count);
    if (!s && y_count) s = check_range(y, y_offset, (size_t)16 * y_count);
    if (s) return s;
    if ((x_count && overlaps(out, out_offset, ob, x, x_offset, (size_t)16 * x_count)) ||
        (y_count && overlaps(out, out_offset, ob, y, y_offset, (size_t)16 * y_count)))
        return fail(AZK_ERR_INVALID_LAYOUT, "output range aliases an input range");
    const uint8_t *xp = x_count ? (const uint8_t *)x->ptr + x_offset : nullptr;
    const uint8_t *yp = y_count ? (const uint8_t *)y->ptr + y_offset : nullptr;
    uint8_t *op = (uint8_t *)out->ptr + out_offset;
    if ((uintptr_t)xp % 16 || (uintptr_t)yp % 16 || (uintptr_t)op % 16)
        return fail(AZK_ERR_INVALID_LAYOUT, "QM31 buffers must be 16-byte aligned");
    HIP_TRY(hipSetDevice(p->ctx->device));
    {
        std::lock_guard<std::mutex> gc(p->ctx->mu);
        s = run_timed(p->ctx, stats, [&]() {
            return basis_enqueue(p, (const uint4 *)xp, x_count, (const uint4 *)yp, y_count, (uint4 *)op);
        });
    }
    return s ? s : read_flag(p);
}

azk_status azk_witness_basis_execute(azk_basis_plan *p, const uint32_t *x, uint32_t x_count, const uint32_t *y,
                                     uint32_t y_count, uint32_t *out, size_t out_words, azk_stats *stats)
{
    CHECK_ARG(p && out && (x || !x_count) && (y || !y_count));
    std::lock_guard<std::mutex> g(p->mu);
    const size_t N = (size_t)1 << (p->log_n - 1), ob = 32 * N;
    if (x_count > N || y_count > N || x_count + (uint64_t)y_count == 0)
        return fail(AZK_ERR_INVALID_LAYOUT, "coefficient counts %u, %u: each at most %zu, not both zero", x_count,
                    y_count, N);
    if (out_words < 8 * N) return fail(AZK_ERR_INVALID_LAYOUT, "output needs %zu words", 8 * N);
    HIP_TRY(hipSetDevice(p->ctx->device));
    azk_status s = p->in_x.ensure((size_t)16 * std::max<uint32_t>(x_count, 1));
    if (!s) s = p->in_y.ensure((size_t)16 * std::max<uint32_t>(y_count, 1));
    if (!s) s = p->out.ensure(ob);
    if (s) return s;
    {
        std::lock_guard<std::mutex> gc(p->ctx->mu);
        s = run_timed(p->ctx, stats, [&]() -> azk_status {
            hipStream_t st = p->ctx->stream;
            if (x_count) HIP_TRY(hipMemcpyAsync(p->in_x.p, x, (size_t)16 * x_count, hipMemcpyHostToDevice, st));
            if (y_count) HIP_TRY(hipMemcpyAsync(p->in_y.p, y, (size_t)16 * y_count, hipMemcpyHostToDevice, st));
            azk_status e = basis_enqueue(p, (const uint4 *)p->in_x.p, x_count, (const uint4 *)p->in_y.p, y_count,
                                         (uint4 *)p->out.p);
            if (e) return e;
            HIP_TRY(hipMemcpyAsync(out, p->out.p, ob, hipMemcpyDeviceToHost, st));
            return AZK_OK;
        });
    }
    return s ? s : read_flag(p);
}

// executeVerified (CircleWitnessFFTBasis.swift:483-517): the device basis against the
// monomial columns at pseudo-random QM31 points on the host (the reference recomputes the
// basis with its O(n^2) oracle; a point check costs O(n) and misses a wrong basis with
// probability below 2^-98 per point).
azk_status azk_witness_basis_execute_verified(azk_basis_plan *p, const uint32_t *x, uint32_t x_count,
                                              const uint32_t *y, uint32_t y_count, uint32_t *out, size_t out_words,
                                              azk_stats *stats)
{
    azk_status s = azk_witness_basis_execute(p, x, x_count, y, y_count, out, out_words, stats);
    if (s) return s;
    if (!host::basis_consistent(x, x_count, y, y_count, p->log_n, out))
        return fail(AZK_ERR_CORRECTNESS, "witness basis device result did not match the CPU gate");
    return AZK_OK;
}

} // extern "C"
