// CPU gates of the *_verified entry points: host computations of the path's results,
// compared with what the device produced (the reference's *Verified variants compare
// with its CPU oracles, e.g. CircleCodeword.swift:476-490, CircleFRIFoldPlan.swift:209-225,
// CircleWitnessFFTBasis.swift:483-517, CircleCodeword.swift:1082-1113).  QM31 arrays are
// 4 LE u32 words per element.
#pragma once
#include <stddef.h>
#include <stdint.h>

#include <vector>

#include "../../include/azk.h"
#include "host_field.hpp"

namespace azk::host {

// Circle FFT of N = 2^log_n interleaved line-basis coefficients on the canonical
// bit-reversed domain (CircleCodewordOracle.evaluateWithCircleFFT, CircleCodeword.swift:72-98).
void circle_fft(const uint32_t *coeffs, uint32_t log_n, uint32_t *out);

// `rounds` folds of N = 2^log_n evaluations with explicit challenges
// (CircleFRILayerOracleV1.fold, CircleDomain.swift:513-544): N >> rounds outputs; false
// on a zero domain coordinate (the reference throws).
bool fold_chain(const uint32_t *evals, uint32_t log_n, const uint32_t *challenges, uint32_t rounds, uint32_t *out);

// Fold round r of one layer (2 n_out QM31 in, n_out out) with challenge ch, pair coordinates
// from `pts`, the storage-order points of the original domain; false on a zero coordinate.
bool fold_round(const std::vector<Pt> &pts, uint32_t r, const uint32_t *in, uint64_t n_out, const uint32_t ch[4],
                uint32_t *out);

// SHA3-256 Merkle root over `count` (a power of two) 16-byte QM31 leaves: leaf =
// SHA3(16 bytes), parent = SHA3(left || right) (MerkleOracle.rootSHA3_256, MerkleOracle.swift:4-52).
void qm31_merkle_root(const uint32_t *leaves, uint64_t count, uint8_t root[32]);

// Whether `basis` (2^log_n interleaved line-basis coefficients) represents the x / y
// monomial columns: both sides evaluated at `points` pseudo-random QM31 points per column
// (derived from the inputs and the basis).  A wrong basis passes a point with probability
// below 2^log_n / p^4 < 2^-98.
bool basis_consistent(const uint32_t *x, uint32_t nx, const uint32_t *y, uint32_t ny, uint32_t log_n,
                      const uint32_t *basis, uint32_t points = 2);

// CircleFRIProofBuilderV1.prove (CircleProofFormat.swift:1172-1395) of the Circle FFT of the
// line-basis coefficients, on the host: the proof bytes the prover with descriptor `d` must
// emit (no claimed openings).  False on a zero fold coordinate.
bool circle_prove(const uint32_t *coeffs, const azk_circle_prover_desc &d, std::vector<uint8_t> &proof);

} // namespace azk::host
