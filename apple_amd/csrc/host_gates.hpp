// CPU gates of the *_verified entry points: host computations of the path's results,
// compared with what the device produced (the reference's *Verified variants compare
// with its CPU oracles, e.g. CircleCodeword.swift:476-490, CircleFRIFoldPlan.swift:209-225,
// CircleWitnessFFTBasis.swift:483-517, CircleCodeword.swift:1082-1113).  QM31 arrays are
// 4 LE u32 words per element.
#pragma once
#include <stddef.h>
#include <stdint.h>

namespace azk::host {

// Circle FFT of N = 2^log_n interleaved line-basis coefficients on the canonical
// bit-reversed domain (CircleCodewordOracle.evaluateWithCircleFFT, CircleCodeword.swift:72-98).
void circle_fft(const uint32_t *coeffs, uint32_t log_n, uint32_t *out);

// `rounds` folds of N = 2^log_n evaluations with explicit challenges
// (CircleFRILayerOracleV1.fold, CircleDomain.swift:513-544): N >> rounds outputs; false
// on a zero domain coordinate (the reference throws).
bool fold_chain(const uint32_t *evals, uint32_t log_n, const uint32_t *challenges, uint32_t rounds, uint32_t *out);

// Whether `basis` (2^log_n interleaved line-basis coefficients) represents the x / y
// monomial columns: both sides evaluated at `points` pseudo-random QM31 points per column
// (derived from the inputs and the basis).  A wrong basis passes a point with probability
// below 2^log_n / p^4 < 2^-98.
bool basis_consistent(const uint32_t *x, uint32_t nx, const uint32_t *y, uint32_t ny, uint32_t log_n,
                      const uint32_t *basis, uint32_t points = 2);

} // namespace azk::host
