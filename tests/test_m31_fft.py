"""The M31-input paths: the codeword of M31 coefficients (azk_codeword_execute_m31*) and the
prover from M31 coefficients (azk_circle_prove_fft_coefficients_m31*).

The reference transforms the QM31 lift (c, 0, 0, 0) of M31 coefficients
(CircleCodewordOracle.evaluateWithCircleFFT, CircleCodeword.swift:72-98; the benchmark's P, Q
are M31, lifted as in apple_amd/inputs.py).  Its butterflies multiply by M31 twiddles, so the
lift's upper components stay zero: the M31 codeword must equal the oracle's QM31 codeword's
first component with the other three all zero, and the proof from M31 coefficients must be
byte-identical to the QM31 path's (and the oracle's) on the lifted array.
"""
import numpy as np
import pytest

import pyoracle as O
from apple_amd import azk, inputs

P = O.P


@pytest.fixture(scope="module")
def ctx():
    c = azk.Context(0)
    yield c
    c.close()


def lift(c):
    q = np.zeros((c.size, 4), np.uint32)
    q[:, 0] = c
    return q


@pytest.mark.gpu
@pytest.mark.parametrize("log_n", [1, 2, 3, 5, 8, 11, 12, 13, 14, 17, 20])
def test_codeword_m31_is_the_lifted_codeword(ctx, log_n):
    print(f"m31 fft log_n={log_n}", flush=True)
    rng = np.random.default_rng(700 + log_n)
    c = rng.integers(0, P, size=1 << log_n, dtype=np.uint32)
    want = O.circle_fft(lift(c), log_n).reshape(-1, 4)
    assert not want[:, 1:].any()  # the lift's upper components stay zero through the FFT
    plan = azk.CircleCodewordPlan(ctx, log_n)
    assert np.array_equal(plan.execute_m31(c), want[:, 0])


@pytest.mark.gpu
@pytest.mark.parametrize("log_n", [22, 24])
def test_codeword_m31_at_scale_against_the_qm31_path(ctx, log_n):
    """Sizes where the M31 passes take two or three HBM passes; the QM31 device path (itself
    checked against the oracle up to 2^20 and against the host FFT at 2^24 and 2^26) is the
    reference here."""
    print(f"m31 fft log_n={log_n}", flush=True)
    rng = np.random.default_rng(800 + log_n)
    c = rng.integers(0, P, size=1 << log_n, dtype=np.uint32)
    plan = azk.CircleCodewordPlan(ctx, log_n)
    got = plan.execute_m31(c)
    q = plan.execute(lift(c)).reshape(-1, 4)
    assert not q[:, 1:].any()
    assert np.array_equal(got, q[:, 0])


@pytest.mark.gpu
def test_codeword_m31_resident_offsets_and_errors(ctx):
    log_n = 12
    n = 1 << log_n
    rng = np.random.default_rng(9)
    c = rng.integers(0, P, size=n, dtype=np.uint32)
    plan = azk.CircleCodewordPlan(ctx, log_n)
    buf_in = ctx.upload(np.concatenate([np.zeros(3, np.uint32), c]))
    buf_out = ctx.buffer(4 * n + 8)
    plan.execute_m31_resident(buf_in, buf_out, coeff_offset=12, out_offset=8)
    got = buf_out.download(4 * n, offset=8).view(np.uint32)
    assert np.array_equal(got, plan.execute_m31(c))
    with pytest.raises(azk.AZKError):  # misaligned
        plan.execute_m31_resident(buf_in, buf_out, coeff_offset=2)
    with pytest.raises(azk.AZKError):  # too short
        plan.execute_m31_resident(buf_in, buf_out, coeff_offset=16)
    with pytest.raises(azk.AZKError):  # aliased
        plan.execute_m31_resident(buf_in, buf_in)


@pytest.mark.gpu
@pytest.mark.parametrize("log_n,rounds,queries", [(3, 2, 3), (9, 5, 7), (12, 11, 16), (15, 10, 32), (17, 6, 5),
                                                  (21, 10, 32)])
def test_prover_m31_is_the_qm31_proof(ctx, log_n, rounds, queries):
    print(f"m31 prover log_n={log_n} rounds={rounds} q={queries}", flush=True)
    sec = (2, queries, 1, 0)
    digest = inputs.random_bytes(log_n + 1, 32).tobytes()
    coeffs = inputs.circle_fft_coefficients(seed=log_n * 37, log_n=log_n)
    m31 = np.ascontiguousarray(coeffs[:, 0])
    prover = azk.CircleCodewordPCSFRIProverV1(ctx, log_n, rounds, sec, digest)
    want = prover.prove_fft_coefficients_resident(ctx.upload(coeffs))
    if log_n <= 17:
        assert want == O.fri_prove(O.circle_fft(coeffs, log_n).reshape(-1, 4), log_n, sec, digest, rounds)[0]
    buf = ctx.upload(m31)
    assert prover.prove_fft_coefficients_m31_resident(buf) == want
    for _ in range(2):  # graph replays
        assert prover.prove_fft_coefficients_m31_resident(buf) == want
    prover.set_graph(False)
    assert prover.prove_fft_coefficients_m31_resident(buf) == want
    prover.set_graph(True)
    assert prover.prove_fft_coefficients_m31(m31) == want
    # switching back to the QM31 input rebuilds the program
    assert prover.prove_fft_coefficients_resident(ctx.upload(coeffs)) == want


@pytest.mark.gpu
def test_prover_m31_input_errors(ctx):
    prover = azk.CircleCodewordPCSFRIProverV1(ctx, 10, 3, (2, 4, 1, 0), bytes(32))
    buf = ctx.upload(np.zeros(1 << 10, np.uint32))
    with pytest.raises(azk.AZKError):
        prover.prove_fft_coefficients_m31_resident(buf, offset=4)  # one word short
    with pytest.raises(azk.AZKError):
        prover.prove_fft_coefficients_m31(np.zeros(100, np.uint32))
