"""Monomial witness columns -> Circle-FFT-basis coefficients on the device
(azk_witness_basis_*, CircleWitnessToFFTBasisPlanV1 CircleWitnessFFTBasis.swift:230-517) against
the oracle's circleFFTCoefficients restatement (CircleCodeword.swift:100-127, O(n^2)) at sizes it
finishes in seconds, against the reference's 2^14 codeword digest and the proof corpus, and, at
2^21-2^23, through the codeword: FFT(basis) must equal direct Horner evaluation of the monomial
polynomial at sampled domain points."""
import hashlib
import json
import os

import numpy as np
import pytest

import pyoracle as O
from apple_amd import azk

P = O.P
HERE = os.path.dirname(__file__)
V = json.load(open(os.path.join(HERE, "golden", "reference_vectors.json")))


def _m31vec(count, salt):  # makeDeterministicM31Vector, main.swift:1336-1350
    i = np.arange(count, dtype=np.uint64)
    v = (i + 1) * np.uint64(1048573) + np.uint64(salt) * np.uint64(65537) + i * np.uint64(17) * np.uint64(salt | 1)
    return (v % np.uint64(P)).astype(np.uint32)


def _qm31vec(count, a, b, c, d):
    return np.stack([_m31vec(count, a), _m31vec(count, b), _m31vec(count, c), _m31vec(count, d)], 1)


def test_oracle_evaluate_at_matches_direct():
    rng = np.random.default_rng(3)
    for log_n, nx, ny in [(1, 1, 1), (5, 16, 3), (9, 200, 256), (11, 1024, 0)]:
        x = rng.integers(0, P, size=(nx, 4), dtype=np.uint32)
        y = rng.integers(0, P, size=(ny, 4), dtype=np.uint32)
        full = O.evaluate_direct(x, y, log_n).reshape(-1, 4)
        idx = rng.integers(0, 1 << log_n, size=17)
        assert np.array_equal(O.evaluate_at(x, y, log_n, idx), full[idx])


@pytest.fixture(scope="module")
def ctx():
    c = azk.Context(0)
    yield c
    c.close()


CASES = [(1, (1, 1)), (2, (2, 1)), (3, (0, 4)), (7, (64, 64)), (7, (3, 0)), (8, (128, 17)), (11, (1024, 1024)),
         (12, (2048, 2048)), (12, (1000, 5)), (13, (4096, 4096)), (14, (8192, 8192)), (15, (16384, 9000))]


@pytest.mark.gpu
@pytest.mark.parametrize("log_n,counts", CASES)
def test_basis_matches_oracle(ctx, log_n, counts):
    print(f"witness basis log_n={log_n} counts={counts}", flush=True)
    rng = np.random.default_rng(log_n * 1000 + counts[0])
    x = rng.integers(0, P, size=(counts[0], 4), dtype=np.uint32)
    y = rng.integers(0, P, size=(counts[1], 4), dtype=np.uint32)
    plan = azk.CircleWitnessToFFTBasisPlanV1(ctx, log_n)
    got = plan.execute(x, y)
    assert np.array_equal(got, O.basis_from_monomial(x, y, log_n))
    assert np.array_equal(plan.execute(x, y), got)  # repeat on the same plan


@pytest.mark.gpu
def test_edge_polynomials(ctx):
    """Constants, single high-degree terms and all-(p-1) columns at a size with chunk and tail levels."""
    log_n, N = 12, 1 << 11
    plan = azk.CircleWitnessToFFTBasisPlanV1(ctx, log_n)
    one = np.array([[1, 0, 0, 0]], np.uint32)
    for x, y in [(one, one), (np.zeros((N, 4), np.uint32), one),
                 (np.vstack([np.zeros((N - 1, 4), np.uint32), [[5, 6, 7, 8]]]), np.zeros((0, 4), np.uint32)),
                 (np.full((N, 4), P - 1, np.uint32), np.full((N, 4), P - 1, np.uint32))]:
        assert np.array_equal(plan.execute(x, y), O.basis_from_monomial(x, y, log_n))


@pytest.mark.gpu
def test_reference_codeword_digest(ctx):
    """makeDeterministicCircleCodewordPolynomial (8 x / 7 y coefficients, main.swift:1376-1395) at
    2^14: device basis, then the device FFT, hashes to the reference's published codeword digest."""
    x = _qm31vec(8, 0xE01, 0xE03, 0xE07, 0xE0B)
    y = _qm31vec(7, 0xE11, 0xE17, 0xE1D, 0xE23)
    basis = azk.CircleWitnessToFFTBasisPlanV1(ctx, 14).execute(x, y)
    assert np.array_equal(basis, O.basis_from_monomial(x, y, 14))
    cw = azk.CircleCodewordPlan(ctx, 14).execute(basis)
    assert hashlib.sha3_256(cw.astype("<u4").tobytes()).hexdigest() == V["proofs"]["codeword_2^14"]["codeword"]


@pytest.mark.gpu
def test_corpus_proof_from_monomials(ctx):
    """The proof corpus valid case (Resources/CirclePCSFRIProofCorpusV1.json): 2 x + 2 y monomial
    coefficients at log N 6 through the device basis and the device prover give the corpus proof
    (up to its claimed-openings section)."""
    corpus = json.load(open(os.path.join(HERE, "golden", "circle_pcs_fri_proof_corpus_v1.json")))
    vc, prof = corpus["validCase"], corpus["profile"]
    x = np.frombuffer(b"".join(bytes.fromhex(h) for h in vc["xCoefficientHex"]), "<u4").reshape(-1, 4)
    y = np.frombuffer(b"".join(bytes.fromhex(h) for h in vc["yCoefficientHex"]), "<u4").reshape(-1, 4)
    log_n = vc["domainLogSize"]
    basis = azk.CircleWitnessToFFTBasisPlanV1(ctx, log_n).execute(x, y)
    assert np.array_equal(basis, O.basis_from_monomial(x, y, log_n))
    sec = (prof["logBlowupFactor"], prof["queryCount"], prof["foldingStep"], prof["grindingBits"])
    prover = azk.CircleCodewordPCSFRIProverV1(ctx, log_n, vc["roundCount"], sec,
                                              bytes.fromhex(vc["publicInputDigestHex"]))
    proof = prover.prove_fft_coefficients(basis)
    assert bytes.fromhex(vc["proofHex"]).startswith(proof)


@pytest.mark.gpu
@pytest.mark.parametrize("log_n", [17, 21, 23])
def test_large_through_codeword(ctx, log_n):
    """Sizes past the oracle's reach: FFT(device basis) equals Horner evaluation of the monomial
    polynomial at 48 sampled storage indices; resident execution with offsets equals the host one."""
    print(f"witness basis large log_n={log_n}", flush=True)
    N = 1 << (log_n - 1)
    rng = np.random.default_rng(log_n)
    x = rng.integers(0, P, size=(N, 4), dtype=np.uint32)
    y = rng.integers(0, P, size=(N - 3, 4), dtype=np.uint32)
    plan = azk.CircleWitnessToFFTBasisPlanV1(ctx, log_n)
    basis = plan.execute(x, y)
    cw = azk.CircleCodewordPlan(ctx, log_n).execute(basis).reshape(-1, 4)
    idx = np.concatenate([[0, 1, (1 << log_n) - 1], rng.integers(0, 1 << log_n, size=45)])
    assert np.array_equal(cw[idx], O.evaluate_at(x, y, log_n, idx))
    bx = ctx.upload(np.vstack([np.zeros((2, 4), np.uint32), x]))
    by = ctx.upload(y)
    out = azk.Buffer(ctx, 16 * (2 * N + 4))
    plan.execute_resident(bx, len(x), by, len(y), out, x_offset=32, out_offset=64)
    assert np.array_equal(out.download(32 * N, 64).view(np.uint32), basis)


@pytest.mark.gpu
def test_rejects_bad_layouts(ctx):
    plan = azk.CircleWitnessToFFTBasisPlanV1(ctx, 6)
    good = np.ones((4, 4), np.uint32)
    bad = good.copy()
    bad[2, 1] = P  # QM31Field.modulus limb (CircleDomainTests.swift:1981-1999)
    for args in [(bad, good), (good, bad), (np.zeros((0, 4), np.uint32), np.zeros((0, 4), np.uint32)),
                 (np.ones((33, 4), np.uint32), good)]:
        with pytest.raises(azk.AZKError) as e:
            plan.execute(*args)
        assert e.value.code == 6
    assert np.array_equal(plan.execute(good, good), O.basis_from_monomial(good, good, 6))  # plan still usable
    buf = ctx.upload(np.ones((32, 4), np.uint32))
    with pytest.raises(azk.AZKError) as e:  # output aliases the x column
        plan.execute_resident(buf, 4, None, 0, buf)
    assert e.value.code == 6
