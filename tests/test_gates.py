"""The CPU gates of the *_verified entry points (host code in libazk): the host Circle FFT
and fold chain against the oracle and against the reference's published digests
(codeword 2^14 main.swift:1376-1395, fold baselines), the witness-basis point check and the
proof check with coefficients, each with tampered inputs that must be refused; on the GPU,
every *_verified entry point on valid inputs."""
import hashlib
import json
import os

import numpy as np
import pytest

import pyoracle as O
from apple_amd import azk, inputs

P = O.P
HERE = os.path.dirname(__file__)
V = json.load(open(os.path.join(HERE, "golden", "reference_vectors.json")))
SEC = (2, 4, 1, 0)
DIGEST = bytes(0xA0 + i for i in range(32))


def sha3(b):
    return hashlib.sha3_256(b).hexdigest()


def _m31vec(count, salt):  # makeDeterministicM31Vector, main.swift:1336-1350
    i = np.arange(count, dtype=np.uint64)
    v = (i + 1) * np.uint64(1048573) + np.uint64(salt) * np.uint64(65537) + i * np.uint64(17) * np.uint64(salt | 1)
    return (v % np.uint64(P)).astype(np.uint32)


def _qm31vec(count, a, b, c, d):
    return np.stack([_m31vec(count, a), _m31vec(count, b), _m31vec(count, c), _m31vec(count, d)], 1)


@pytest.mark.parametrize("log_n", list(range(1, 13)))
def test_host_fft_matches_oracle(log_n):
    c = np.random.default_rng(log_n).integers(0, P, size=(1 << log_n, 4), dtype=np.uint32)
    assert np.array_equal(azk.host_circle_fft(c, log_n), O.circle_fft(c, log_n))


def test_host_fft_reference_digest():
    x = _qm31vec(8, 0xE01, 0xE03, 0xE07, 0xE0B)
    y = _qm31vec(7, 0xE11, 0xE17, 0xE1D, 0xE23)
    cw = azk.host_circle_fft(O.basis_from_monomial(x, y, 14), 14)
    assert sha3(cw.astype("<u4").tobytes()) == V["proofs"]["codeword_2^14"]["codeword"]


@pytest.mark.parametrize("log_n,rounds", [(1, 1), (2, 2), (6, 3), (9, 9), (13, 5)])
def test_host_fold_chain_matches_oracle(log_n, rounds):
    rng = np.random.default_rng(50 + log_n)
    ev = rng.integers(0, P, size=(1 << log_n, 4), dtype=np.uint32)
    ch = rng.integers(0, P, size=(rounds, 4), dtype=np.uint32)
    assert np.array_equal(azk.host_fold_chain(ev, log_n, ch), O.fold_chain(ev, log_n, ch))


def test_host_fold_reference_digests():
    fb = V["fold_baselines"]
    ev = _qm31vec(1 << 14, 0xC31, 0xC37, 0xC3D, 0xC43)
    out = azk.host_fold_chain(ev, 14, [(41, 43, 47, 53)])
    assert sha3(out.astype("<u4").tobytes()) == fb["first_fold_2^14"]
    ev = _qm31vec(1 << 14, 0xCC1, 0xCC7, 0xCCD, 0xCD3)
    chs = [(41 + 18 * r, 43 + 22 * r, 47 + 26 * r, 53 + 30 * r) for r in range(3)]
    assert sha3(azk.host_fold_chain(ev, 14, chs).astype("<u4").tobytes()) == fb["explicit_chain_2^14_r3"]


def test_host_gates_reject_non_canonical():
    c = np.zeros((4, 4), np.uint32)
    c[1, 2] = P
    for call in (lambda: azk.host_circle_fft(c, 2), lambda: azk.host_fold_chain(c, 2, [(1, 0, 0, 0)])):
        with pytest.raises(azk.AZKError) as e:
            call()
        assert e.value.code == 6


@pytest.mark.parametrize("log_n,counts", [(1, (1, 1)), (3, (4, 0)), (6, (32, 7)), (10, (512, 512))])
def test_basis_check(log_n, counts):
    rng = np.random.default_rng(log_n)
    x = rng.integers(0, P, size=(counts[0], 4), dtype=np.uint32)
    y = rng.integers(0, P, size=(counts[1], 4), dtype=np.uint32)
    basis = O.basis_from_monomial(x, y, log_n).reshape(-1, 4)
    assert azk.witness_basis_check(x, y, log_n, basis)
    n = 1 << log_n
    for k in sorted({0, 1, n // 2, n - 1}):
        bad = basis.copy()
        bad[k, rng.integers(0, 4)] ^= 1
        assert not azk.witness_basis_check(x, y, log_n, bad), k
    if counts[0]:
        xb = x.copy()
        xb[-1, 0] = (int(xb[-1, 0]) + 1) % P
        assert not azk.witness_basis_check(xb, y, log_n, basis)
    if counts[0] == counts[1] and counts[0]:
        assert not azk.witness_basis_check(y, x, log_n, basis)


def _proof_case(log_n, rounds):
    c = np.random.default_rng(log_n).integers(0, P, size=(1 << log_n, 4), dtype=np.uint32)
    proof, _ = O.fri_prove(O.circle_fft(c, log_n).reshape(-1, 4), log_n, SEC, DIGEST, rounds)
    return c, proof


@pytest.mark.parametrize("log_n,rounds", [(4, 2), (8, 5), (12, 10)])
def test_verify_with_coefficients(log_n, rounds):
    c, proof = _proof_case(log_n, rounds)
    assert azk.circle_verify_with_coefficients(proof, DIGEST, c)
    assert not azk.circle_verify_with_coefficients(proof, bytes(32), c)
    bad = c.copy()
    bad[len(c) // 3, 1] = (int(bad[len(c) // 3, 1]) + 5) % P  # changes the codeword everywhere
    assert not azk.circle_verify_with_coefficients(proof, DIGEST, bad)
    with pytest.raises(azk.AZKError) as e:
        azk.circle_verify_with_coefficients(proof, DIGEST, c[:-1])
    assert e.value.code == 6


def test_verify_with_coefficients_corpus():
    """The corpus proof (claimed openings included) against the basis of its monomials."""
    corpus = json.load(open(os.path.join(HERE, "golden", "circle_pcs_fri_proof_corpus_v1.json")))
    vc = corpus["validCase"]
    x = np.frombuffer(b"".join(bytes.fromhex(h) for h in vc["xCoefficientHex"]), "<u4").reshape(-1, 4)
    y = np.frombuffer(b"".join(bytes.fromhex(h) for h in vc["yCoefficientHex"]), "<u4").reshape(-1, 4)
    basis = O.basis_from_monomial(x, y, vc["domainLogSize"])
    proof, digest = bytes.fromhex(vc["proofHex"]), bytes.fromhex(vc["publicInputDigestHex"])
    assert azk.circle_verify_with_coefficients(proof, digest, basis)
    other = O.basis_from_monomial(y, x, vc["domainLogSize"])
    assert not azk.circle_verify_with_coefficients(proof, digest, other)


@pytest.fixture(scope="module")
def ctx():
    c = azk.Context(0)
    yield c
    c.close()


@pytest.mark.gpu
def test_verified_entry_points(ctx):
    rng = np.random.default_rng(9)
    for log_n in (5, 14, 18):
        print(f"verified entry points 2^{log_n}", flush=True)
        c = rng.integers(0, P, size=(1 << log_n, 4), dtype=np.uint32)
        plan = azk.CircleCodewordPlan(ctx, log_n)
        assert np.array_equal(plan.execute(c, verified=True), plan.execute(c))
        rounds = min(4, log_n)
        ch = rng.integers(0, P, size=(rounds, 4), dtype=np.uint32)
        fold = azk.CircleFRIFoldChainPlan(ctx, log_n, rounds)
        assert np.array_equal(fold.execute(c, ch, verified=True), fold.execute(c, ch))
        x = rng.integers(0, P, size=(1 << (log_n - 1), 4), dtype=np.uint32)
        y = rng.integers(0, P, size=(3, 4), dtype=np.uint32)
        wb = azk.CircleWitnessToFFTBasisPlanV1(ctx, log_n)
        assert np.array_equal(wb.execute(x, y, verified=True), wb.execute(x, y))
        prover = azk.CircleCodewordPCSFRIProverV1(ctx, log_n, min(5, log_n - 1), SEC, DIGEST)
        proof = prover.prove_fft_coefficients(c, verified=True)
        assert proof == prover.prove_fft_coefficients(c)
        assert prover.prove_fft_coefficients_resident(ctx.upload(c), verified=True) == proof


@pytest.mark.gpu
def test_verified_prover_at_scale(ctx):
    """The gate at the bench shape (2^21, 10 rounds, 32 queries)."""
    from apple_amd import inputs
    c = inputs.circle_fft_coefficients(seed=1, log_n=21)
    prover = azk.CircleCodewordPCSFRIProverV1(ctx, 21, 10, (2, 32, 1, 0), DIGEST)
    proof = prover.prove_fft_coefficients_resident(ctx.upload(c), verified=True)
    assert azk.circle_verify(proof, DIGEST)


# ------------------------------------------------------------------ host proof builder
@pytest.mark.parametrize("log_n,rounds,sec", [(4, 1, (2, 3, 1, 0)), (5, 3, (2, 5, 1, 0)), (10, 7, (2, 32, 1, 0)),
                                              (12, 10, (2, 32, 1, 0)), (9, 5, (2, 6, 1, 7))])
def test_host_proof_builder_matches_oracle(log_n, rounds, sec):
    """The verified prover's gate: the host rebuild of the whole proof equals the oracle's
    (whose proof builder is pinned by the reference's proof digests and corpus)."""
    digest = bytes(0x40 + i for i in range(32))
    coeffs = inputs.circle_fft_coefficients(seed=30 + log_n, log_n=log_n)
    want = O.fri_prove(O.circle_fft(coeffs, log_n).reshape(-1, 4), log_n, sec, digest, rounds)[0]
    assert azk.host_circle_prove(coeffs, log_n, rounds, sec, digest) == want


def test_host_proof_builder_grinding_and_bounds():
    """Grinding nonce search on the host (8 bits) equals the oracle's; rounds above log N are
    refused as an error, not clamped."""
    digest = bytes(0x50 + i for i in range(32))
    c = inputs.circle_fft_coefficients(seed=5, log_n=6)
    assert azk.host_circle_prove(c, 6, 3, (2, 5, 1, 8), digest) == \
        O.fri_prove(O.circle_fft(c, 6).reshape(-1, 4), 6, (2, 5, 1, 8), digest, 3)[0]
    with pytest.raises(azk.AZKError):
        azk.host_circle_prove(c[:32], 5, 6, (2, 5, 1, 0), digest)  # rounds > log N
