"""Sizes past the oracle's reach, through size-independent checks with the library's CPU gates
(which tests/test_gates.py pins to the oracle and to the reference's digests): the codeword
plan at 2^22, 2^24 and 2^26 (the bench's standalone shape) equal to the host Circle FFT, the fold
chain at 2^24 equal to the host fold, and a 2^24 commit (10 rounds, 32 queries) whose proof
verifies and whose every opened first-layer value equals the host FFT of its coefficients."""
import numpy as np
import pytest

from apple_amd import azk, inputs

P = 2**31 - 1
DIGEST = bytes(0x60 + i for i in range(32))


@pytest.fixture(scope="module")
def ctx():
    c = azk.Context(0)
    yield c
    c.close()


@pytest.mark.gpu
@pytest.mark.parametrize("log_n", [22, 24, 26])
def test_codeword_matches_host_fft(ctx, log_n):
    print(f"codeword 2^{log_n} vs host FFT", flush=True)
    c = np.random.default_rng(log_n).integers(0, P, size=(1 << log_n, 4), dtype=np.uint32)
    buf, out = ctx.upload(c), ctx.buffer(16 << log_n)
    azk.CircleCodewordPlan(ctx, log_n).execute_resident(buf, out)
    got = out.download().view(np.uint32)
    buf.free()
    out.free()
    assert np.array_equal(got, azk.host_circle_fft(c, log_n))


@pytest.mark.gpu
def test_fold_chain_matches_host_fold(ctx):
    log_n, rounds = 24, 10
    print(f"fold chain 2^{log_n} r{rounds} vs host fold", flush=True)
    rng = np.random.default_rng(77)
    ev = rng.integers(0, P, size=(1 << log_n, 4), dtype=np.uint32)
    ch = rng.integers(0, P, size=(rounds, 4), dtype=np.uint32)
    got = azk.CircleFRIFoldChainPlan(ctx, log_n, rounds).execute(ev, ch)
    assert np.array_equal(got, azk.host_fold_chain(ev, log_n, ch))


@pytest.mark.gpu
def test_commit_2_24_verifies_with_coefficients(ctx):
    log_n = 24
    print(f"commit 2^{log_n} r10 q32", flush=True)
    c = inputs.circle_fft_coefficients(seed=24, log_n=log_n)
    prover = azk.CircleCodewordPCSFRIProverV1(ctx, log_n, 10, (2, 32, 1, 0), DIGEST)
    buf = ctx.upload(c)
    proof = prover.prove_fft_coefficients_resident(buf)
    assert prover.prove_fft_coefficients_resident(buf) == proof  # replay of the captured program
    assert azk.circle_verify_with_coefficients(proof, DIGEST, c)
    bad = c.copy()
    bad[12345, 0] = (int(bad[12345, 0]) + 1) % P
    assert not azk.circle_verify_with_coefficients(proof, DIGEST, bad)
