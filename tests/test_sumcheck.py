"""M31 multilinear sum-check (the serving job): device prover and host verifier vs the oracle.

Mirrors testM31MultilinearSumcheckVerifiesEvaluationTableProtocol
(Tests/AppleZKProverTests/CircleDomainTests.swift:267-330): statement fields, verifier
acceptance, codec round trip with a trailing byte rejected, a tampered round polynomial
rejected, and a wrong claimed sum refused by the prover.  The exact challenge bytes have no
reference vector; they are compared with the oracle's restatement of the same transcript.
"""
import numpy as np
import pytest

import pyoracle as O
from apple_amd import azk, inputs

P = O.P
EVALS8 = np.array([3, 5, 7, 11, 13, 17, 19, 23], np.uint32)

# byte offsets inside the codec (SumcheckOracle.swift:1176-1194)
OFF_CLAIMED = 8 + 4 + 4 + 8 + 4
OFF_DIGEST = OFF_CLAIMED + 4
OFF_WORDS = OFF_DIGEST + 32 + 8


def rounds_off(n):
    return OFF_WORDS + 4 * n + 8


def test_host_verifier_accepts_oracle_proof():
    proof = O.sumcheck_multilinear_proof(EVALS8)
    ml = O.sumcheck_multilinear(EVALS8)
    assert ml["claimed_sum"] == int(EVALS8.sum()) % P
    assert len(ml["challenges"]) == 3
    assert azk.sumcheck_verify(proof)


def flip(b: bytes, off: int, delta: int = 1) -> bytes:
    bb = bytearray(b)
    v = int.from_bytes(bb[off:off + 4], "little")
    bb[off:off + 4] = ((v + delta) % P).to_bytes(4, "little")
    return bytes(bb)


@pytest.mark.parametrize("where", ["round_e0", "round_e1", "challenge", "final", "claimed", "word"])
def test_host_verifier_rejects_tampering(where):
    n = len(EVALS8)
    proof = O.sumcheck_multilinear_proof(EVALS8)
    off = {"round_e0": rounds_off(n), "round_e1": rounds_off(n) + 4,
           "challenge": rounds_off(n) + 8 * 3 + 8, "final": len(proof) - 4,
           "claimed": OFF_CLAIMED, "word": OFF_WORDS + 8}[where]
    assert not azk.sumcheck_verify(flip(proof, off))


def test_host_verifier_rejects_digest_and_layout():
    proof = O.sumcheck_multilinear_proof(EVALS8)
    bad = bytearray(proof)
    bad[OFF_DIGEST] ^= 1
    assert not azk.sumcheck_verify(bytes(bad))
    with pytest.raises(azk.AZKError):
        azk.sumcheck_verify(proof + b"\x00")       # trailing bytes: decode throws
    with pytest.raises(azk.AZKError):
        azk.sumcheck_verify(proof[:-1])
    nc = bytearray(proof)
    nc[OFF_WORDS:OFF_WORDS + 4] = P.to_bytes(4, "little")  # non-canonical word
    with pytest.raises(azk.AZKError):
        azk.sumcheck_verify(bytes(nc))


@pytest.mark.parametrize("log_n", [1, 5, 9])
def test_host_verifier_random(log_n):
    ev = inputs.random_m31(100 + log_n, 1 << log_n)
    assert azk.sumcheck_verify(O.sumcheck_multilinear_proof(ev))


# ------------------------------------------------------------------ device prover
@pytest.fixture(scope="module")
def ctx():
    c = azk.Context(0)
    yield c
    c.close()


@pytest.mark.gpu
def test_gpu_reference_shape(ctx):
    print("sumcheck 8 lanes", flush=True)
    plan = azk.M31MultilinearSumcheckPlan(ctx, 3)
    proof = plan.prove(EVALS8)
    assert proof == O.sumcheck_multilinear_proof(EVALS8)
    assert azk.sumcheck_verify(proof)
    with pytest.raises(azk.AZKError) as e:
        plan.prove(EVALS8, claimed_sum=(int(EVALS8.sum()) + 1) % P)
    assert e.value.code == 7  # AZK_ERR_CORRECTNESS
    assert plan.prove(EVALS8, claimed_sum=int(EVALS8.sum()) % P) == proof


# 1..12: the single-block tail kernel alone; 13..: reduce / round / fold kernels first.
@pytest.mark.gpu
@pytest.mark.parametrize("log_n", [1, 2, 4, 7, 11, 12, 13, 14, 16, 19, 22])
def test_gpu_matches_oracle(ctx, log_n):
    print(f"sumcheck 2^{log_n}", flush=True)
    ev = inputs.random_m31(7000 + log_n, 1 << log_n)
    plan = azk.M31MultilinearSumcheckPlan(ctx, log_n)
    proof = plan.prove(ev)
    assert proof == O.sumcheck_multilinear_proof(ev)
    assert azk.sumcheck_verify(proof)


@pytest.mark.gpu
def test_gpu_edge_values(ctx):
    for name, ev in [("zeros", np.zeros(1 << 13, np.uint32)),
                     ("p-1", np.full(1 << 13, P - 1, np.uint32)),
                     ("alternating", np.tile(np.array([0, P - 1], np.uint32), 1 << 12))]:
        print("sumcheck edge", name, flush=True)
        plan = azk.M31MultilinearSumcheckPlan(ctx, 13)
        assert plan.prove(ev) == O.sumcheck_multilinear_proof(ev), name


@pytest.mark.gpu
@pytest.mark.parametrize("log_n", [6, 15])
def test_gpu_non_canonical_rejected(ctx, log_n):
    ev = inputs.random_m31(5, 1 << log_n)
    ev[(1 << log_n) - 3] = P
    plan = azk.M31MultilinearSumcheckPlan(ctx, log_n)
    with pytest.raises(azk.AZKError) as e:
        plan.prove(ev)
    assert e.value.code == 6  # AZK_ERR_INVALID_LAYOUT


@pytest.mark.gpu
@pytest.mark.parametrize("log_n,offset", [(10, 4), (17, 12)])
def test_gpu_resident(ctx, log_n, offset):
    print(f"sumcheck resident 2^{log_n} offset {offset}", flush=True)
    n = 1 << log_n
    ev = inputs.random_m31(300 + log_n, n)
    raw = np.concatenate([np.zeros(offset // 4, np.uint32), ev, np.zeros(3, np.uint32)])
    buf = ctx.upload(raw)
    plan = azk.M31MultilinearSumcheckPlan(ctx, log_n)
    ml = O.sumcheck_multilinear(ev)
    for dg in (None, ml["digest"]):
        r = plan.prove_resident(buf, offset, initial_digest=dg)
        assert r["initial_digest"] == ml["digest"]
        assert r["claimed_sum"] == ml["claimed_sum"]
        assert r["final_evaluation"] == ml["final"]
        assert r["round_polys"].reshape(-1, 2).tolist() == ml["round_polys"]
        assert r["challenges"].tolist() == ml["challenges"]
    # the caller's words are left as they were
    assert np.array_equal(buf.download().view(np.uint32)[offset // 4: offset // 4 + n], ev)


@pytest.mark.gpu
def test_gpu_repeat_and_plan_reuse(ctx):
    plan = azk.M31MultilinearSumcheckPlan(ctx, 15)
    a = inputs.random_m31(1, 1 << 15)
    b = inputs.random_m31(2, 1 << 15)
    pa, pb = plan.prove(a), plan.prove(b)
    assert pa == O.sumcheck_multilinear_proof(a)
    assert pb == O.sumcheck_multilinear_proof(b)
    assert plan.prove(a) == pa
