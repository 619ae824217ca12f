"""M31 multilinear sum-check (the serving job): device prover and host verifier vs the oracle.

Mirrors testM31MultilinearSumcheckVerifiesEvaluationTableProtocol
(Tests/AppleZKProverTests/CircleDomainTests.swift:267-330): statement fields, verifier
acceptance, codec round trip with a trailing byte rejected, a tampered round polynomial
rejected, and a wrong claimed sum refused by the prover.

The transcript bytes are pinned by the reference's integrated-artifact corpus
(Resources/ApplicationPublicTheoremIntegratedArtifactCorpusV1.json, asserted at
ApplicationTheoremTests.swift:1766-1771): its AIR multilinear sum-check proof is the 16-lane
proof of the Fibonacci trace's padded constraint evaluations inside a 96-byte wrapper, and the
SHA3-256 of those 300 bytes is given.  The proof carries every round polynomial and challenge,
so that digest fixes the header, round and challenge frames and the codec.
"""
import hashlib
import json
import os
import struct

import numpy as np
import pytest

import pyoracle as O
from apple_amd import azk, inputs

P = O.P
EVALS8 = np.array([3, 5, 7, 11, 13, 17, 19, 23], np.uint32)
AIR = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.json")))[
    "air_sumcheck_integrated"]

# ------------------------------------------------------------------ AIR artifact restatement
# Fibonacci AIR of the corpus (ApplicationTheoremTests.swift:2392-2436): transitions
# next.a - cur.b and next.b - cur.a - cur.b, boundaries a0 = 1, b0 = 1, b3 = 5.  A term is
# (coefficient, [(kind, column)]) with kind 0 = current row, 1 = next row.
FIB_TRANSITIONS = [[(1, [(1, 0)]), (P - 1, [(0, 1)])],
                   [(1, [(1, 1)]), (P - 1, [(0, 0)]), (P - 1, [(0, 1)])]]
FIB_BOUNDARIES = [(row, [(1, [(0, col)]), ((P - val) % P, [])]) for row, col, val in [(0, 0, 1), (0, 1, 1), (3, 1, 5)]]


def u32(v):
    return struct.pack("<I", v)


def u64(v):
    return struct.pack("<Q", v)


def air_definition_digest():
    """AIRDefinitionDigestV1.digest (ApplicationTheorem.swift:2213-2251)."""
    def poly(terms):
        b = u32(len(terms))
        for c, factors in terms:
            b += u32(c) + u32(len(factors)) + b"".join(u32(k) + u32(col) for k, col in factors)
        return b
    dom = b"AppleZKProver.AIRDefinition.V1"
    d = u32(len(dom)) + dom + u32(P) + u32(1) + u32(2) + u32(len(FIB_TRANSITIONS))
    d += b"".join(poly(t) for t in FIB_TRANSITIONS) + u32(len(FIB_BOUNDARIES))
    d += b"".join(u64(row) + poly(t) for row, t in FIB_BOUNDARIES)
    return hashlib.sha3_256(d).digest()


def padded_constraint_evaluations(columns):
    """AIRSemanticVerifierV1.constraintEvaluations + AIRToSumcheckReductionV1.paddedEvaluationVector
    (ApplicationTheorem.swift:2155-2210,2276-2292): every transition on rows 0..n-2, then the
    boundaries, zero-padded to a power of two of at least 2."""
    rows = len(columns[0])

    def ev(terms, r):
        acc = 0
        for c, factors in terms:
            prod = c
            for kind, col in factors:
                prod = prod * columns[col][r + kind] % P
            acc = (acc + prod) % P
        return acc
    out = [ev(t, r) for r in range(rows - 1) for t in FIB_TRANSITIONS] + [ev(t, r) for r, t in FIB_BOUNDARIES]
    target = 2
    while target < len(out):
        target <<= 1
    return np.array(out + [0] * (target - len(out)), np.uint32)


def air_wrapper(sumcheck_proof: bytes, evaluation_digest: bytes, rows: int, cols: int) -> bytes:
    """AIRConstraintMultilinearSumcheckProofCodecV1.encode (ApplicationTheorem.swift:6947-6962)."""
    return (b"AZKACM1\x00" + u32(1) + air_definition_digest() + u64(rows) + u64(cols) + evaluation_digest +
            u32(len(sumcheck_proof)) + sumcheck_proof)


def test_air_reduction_inputs_match_the_corpus():
    cols = AIR["witness_columns"]
    assert air_definition_digest().hex() == AIR["air_definition_digest"]
    ev = padded_constraint_evaluations(cols)
    assert len(ev) == 16 and not ev.any()           # 9 evaluations of a valid trace, padded
    bad = padded_constraint_evaluations([[1, 1, 2, 3], [1, 2, 3, 6]])
    assert bad.any()                                # a broken trace moves the vector
    assert O.sumcheck_multilinear(ev)["digest"].hex() == AIR["sumcheck_initial_digest"]
    fin, _, _ = O.sumcheck_chunk(ev, 4)             # M31SumcheckProofBuilderV1 at rounds = log2(16)
    fw = np.array(fin, np.uint32)
    dom = b"AppleZKProver.M31Sumcheck.Words.V1"
    framed = u32(len(dom)) + dom + u32(1) + u64(len(fw)) + fw.astype("<u4").tobytes()
    assert hashlib.sha3_256(framed).hexdigest() == AIR["sumcheck_final_digest"]


def test_oracle_multilinear_proof_reproduces_the_integrated_artifact_digest():
    ev = padded_constraint_evaluations(AIR["witness_columns"])
    inner = O.sumcheck_multilinear_proof(ev)
    digest = O.sumcheck_multilinear(ev)["digest"]
    wrapped = air_wrapper(inner, digest, 4, 2)
    assert len(wrapped) == AIR["air_sumcheck_byte_count"]
    assert hashlib.sha3_256(wrapped).hexdigest() == AIR["air_sumcheck_digest"]
    dom = b"AppleZKProver.AIRConstraintMultilinearSumcheckProof.V1"  # ProofDigestV1, :6990-7002
    assert hashlib.sha3_256(u32(len(dom)) + dom + u32(len(wrapped)) + wrapped).hexdigest() == \
        AIR["air_sumcheck_domain_digest"]
    assert azk.sumcheck_verify(inner)
    # a one-bit change of any challenge word breaks the pin
    for k in range(4):
        off = len(inner) - 4 - 4 * (4 - k)
        if 0 <= off:
            assert hashlib.sha3_256(air_wrapper(flip(inner, off), digest, 4, 2)).hexdigest() != AIR["air_sumcheck_digest"]

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
def test_gpu_integrated_artifact_sumcheck(ctx):
    """The device plan's 16-lane proof of the corpus trace is the pinned inner proof."""
    print("sumcheck integrated artifact 16 lanes", flush=True)
    ev = padded_constraint_evaluations(AIR["witness_columns"])
    proof = azk.M31MultilinearSumcheckPlan(ctx, 4).prove(ev)
    assert proof == O.sumcheck_multilinear_proof(ev)
    wrapped = air_wrapper(proof, O.sumcheck_multilinear(ev)["digest"], 4, 2)
    assert hashlib.sha3_256(wrapped).hexdigest() == AIR["air_sumcheck_digest"]


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
