"""GPU parity: every device path against the CPU oracle, byte for byte, through the C ABI.

Sizes are chosen so the oracle finishes in seconds; the full 2^21 commit (config 3 shape)
is compared whole, and re-checked through the independent verifier.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import pyoracle as O
from apple_amd import azk, inputs

pytestmark = pytest.mark.gpu
P = O.P
V = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.json")))


@pytest.fixture(scope="module")
def ctx():
    c = azk.Context(0)
    print("device:", c.capabilities(), flush=True)
    yield c
    c.close()


def sha3(b) -> str:
    return hashlib.sha3_256(bytes(b)).hexdigest()


# ------------------------------------------------------------------ Keccak-f / one-block
@pytest.mark.parametrize("count,in_stride,out_stride", [(1, 200, 200), (1000, 200, 200), (777, 208, 216)])
def test_keccakf_batch(ctx, count, in_stride, out_stride):
    print(f"keccakf count={count} strides={in_stride}/{out_stride}", flush=True)
    raw = inputs.random_bytes(count, count * in_stride)
    plan = azk.KeccakF1600PermutationBatchPlan(ctx, count, in_stride, out_stride)
    got = plan.permute(raw)
    assert np.array_equal(got, O.keccakf_batch(raw, in_stride, out_stride, count))
    plan.permute(raw, verified=True)


def test_keccakf_resident_in_place(ctx):
    count = 4096
    raw = inputs.random_bytes(9, count * 200)
    buf = ctx.upload(raw)
    plan = azk.KeccakF1600PermutationBatchPlan(ctx, count)
    plan.permute_resident(buf, buf)
    assert np.array_equal(buf.download(), O.keccakf_batch(raw, 200, 200, count))


@pytest.mark.parametrize("length", [0, 1, 31, 32, 64, 100, 135, 136])
@pytest.mark.parametrize("stride", ["tight", "aligned", "odd"])
def test_oneblock_sha3_and_keccak(ctx, length, stride):
    st = {"tight": max(length, 1), "aligned": 144, "odd": 139}[stride]
    st = max(st, length)
    count = 513
    print(f"oneblock len={length} stride={st}", flush=True)
    msgs = inputs.random_bytes(length * 7 + 1, count * st)
    for cls, suffix in [(azk.SHA3FixedOneBlockHashPlan, 0x06), (azk.Keccak256FixedOneBlockHashPlan, 0x01)]:
        for out_stride in (32, 40):
            plan = cls(ctx, count, st, length, out_stride)
            got = plan.hash(msgs)
            want = O.hash_oneblock_batch(msgs, st, length, count, suffix, out_stride)
            assert np.array_equal(got, want)
    plan = azk.SHA3FixedOneBlockHashPlan(ctx, count, st, length)
    plan.hash(msgs, verified=True)


def test_oneblock_rejects_137(ctx):
    with pytest.raises(azk.AZKError) as e:
        azk.SHA3FixedOneBlockHashPlan(ctx, 4, 137, 137)
    assert e.value.status == "unsupportedOneBlockLength"


def test_oneblock_reference_boundary_kats(ctx):
    for length, want in V["sha3_256"]["rate_boundary"].items():
        n = int(length)
        if n > 136:
            continue
        msg = np.frombuffer(bytes((i * 11 + 7) & 255 for i in range(n)), np.uint8)
        plan = azk.SHA3FixedOneBlockHashPlan(ctx, 1, max(n, 1), n)
        assert plan.hash(msg if n else np.zeros(1, np.uint8)).tobytes().hex() == want
        kplan = azk.Keccak256FixedOneBlockHashPlan(ctx, 1, max(n, 1), n)
        assert kplan.hash(msg if n else np.zeros(1, np.uint8)).tobytes().hex() == V["keccak_256"]["rate_boundary"][length]


# ------------------------------------------------------------------ Merkle
@pytest.mark.parametrize("count,length,stride", [(1, 64, 64), (2, 64, 64), (8, 32, 32), (1024, 17, 21),
                                                  (4096, 136, 136), (1 << 16, 64, 64)])
def test_merkle_commit_and_openings(ctx, count, length, stride):
    print(f"merkle count={count} len={length} stride={stride}", flush=True)
    raw = inputs.random_bytes(count + length, count * stride)
    plan = azk.SHA3RawLeavesMerkleCommitPlan(ctx, count, stride, length)
    root = plan.commit(raw)
    assert root == O.merkle_root(raw, count, stride, length)
    idx = inputs.splitmix64(count, 16) % np.uint64(count)
    order = [0, count - 1, *[int(x) for x in idx]]
    for i in order:
        sib, r2 = plan.open(i)
        want_sib, want_root = O.merkle_opening(raw, count, stride, length, i)
        assert sib == want_sib and r2 == want_root == root
        assert O.merkle_verify(raw[i * stride:i * stride + length].tobytes(), i, sib, root)
    # every opening in one call (repeats included) equals the one-at-a-time form
    paths, r3 = plan.open_batch(order + order[:3])
    assert r3 == root and paths == [plan.open(i)[0] for i in order + order[:3]]


@pytest.mark.parametrize("length", [0, 32, 64, 128, 135, 136])
def test_merkle_reference_suite_roots(ctx, length):
    count = 16384
    leaf = np.arange(count, dtype=np.int64)[:, None]
    j = np.arange(length, dtype=np.int64)[None, :]
    raw = ((leaf * 131 + j * 17 + 0x5A) & 255).astype(np.uint8)
    plan = azk.SHA3RawLeavesMerkleCommitPlan(ctx, count, length, length)
    assert plan.commit(raw).hex() == V["merkle_suite_roots_16384"][str(length)]


@pytest.mark.parametrize("subtree_log", [1, 4, 8, 11])
def test_merkle_subtree_shapes_agree(ctx, subtree_log):
    count, length = 1 << 13, 48
    raw = inputs.random_bytes(subtree_log, count * length)
    plan = azk.SHA3RawLeavesMerkleCommitPlan(ctx, count, length, length, subtree_log)
    assert plan.commit(raw, verified=True) == O.merkle_root(raw, count, length, length)


def test_merkle_resident_offset_and_clear(ctx):
    count, length = 512, 64
    raw = inputs.random_bytes(3, count * length)
    buf = ctx.buffer(count * length + 256)
    buf.upload(raw, offset=256)
    plan = azk.SHA3RawLeavesMerkleCommitPlan(ctx, count, length, length)
    assert plan.commit_resident(buf, 256) == O.merkle_root(raw, count, length, length)
    assert plan.open_batch([])[0] == []
    with pytest.raises(azk.AZKError):
        plan.open_batch([1, count])  # one index out of range rejects the call
    plan.clear()
    with pytest.raises(azk.AZKError):
        plan.open(0)
    with pytest.raises(azk.AZKError):
        plan.open_batch([0])
    with pytest.raises(azk.AZKError):
        plan.commit_resident(buf, 512)  # range past the end


def test_prehashed_root(ctx):
    count = 1 << 12
    hashes = inputs.random_bytes(5, count * 32)
    plan = azk.SHA3PrehashedLeavesMerkleCommitPlan(ctx, count)
    level = [hashes[32 * k:32 * k + 32].tobytes() for k in range(count)]
    while len(level) > 1:
        level = [hashlib.sha3_256(level[i] + level[i + 1]).digest() for i in range(0, len(level), 2)]
    assert plan.commit_resident(ctx.upload(hashes)) == level[0]


# ------------------------------------------------------------------ Circle FFT / fold
@pytest.mark.parametrize("log_n", [1, 2, 3, 5, 8, 11, 12, 14, 17, 20])
def test_codeword_fft(ctx, log_n):
    print(f"fft log_n={log_n}", flush=True)
    rng = np.random.default_rng(log_n)
    coeffs = rng.integers(0, P, size=(1 << log_n, 4), dtype=np.uint32)
    plan = azk.CircleCodewordPlan(ctx, log_n)
    assert np.array_equal(plan.read_twiddles(), O.fft_twiddles(log_n))
    assert np.array_equal(plan.execute(coeffs), O.circle_fft(coeffs, log_n))


def _qm31vec(count, a, b, c, d):
    def m31vec(salt):
        i = np.arange(count, dtype=np.uint64)
        v = (i + 1) * np.uint64(1048573) + np.uint64(salt) * np.uint64(65537) + i * np.uint64(17) * np.uint64(salt | 1)
        return (v % np.uint64(P)).astype(np.uint32)
    return np.stack([m31vec(a), m31vec(b), m31vec(c), m31vec(d)], 1)


def test_codeword_reference_digest(ctx):
    x = _qm31vec(8, 0xE01, 0xE03, 0xE07, 0xE0B)
    y = _qm31vec(7, 0xE11, 0xE17, 0xE1D, 0xE23)
    basis = O.basis_from_monomial(x, y, 14)
    cw = azk.CircleCodewordPlan(ctx, 14).execute(basis)
    assert sha3(cw.astype("<u4").tobytes()) == V["proofs"]["codeword_2^14"]["codeword"]


def test_fold_reference_digests(ctx):
    fb = V["fold_baselines"]
    ev = _qm31vec(1 << 14, 0xC31, 0xC37, 0xC3D, 0xC43)
    out = azk.CircleFRIFoldChainPlan(ctx, 14, 1).execute(ev, [(41, 43, 47, 53)])
    assert sha3(out.astype("<u4").tobytes()) == fb["first_fold_2^14"]
    ev = _qm31vec(1 << 14, 0xCC1, 0xCC7, 0xCCD, 0xCD3)
    chs = [(41 + 18 * r, 43 + 22 * r, 47 + 26 * r, 53 + 30 * r) for r in range(3)]
    out = azk.CircleFRIFoldChainPlan(ctx, 14, 3).execute(ev, chs)
    assert sha3(out.astype("<u4").tobytes()) == fb["explicit_chain_2^14_r3"]


@pytest.mark.parametrize("log_n,rounds", [(2, 1), (6, 6), (13, 4), (18, 10)])
def test_fold_chain_random(ctx, log_n, rounds):
    rng = np.random.default_rng(100 + log_n)
    ev = rng.integers(0, P, size=(1 << log_n, 4), dtype=np.uint32)
    ch = rng.integers(0, P, size=(rounds, 4), dtype=np.uint32)
    got = azk.CircleFRIFoldChainPlan(ctx, log_n, rounds).execute(ev, ch)
    assert np.array_equal(got, O.fold_chain(ev, log_n, ch))


# ------------------------------------------------------------------ Circle PCS prover
def _oracle_proof_from_coeffs(coeffs, log_n, sec, digest, rounds):
    cw = O.circle_fft(coeffs, log_n)
    return O.fri_prove(cw.reshape(-1, 4), log_n, sec, digest, rounds)[0]


@pytest.mark.parametrize("key,log_n", [("codeword_2^14", 14), ("codeword_2^10", 10)])
def test_prover_reference_proof_digests(ctx, key, log_n):
    n = 1 << log_n
    x = _qm31vec(min(8, n // 2), 0xE01, 0xE03, 0xE07, 0xE0B)
    y = _qm31vec(min(7, n // 2), 0xE11, 0xE17, 0xE1D, 0xE23)
    digest = bytes(0xA0 + i for i in range(32))
    prover = azk.CircleCodewordPCSFRIProverV1(ctx, log_n, 3, (2, 4, 1, 0), digest)
    proof = prover.prove_fft_coefficients(O.basis_from_monomial(x, y, log_n))
    assert len(proof) == V["proofs"][key]["proof_bytes"]
    assert sha3(proof) == V["proofs"][key]["proof"]


def _stable(count):
    i = np.arange(count)
    return np.stack([17 + i * 3, 29 + i * 5, 31 + i * 7, 43 + i * 11], 1).astype(np.uint32)


def test_prover_from_evaluations_reference_digests(ctx):
    d1 = bytes(0x10 + i for i in range(32))
    p1 = azk.CircleCodewordPCSFRIProverV1(ctx, 4, 1, (2, 3, 1, 0), d1)
    assert sha3(p1.prove_evaluations_resident(ctx.upload(_stable(16)))) == V["proofs"]["first_fold_log4"]
    d2 = bytes(0x50 + i for i in range(32))
    p2 = azk.CircleCodewordPCSFRIProverV1(ctx, 5, 3, (2, 5, 1, 0), d2)
    assert sha3(p2.prove_evaluations_resident(ctx.upload(_stable(32)))) == V["proofs"]["three_round_log5"]


@pytest.mark.parametrize("log_n,rounds,queries", [(3, 2, 3), (9, 5, 7), (12, 11, 16), (15, 10, 32), (17, 6, 5)])
def test_prover_random_shapes(ctx, log_n, rounds, queries):
    print(f"prover log_n={log_n} rounds={rounds} q={queries}", flush=True)
    sec = (2, queries, 1, 0)
    digest = inputs.random_bytes(log_n, 32).tobytes()
    coeffs = inputs.circle_fft_coefficients(seed=log_n * 31, log_n=log_n)
    prover = azk.CircleCodewordPCSFRIProverV1(ctx, log_n, rounds, sec, digest)
    buf = ctx.upload(coeffs)
    proof = prover.prove_fft_coefficients_resident(buf)
    assert proof == _oracle_proof_from_coeffs(coeffs, log_n, sec, digest, rounds)
    prover.set_graph(False)
    assert prover.prove_fft_coefficients_resident(buf) == proof
    prover.set_graph(True)
    assert prover.prove_fft_coefficients_resident(buf) == proof


def test_prover_full_config3_shape(ctx):
    """2^21 codeword from 2 x 2^19 M31 coefficients, 10 rounds, 32 queries, whole proof vs oracle."""
    log_n, rounds, sec = 21, 10, (2, 32, 1, 0)
    digest = bytes(0xA0 + i for i in range(32))
    coeffs = inputs.circle_fft_coefficients(seed=1, log_n=log_n)
    print("prover 2^21 r10 q32", flush=True)
    prover = azk.CircleCodewordPCSFRIProverV1(ctx, log_n, rounds, sec, digest)
    buf = ctx.upload(coeffs)
    proof = prover.prove_fft_coefficients_resident(buf)
    assert len(proof) == 393284
    want = _oracle_proof_from_coeffs(coeffs, log_n, sec, digest, rounds)
    assert proof == want
    assert O.fri_verify(proof, digest) == 1
    for _ in range(3):  # graph replays are stable
        assert prover.prove_fft_coefficients_resident(buf) == proof


def test_prover_offsets_timing_and_clear(ctx):
    log_n, rounds, sec = 11, 4, (2, 9, 1, 0)
    digest = bytes(range(32))
    coeffs = inputs.circle_fft_coefficients(seed=77, log_n=log_n)
    buf = ctx.buffer(coeffs.nbytes + 4096)
    buf.upload(coeffs, offset=4096)
    prover = azk.CircleCodewordPCSFRIProverV1(ctx, log_n, rounds, sec, digest)
    want = _oracle_proof_from_coeffs(coeffs, log_n, sec, digest, rounds)
    assert prover.prove_fft_coefficients_resident(buf, 4096) == want
    prover.set_timing(True)
    assert prover.prove_fft_coefficients_resident(buf, 4096) == want
    t = prover.phase_times()
    assert t["total_ns"] > 0 and t["transcript_ns"] > 0
    prover.set_timing(False)
    prover.clear()
    assert prover.prove_fft_coefficients_resident(buf, 4096) == want
    with pytest.raises(azk.AZKError):
        prover.prove_fft_coefficients_resident(buf, 4100)  # misaligned
    with pytest.raises(azk.AZKError) as e:  # above CircleFRIGrindingV1.maximumLocalSearchBits
        azk.CircleCodewordPCSFRIProverV1(ctx, log_n, rounds, (2, 9, 1, 21), digest)
    assert e.value.code == 6


@pytest.mark.parametrize("log_n,rounds,bits", [(6, 3, 1), (10, 4, 6), (12, 5, 12), (16, 8, 16), (8, 4, 20)])
def test_prover_grinding_matches_oracle(ctx, log_n, rounds, bits):
    """The smallest nonce of the reference's sequential search (findGrindingNonce,
    CircleProofFormat.swift:752-791) in the proof tail: host threads up to 12 bits, the
    device search above."""
    print(f"grinding 2^{log_n} bits {bits}", flush=True)
    sec = (2, 9, 1, bits)
    digest = bytes(0x30 + i for i in range(32))
    coeffs = inputs.circle_fft_coefficients(seed=700 + bits, log_n=log_n)
    want = _oracle_proof_from_coeffs(coeffs, log_n, sec, digest, rounds)
    prover = azk.CircleCodewordPCSFRIProverV1(ctx, log_n, rounds, sec, digest)
    got = prover.prove_fft_coefficients(coeffs)
    assert got == want
    assert azk.circle_verify(got, digest)
    batch = azk.CircleBatchPlan([prover], [ctx.upload(coeffs)])
    assert batch.prove() == [want]


def test_batch_uneven_jobs_match_plain(ctx):
    rng = np.random.default_rng(4)
    logs = [int(x) for x in rng.integers(8, 16, size=12)]
    provers, bufs, wants = [], [], []
    for j, ln in enumerate(logs):
        sec = (2, 8, 1, 0)
        digest = inputs.random_bytes(1000 + j, 32).tobytes()
        coeffs = inputs.circle_fft_coefficients(seed=500 + j, log_n=ln)
        pr = azk.CircleCodewordPCSFRIProverV1(ctx, ln, min(10, ln - 2), sec, digest)
        b = ctx.upload(coeffs)
        provers.append(pr)
        bufs.append(b)
        wants.append(pr.prove_fft_coefficients_resident(b))
    proofs, stats = azk.prove_batch(provers, bufs)
    assert proofs == wants
    for j in (0, len(logs) - 1):
        ln = logs[j]
        coeffs = inputs.circle_fft_coefficients(seed=500 + j, log_n=ln)
        digest = inputs.random_bytes(1000 + j, 32).tobytes()
        assert proofs[j] == _oracle_proof_from_coeffs(coeffs, ln, (2, 8, 1, 0), digest, min(10, ln - 2))


def test_batch_plan_replays_and_evaluation_jobs(ctx):
    logs = [9, 12, 10, 13, 9]
    provers, bufs, wants = [], [], []
    for j, ln in enumerate(logs):
        digest = inputs.random_bytes(2000 + j, 32).tobytes()
        coeffs = inputs.circle_fft_coefficients(seed=900 + j, log_n=ln)
        pr = azk.CircleCodewordPCSFRIProverV1(ctx, ln, ln - 3, (2, 6, 1, 0), digest)
        b = ctx.upload(coeffs)
        provers.append(pr)
        bufs.append(b)
        wants.append(_oracle_proof_from_coeffs(coeffs, ln, (2, 6, 1, 0), digest, ln - 3))
    plan = azk.CircleBatchPlan(provers, bufs)
    for _ in range(3):
        assert plan.prove() == wants
    # plain proves still agree after batch replays
    assert [p.prove_fft_coefficients_resident(b) for p, b in zip(provers, bufs)] == wants
    plan.close()
