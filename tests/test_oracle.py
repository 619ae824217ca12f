"""Pins the CPU oracle (oracle/azk_oracle.c) before anything is compared against it.

Each test checks the oracle against a vector the reference's own tests or baselines
hold, against an independent library (Python hashlib), or against a property the
mathematics fixes (brute force on small inputs, closed forms, invariants).
"""
import hashlib
import json
import os
import random

import numpy as np
import pytest

import pyoracle as O

P = O.P
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
V = json.load(open(os.path.join(GOLDEN, "reference_vectors.json")))


def sha3(b: bytes) -> str:
    return hashlib.sha3_256(b).hexdigest()


def pattern(length: int) -> bytes:  # SHA3OracleTests.swift:54-56
    return bytes((i * 11 + 7) & 255 for i in range(length))


# ---------------------------------------------------------------- Keccak / SHA3
def test_sha3_256_reference_kats():
    assert O.sha3_256(b"").hex() == V["sha3_256"]["empty"]
    assert O.sha3_256(b"abc").hex() == V["sha3_256"]["abc"]
    for length, want in V["sha3_256"]["rate_boundary"].items():
        assert O.sha3_256(pattern(int(length))).hex() == want


def test_keccak_256_reference_kats():
    assert O.keccak_256(b"").hex() == V["keccak_256"]["empty"]
    assert O.keccak_256(b"abc").hex() == V["keccak_256"]["abc"]
    for length, want in V["keccak_256"]["rate_boundary"].items():
        assert O.keccak_256(pattern(int(length))).hex() == want
    assert O.keccak_256(b"abc") != O.sha3_256(b"abc")


def test_sponge_matches_hashlib_random_lengths():
    rng = random.Random(7)
    for length in [0, 1, 31, 32, 64, 135, 136, 137, 271, 272, 273, 1000, 4097]:
        msg = bytes(rng.randrange(256) for _ in range(length))
        assert O.sha3_256(msg) == hashlib.sha3_256(msg).digest()
        assert O.sponge(msg, 72, 0x06, 64) == hashlib.sha3_512(msg).digest()
        # SHAKE256 over several squeeze blocks pins every lane across permutations
        assert O.sponge(msg, 136, 0x1F, 700) == hashlib.shake_256(msg).digest(700)
        assert O.sponge(msg, 168, 0x1F, 400) == hashlib.shake_128(msg).digest(400)


def test_keccak_f1600_zero_state_kat():
    # FIPS 202 / Keccak team intermediate values: first lane of f(0).
    assert O.keccak_f1600([0] * 25)[0] == 0xF1258F7940E1DDE7


def test_keccak_f1600_on_arbitrary_states_via_hashlib():
    # A padded one-block message absorbed into the zero state is an arbitrary
    # 17-lane state; its permutation's first 4 lanes are the SHA3-256 digest.
    rng = random.Random(11)
    for _ in range(64):
        msg = bytes(rng.randrange(256) for _ in range(rng.randrange(136)))
        block = bytearray(136)
        block[: len(msg)] = msg
        block[len(msg)] ^= 0x06
        block[135] ^= 0x80
        lanes = list(np.frombuffer(bytes(block), "<u8")) + [0] * 8
        out = O.keccak_f1600([int(x) for x in lanes])
        assert np.array(out[:4], "<u8").tobytes() == hashlib.sha3_256(msg).digest()


def test_keccakf_batch_strided_and_padding_zeroed():
    rng = np.random.default_rng(3)
    states = rng.integers(0, 2**64, size=(5, 25), dtype=np.uint64)
    raw = np.zeros((5, 208), np.uint8)
    raw[:, :200] = states.view(np.uint8).reshape(5, 200)
    out = O.keccakf_batch(raw, in_stride=208, out_stride=216, count=5).reshape(5, 216)
    for k in range(5):
        want = O.keccak_f1600([int(x) for x in states[k]])
        assert np.array_equal(out[k, :200].view("<u8"), np.array(want, np.uint64))
        assert not out[k, 200:].any()


def test_oneblock_batch_matches_hashlib():
    rng = np.random.default_rng(5)
    for length in [0, 1, 64, 135, 136]:
        msgs = rng.integers(0, 256, size=(9, 140), dtype=np.uint8)
        out = O.hash_oneblock_batch(msgs, 140, length, 9, 0x06, out_stride=40).reshape(9, 40)
        for k in range(9):
            assert out[k, :32].tobytes() == hashlib.sha3_256(msgs[k, :length].tobytes()).digest()
            assert not out[k, 32:].any()
    with pytest.raises(ValueError):
        O.hash_oneblock_batch(np.zeros(137, np.uint8), 137, 137, 1, 0x06)


# ---------------------------------------------------------------- Merkle
def test_merkle_reference_roots():
    def leaves(count, length, salt=0):  # MerkleTests.swift:592-600
        return bytes((leaf * 29 + j * 7 + salt + 3) & 255 for leaf in range(count) for j in range(length))

    assert O.merkle_root(leaves(8, 32), 8, 32, 32).hex() == V["merkle_test_roots"]["8x32_salt0"]
    assert O.merkle_root(leaves(1024, 32, 71), 1024, 32, 32).hex() == V["merkle_test_roots"]["1024x32_salt71"]


@pytest.mark.parametrize("length", [0, 32, 64, 128, 135, 136])
def test_merkle_suite_roots(length):
    count = 16384  # main.swift:1300-1314 generator
    leaf = np.arange(count, dtype=np.int64)[:, None]
    j = np.arange(length, dtype=np.int64)[None, :]
    raw = ((leaf * 131 + j * 17 + 0x5A) & 255).astype(np.uint8)
    assert O.merkle_root(raw, count, length, length).hex() == V["merkle_suite_roots_16384"][str(length)]


def _hashlib_root(leaves):
    level = [hashlib.sha3_256(x).digest() for x in leaves]
    while len(level) > 1:
        level = [hashlib.sha3_256(level[i] + level[i + 1]).digest() for i in range(0, len(level), 2)]
    return level[0]


def test_merkle_matches_hashlib_bruteforce_and_openings():
    rng = random.Random(9)
    for count in [1, 2, 4, 32]:
        length, stride = 17, 24
        raw = bytes(rng.randrange(256) for _ in range(count * stride))
        leaves = [raw[k * stride:k * stride + length] for k in range(count)]
        root = O.merkle_root(raw, count, stride, length)
        assert root == _hashlib_root(leaves)
        for idx in range(count):
            sib, r2 = O.merkle_opening(raw, count, stride, length, idx)
            assert r2 == root
            assert O.merkle_verify(leaves[idx], idx, sib, root)
            if sib:
                bad = bytearray(sib)
                bad[rng.randrange(len(bad))] ^= 1
                assert not O.merkle_verify(leaves[idx], idx, bytes(bad), root)
    with pytest.raises(ValueError):
        O.merkle_root(b"\0" * 96, 3, 32, 32)


# ---------------------------------------------------------------- transcript
def test_transcript_reference_squeeze_vector():
    t = O.TranscriptState()
    t.absorb(bytes((i * 31) & 255 for i in range(513)))
    words = t.squeeze_u32(40, P)
    assert words == V["transcript_squeeze_513"]["words"]
    assert words[:8] == t.squeeze_u32(8, P)  # squeezing reads a copy


def test_transcript_squeeze_agrees_with_hashlib_digest_words():
    rng = random.Random(2)
    for _ in range(20):
        msg = bytes(rng.randrange(256) for _ in range(rng.randrange(300)))
        t = O.TranscriptState()
        t.absorb(msg)
        d = np.frombuffer(hashlib.sha3_256(msg).digest(), "<u4")
        m = rng.choice([P, 1 << 20, 3, 1000003])
        limit = (1 << 32) - (1 << 32) % m
        expect = [int(w) % m for w in d if w < limit]
        assert t.squeeze_u32(len(expect), m) == expect
        assert t.squeeze_bytes(32) == hashlib.sha3_256(msg).digest()


# ---------------------------------------------------------------- fields
def _qm_ref(a, b):
    """QM31 = F_p[i,u]/(i^2+1, u^2-2-i) by generic polynomial product + reduction."""
    # monomials i^e1 u^e2, e1,e2 in {0,1}: index e1 + 2 e2 -> (a, b, c, d) = (1, i, u, iu)
    prod = {}
    for ea in range(4):
        for eb in range(4):
            e1, e2 = (ea & 1) + (eb & 1), (ea >> 1) + (eb >> 1)
            prod[(e1, e2)] = (prod.get((e1, e2), 0) + a[ea] * b[eb]) % P
    out = [0, 0, 0, 0]
    for (e1, e2), c in prod.items():
        terms = {(e1, e2): c}
        changed = True
        while changed:
            changed = False
            nt = {}
            for (x, y), v in terms.items():
                if y >= 2:  # u^2 = 2 + i
                    nt[(x, y - 2)] = (nt.get((x, y - 2), 0) + 2 * v) % P
                    nt[(x + 1, y - 2)] = (nt.get((x + 1, y - 2), 0) + v) % P
                    changed = True
                elif x >= 2:  # i^2 = -1
                    nt[(x - 2, y)] = (nt.get((x - 2, y), 0) - v) % P
                    changed = True
                else:
                    nt[(x, y)] = (nt.get((x, y), 0) + v) % P
            terms = nt
        for (x, y), v in terms.items():
            out[x + 2 * y] = (out[x + 2 * y] + v) % P
    return tuple(out)


def test_qm31_reference_vectors_and_bruteforce():
    q = V["qm31"]
    assert O.qm31_mul(q["lhs"], q["rhs"]) == tuple(q["product"])
    assert O.qm31_square(q["lhs"]) == tuple(q["square_lhs"])
    assert O.qm31_mul((0, 0, 1, 0), (0, 0, 1, 0)) == (2, 1, 0, 0)  # u^2 = 2 + i
    rng = random.Random(4)
    edge = [0, 1, 2, P - 1, P - 2, 1 << 30, (1 << 31) - 2]
    for _ in range(300):
        a = tuple(rng.choice(edge + [rng.randrange(P)]) for _ in range(4))
        b = tuple(rng.choice(edge + [rng.randrange(P)]) for _ in range(4))
        assert O.qm31_mul(a, b) == _qm_ref(a, b)
        assert O.qm31_square(a) == _qm_ref(a, a)
        if any(a):
            assert O.qm31_mul(a, O.qm31_inv(a)) == (1, 0, 0, 0)
    for x in [1, 2, P - 1, 12345]:
        assert O.m31_mul(x, O.m31_inv(x)) == 1
        assert O.m31_mul(x, P - 1) == (P - x) % P


def test_fold_reference_vector_and_identity():
    f = V["fold_pair"]
    vals = np.array([f["positive"], f["negative"]], np.uint32)
    assert O.fold_layer(vals, np.array([f["twiddle"]], np.uint32), f["challenge"]).tolist() == f["folded"]
    x = np.array([[9, 8, 7, 6], [9, 8, 7, 6]], np.uint32)
    assert O.fold_layer(x, np.array([77], np.uint32), (9, 7, 5, 3)).tolist() == [9, 8, 7, 6]


# ---------------------------------------------------------------- Circle domain
def test_circle_domain_reference_vectors():
    cd = V["circle_domain"]
    assert [O.point_index(3, i) >> 27 for i in range(8)] == cd["log3_natural_indices_over_2^27"]
    assert sha3(O.domain_points(5).astype("<u4").tobytes()) == cd["log5_points_sha3"]
    assert sha3(O.fft_twiddles(5).astype("<u4").tobytes()) == cd["log5_fft_twiddles_sha3"]
    layers = O.fold_twiddles(5, 4)
    q = np.zeros((layers.size, 4), np.uint32)
    q[:, 0] = layers
    assert sha3(q.tobytes()) == cd["log5_inverse_layers_4rounds_sha3"]


def test_circle_group_invariants():
    g = O.point_from_index(1)
    assert g == (2, 1268011823)
    assert O.point_from_index(1 << 30) == (P - 1, 0)  # order-2 element
    assert O.point_from_index(0) == (1, 0)
    for log_n in [1, 3, 7, 10]:
        pts = O.domain_points(log_n).reshape(-1, 2).astype(np.int64)
        assert np.all((pts[:, 0] * pts[:, 0] + pts[:, 1] * pts[:, 1]) % P == 1)
        assert len({tuple(p) for p in pts.tolist()}) == len(pts)
        # fold pairs (2k, 2k+1) are conjugate points (x, y), (x, -y)
        assert np.all(pts[0::2, 0] == pts[1::2, 0])
        assert np.all((pts[0::2, 1] + pts[1::2, 1]) % P == 0)


def test_descriptor_bytes():
    d = O.domain_descriptor(21)
    assert d[:8] == b"AZKCDV1\x00"
    assert np.frombuffer(d[8:], "<u4").tolist() == [1, P, 31, 2, 1268011823, 21, 1 << 9, 20, 1]


# ---------------------------------------------------------------- codeword
def _rand_qm(rng, n):
    return [[rng.randrange(P) for _ in range(4)] for _ in range(n)]


@pytest.mark.parametrize("log_n", [1, 2, 3, 4, 5, 6, 7])
def test_fft_equals_direct_evaluation_bruteforce(log_n):
    rng = random.Random(100 + log_n)
    half = 1 << (log_n - 1)
    for nx, ny in [(half, half), (1, 0), (0, 1), (half, 1), (max(1, half // 2), half)]:
        x, y = _rand_qm(rng, nx), _rand_qm(rng, ny)
        direct = O.evaluate_direct(x, y, log_n)
        basis = O.basis_from_monomial(x, y, log_n)
        assert np.array_equal(O.circle_fft(basis, log_n), direct)


def _m31vec(count, salt):  # main.swift:1316-1323
    i = np.arange(count, dtype=np.uint64)
    v = (i + 1) * np.uint64(1048573) + np.uint64(salt) * np.uint64(65537) + i * np.uint64(17) * np.uint64(salt | 1)
    return (v % np.uint64(P)).astype(np.uint32)


def _qm31vec(count, a, b, c, d):  # main.swift:1336-1350
    return np.stack([_m31vec(count, a), _m31vec(count, b), _m31vec(count, c), _m31vec(count, d)], 1)


@pytest.mark.parametrize("log_n,key", [(14, "codeword_2^14"), (10, "codeword_2^10")])
def test_codeword_prover_baselines(log_n, key):
    n = 1 << log_n  # makeDeterministicCircleCodewordPolynomial main.swift:1376-1395
    x = _qm31vec(min(8, n // 2), 0xE01, 0xE03, 0xE07, 0xE0B)
    y = _qm31vec(min(7, n // 2), 0xE11, 0xE17, 0xE1D, 0xE23)
    cw = O.evaluate_direct(x, y, log_n)
    assert sha3(cw.astype("<u4").tobytes()) == V["proofs"][key]["codeword"]
    assert np.array_equal(O.circle_fft(O.basis_from_monomial(x, y, log_n), log_n), cw)
    digest = bytes(0xA0 + i for i in range(32))  # main.swift:4191-4193
    proof, _ = O.fri_prove(cw.reshape(-1, 4), log_n, (2, 4, 1, 0), digest, 3)
    assert len(proof) == V["proofs"][key]["proof_bytes"]
    assert sha3(proof) == V["proofs"][key]["proof"]
    assert O.fri_verify(proof, digest) == 1


# ---------------------------------------------------------------- fold baselines
def test_fold_baselines():
    fb = V["fold_baselines"]
    ev = _qm31vec(1 << 14, 0xC31, 0xC37, 0xC3D, 0xC43)
    assert sha3(O.fold_chain(ev, 14, [(41, 43, 47, 53)]).astype("<u4").tobytes()) == fb["first_fold_2^14"]
    ev = _qm31vec(1 << 14, 0xCC1, 0xCC7, 0xCCD, 0xCD3)
    chs = [(41 + 18 * r, 43 + 22 * r, 47 + 26 * r, 53 + 30 * r) for r in range(3)]
    assert sha3(O.fold_chain(ev, 14, chs).astype("<u4").tobytes()) == fb["explicit_chain_2^14_r3"]
    digest = bytes(0x90 + i for i in range(32))
    proof, meta = O.fri_prove(ev, 14, (2, 4, 1, 0), digest, 3)
    fin = O.fold_chain(ev, 14, meta["challenges"])
    assert sha3(fin.astype("<u4").tobytes()) == fb["merkle_transcript_chain_2^14_r3_final_layer"]
    assert meta["challenges"] == O.transcript_challenges(14, (2, 4, 1, 0), digest, meta["commitments"])


# ---------------------------------------------------------------- proofs
def _stable(count):  # CircleDomainTests.swift:2825-2837
    i = np.arange(count)
    return np.stack([17 + i * 3, 29 + i * 5, 31 + i * 7, 43 + i * 11], 1).astype(np.uint32)


def test_proof_reference_digests_and_verifier():
    d1 = bytes(0x10 + i for i in range(32))
    p1, _ = O.fri_prove(_stable(16), 4, (2, 3, 1, 0), d1, 1)
    assert sha3(p1) == V["proofs"]["first_fold_log4"]
    assert O.fri_verify(p1, d1) == 1
    d2 = bytes(0x50 + i for i in range(32))
    p2, _ = O.fri_prove(_stable(32), 5, (2, 5, 1, 0), d2, 3)
    assert sha3(p2) == V["proofs"]["three_round_log5"]
    assert O.fri_verify(p2, d2) == 1
    assert O.fri_verify(p2, d1) == 0  # wrong public input
    rng = random.Random(1)
    for _ in range(40):  # any single-byte tamper is rejected or malformed
        bad = bytearray(p2)
        k = rng.randrange(len(bad))
        bad[k] ^= 1 << rng.randrange(8)
        assert O.fri_verify(bytes(bad), d2) != 1
    assert O.fri_verify(p2 + b"\0", d2) != 1


def test_proof_corpus_valid_case():
    corpus = json.load(open(os.path.join(GOLDEN, "circle_pcs_fri_proof_corpus_v1.json")))
    vc = corpus["validCase"]
    x = np.frombuffer(b"".join(bytes.fromhex(h) for h in vc["xCoefficientHex"]), "<u4").reshape(-1, 4)
    y = np.frombuffer(b"".join(bytes.fromhex(h) for h in vc["yCoefficientHex"]), "<u4").reshape(-1, 4)
    log_n = vc["domainLogSize"]
    cw = O.evaluate_direct(x, y, log_n)
    prof = corpus["profile"]
    sec = (prof["logBlowupFactor"], prof["queryCount"], prof["foldingStep"], prof["grindingBits"])
    digest = bytes.fromhex(vc["publicInputDigestHex"])
    proof, _ = O.fri_prove(cw.reshape(-1, 4), log_n, sec, digest, vc["roundCount"], vc["claimedStorageIndices"])
    assert proof == bytes.fromhex(vc["proofHex"])
    assert sha3(proof) == vc["proofDigestHex"]
    assert O.fri_verify(proof, digest) == 1


def test_grinding_roundtrip():
    digest = bytes(range(32))
    proof, _ = O.fri_prove(_stable(32), 5, (2, 3, 1, 6), digest, 2)
    assert O.fri_verify(proof, digest) == 1
    bad = proof[:-8] + (int.from_bytes(proof[-8:], "little") + 1).to_bytes(8, "little")
    assert O.fri_verify(bad, digest) != 1


@pytest.mark.parametrize("bits", [3, 7, 10])
def test_grinding_nonce_is_the_smallest_hit(bits):
    """findGrindingNonce (CircleProofFormat.swift:753-794) counts up from 0: every nonce below
    the proof's misses the target and the proof's meets it.  The target test here is a
    leading-zero count of the digest as a big-endian integer, not the oracle's predicate."""
    digest = bytes(range(32))
    proof, meta = O.fri_prove(_stable(32), 5, (2, 3, 1, bits), digest, 2)
    nonce = meta["grinding_nonce"]
    assert int.from_bytes(proof[-8:], "little") == nonce
    for k in range(nonce + 1):
        d = O.grinding_digest(meta["pre_grinding"], bits, k)
        lead = 256 - int.from_bytes(d, "big").bit_length()
        assert (lead >= bits) == (k == nonce), k
        assert O.digest_meets_target(d, bits) == (k == nonce)


# ---------------------------------------------------------------- sum-check
def test_sumcheck_chunk_reference_vector():
    sc = V["sumcheck_chunk"]
    fin, coeffs, ch = O.sumcheck_chunk([(i + 1) * 17 for i in range(16)], 3)
    assert fin == sc["final"] and ch == sc["challenges"] and coeffs == sc["coefficients"]


def test_sumcheck_multilinear_self_consistency():
    rng = random.Random(8)
    for k in [1, 2, 5, 9]:
        ev = [rng.randrange(P) for _ in range(1 << k)]
        r = O.sumcheck_multilinear(ev)
        assert r["claimed_sum"] == sum(ev) % P
        claim = r["claimed_sum"]
        for (e0, e1), c in zip(r["round_polys"], r["challenges"]):
            assert (e0 + e1) % P == claim
            claim = ((1 - c) * e0 + c * e1) % P
        assert r["final"] == claim
        # final = multilinear extension at the challenges (bit j of the index <-> round j)
        mle = 0
        for idx, v in enumerate(ev):
            w = 1
            for j, c in enumerate(r["challenges"]):
                w = w * (c if (idx >> j) & 1 else (1 - c)) % P
            mle = (mle + v * w) % P
        assert r["final"] == mle
        words = np.array(ev, "<u4").tobytes()
        dom = b"AppleZKProver.M31Sumcheck.Words.V1"
        framed = len(dom).to_bytes(4, "little") + dom + (1).to_bytes(4, "little") + len(ev).to_bytes(8, "little") + words
        assert r["digest"] == hashlib.sha3_256(framed).digest()
