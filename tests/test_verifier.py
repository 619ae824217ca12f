"""The product's host verifier of Circle PCS proofs (azk_circle_verify) against proofs whose
bytes are pinned by the reference: the first-fold and three-round proofs
(CircleDomainTests.swift:702-779), the proof corpus valid case with claimed openings
(Resources/CirclePCSFRIProofCorpusV1.json), a grinding proof, and the device prover's own
proofs.  Every tampered proof must be refused, and the verifier's decision must equal the
oracle verifier's."""
import hashlib
import json
import os
import random

import numpy as np
import pytest

import pyoracle as O
from apple_amd import azk, inputs

HERE = os.path.dirname(__file__)
V = json.load(open(os.path.join(HERE, "golden", "reference_vectors.json")))


def stable(count):  # makeStableCircleEvaluations, CircleDomainTests.swift:2825-2837
    i = np.arange(count)
    return np.stack([17 + i * 3, 29 + i * 5, 31 + i * 7, 43 + i * 11], 1).astype(np.uint32)


def reference_proofs():
    d1 = bytes(0x10 + i for i in range(32))
    p1, _ = O.fri_prove(stable(16), 4, (2, 3, 1, 0), d1, 1)
    d2 = bytes(0x50 + i for i in range(32))
    p2, _ = O.fri_prove(stable(32), 5, (2, 5, 1, 0), d2, 3)
    assert hashlib.sha3_256(p1).hexdigest() == V["proofs"]["first_fold_log4"]
    assert hashlib.sha3_256(p2).hexdigest() == V["proofs"]["three_round_log5"]
    return [(p1, d1), (p2, d2)]


def corpus_proof():
    corpus = json.load(open(os.path.join(HERE, "golden", "circle_pcs_fri_proof_corpus_v1.json")))
    vc = corpus["validCase"]
    return bytes.fromhex(vc["proofHex"]), bytes.fromhex(vc["publicInputDigestHex"])


def corpus_tail_start():
    """Byte offset of the corpus proof's claimed-evaluation-openings section: the same
    proof built without claimed openings is exactly the prefix before it."""
    corpus = json.load(open(os.path.join(HERE, "golden", "circle_pcs_fri_proof_corpus_v1.json")))
    vc, prof = corpus["validCase"], corpus["profile"]
    x = np.frombuffer(b"".join(bytes.fromhex(h) for h in vc["xCoefficientHex"]), "<u4").reshape(-1, 4)
    y = np.frombuffer(b"".join(bytes.fromhex(h) for h in vc["yCoefficientHex"]), "<u4").reshape(-1, 4)
    cw = O.evaluate_direct(x, y, vc["domainLogSize"])
    sec = (prof["logBlowupFactor"], prof["queryCount"], prof["foldingStep"], prof["grindingBits"])
    plain, _ = O.fri_prove(cw.reshape(-1, 4), vc["domainLogSize"], sec, bytes.fromhex(vc["publicInputDigestHex"]),
                           vc["roundCount"])
    full = bytes.fromhex(vc["proofHex"])
    assert full.startswith(plain)
    return len(plain)


def decision(proof, digest):
    """1 accepted, 0 refused, -1 malformed (the reference's decode throws)."""
    try:
        return 1 if azk.circle_verify(proof, digest) else 0
    except azk.AZKError as e:
        assert e.code == 6
        return -1


def test_accepts_reference_proofs():
    for proof, digest in reference_proofs() + [corpus_proof()]:
        assert azk.circle_verify(proof, digest)


def test_grinding_proof():
    digest = bytes(range(32))
    proof, _ = O.fri_prove(stable(32), 5, (2, 3, 1, 6), digest, 2)
    assert azk.circle_verify(proof, digest)
    bad = proof[:-8] + (int.from_bytes(proof[-8:], "little") + 1).to_bytes(8, "little")
    assert decision(bad, digest) != 1
    assert decision(proof[:-8], digest) == -1  # grinding bits without a nonce


def test_wrong_digest_and_layout():
    proof, digest = reference_proofs()[1]
    assert not azk.circle_verify(proof, bytes(32))
    assert decision(proof + b"\0", digest) == -1
    assert decision(proof[:-1], digest) == -1
    assert decision(b"", digest) == -1


@pytest.mark.parametrize("which", [0, 1, 2])
def test_bit_flips_refused_like_the_oracle(which):
    """Every flip is refused, except inside the corpus proof's claimed-openings section,
    which CirclePCSFRIProofVerifierV1 decodes but does not check (the polynomial verifier
    does, CircleProofFormat.swift:1548-1600); there the oracle accepts as well."""
    proof, digest = (reference_proofs() + [corpus_proof()])[which]
    tail = corpus_tail_start() if which == 2 else len(proof)
    rng = random.Random(which)
    for _ in range(60):
        bad = bytearray(proof)
        k = rng.randrange(len(bad))
        bad[k] ^= 1 << rng.randrange(8)
        d = decision(bytes(bad), digest)
        assert (O.fri_verify(bytes(bad), digest) == 1) == (d == 1), k
        if k < tail:
            assert d != 1, k


def test_random_shapes():
    rng = np.random.default_rng(4)
    for log_n, rounds, sec in [(6, 3, (2, 4, 1, 0)), (9, 5, (1, 7, 1, 0)), (12, 10, (3, 9, 1, 0))]:
        ev = rng.integers(0, O.P, size=(1 << log_n, 4), dtype=np.uint32)
        digest = bytes(rng.integers(0, 256, 32, dtype=np.uint8))
        proof, _ = O.fri_prove(ev, log_n, sec, digest, rounds)
        assert azk.circle_verify(proof, digest)


@pytest.mark.gpu
def test_device_proofs_verify():
    ctx = azk.Context(0)
    digest = bytes(0xA0 + i for i in range(32))
    for log_n, rounds in [(10, 3), (16, 10), (21, 10)]:
        print(f"device proof 2^{log_n}", flush=True)
        c = inputs.circle_fft_coefficients(seed=log_n, log_n=log_n)
        p = azk.CircleCodewordPCSFRIProverV1(ctx, log_n, rounds, (2, 32, 1, 0), digest)
        proof = p.prove_fft_coefficients_resident(ctx.upload(c))
        assert azk.circle_verify(proof, digest)
        assert not azk.circle_verify(proof, bytes(32))
    ctx.close()
