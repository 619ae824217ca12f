"""Statement-level verification on the host: the polynomial-claim digest, the polynomial
verifier and the contract verifier (azk_circle_polynomial_claim_digest,
azk_circle_verify_polynomial, azk_circle_verify_contract), against the reference's proof
corpora.

Mirrors testCirclePCSProofCorpusV1PinsCanonicalBytesDigestsAndRejections
(Tests/AppleZKProverTests/CircleDomainTests.swift:1108-1166): the corpus statement's claim digest
equals its public-input digest, the valid proof is accepted, every tamper vector is refused by
the contract verifier, and the developer-parameter proof is still accepted by the polynomial
verifier.  The Circle PCS proof embedded in Resources/ApplicationProofCorpusV1.json
(CircleDomainTests.swift:1168-1194) is a second statement of the same kind.  Also: the
grinding predicate's reference vectors (CircleDomainTests.swift:598-605) and decode bounds on
untrusted counts.
"""
import json
import os
import struct

import numpy as np
import pytest

import pyoracle as O
from apple_amd import azk

HERE = os.path.dirname(__file__)
V = json.load(open(os.path.join(HERE, "golden", "reference_vectors.json")))
CORPUS = json.load(open(os.path.join(HERE, "golden", "circle_pcs_fri_proof_corpus_v1.json")))
APP = json.load(open(os.path.join(HERE, "golden", "application_proof_corpus_v1.json")))
P = O.P


def qm31_list(hexes):
    return np.frombuffer(b"".join(bytes.fromhex(h) for h in hexes), "<u4").reshape(-1, 4)


def corpus_claim():
    vc = CORPUS["validCase"]
    return azk.CirclePCSFRIPolynomialClaimV1(vc["domainLogSize"], qm31_list(vc["xCoefficientHex"]),
                                             qm31_list(vc["yCoefficientHex"]), vc["claimedStorageIndices"])


def corpus_params():
    prof = CORPUS["profile"]
    return dict(log_blowup=prof["logBlowupFactor"], queries=prof["queryCount"], grinding_bits=prof["grindingBits"],
                target_soundness_bits=prof["targetSoundnessBits"])


def test_claim_digest_is_the_corpus_public_input_digest():
    assert corpus_claim().digest().hex() == CORPUS["validCase"]["publicInputDigestHex"]


def test_claim_points_and_values_match_the_oracle():
    rng = np.random.default_rng(5)
    for log_n in (3, 6, 9):
        n = 1 << log_n
        x = rng.integers(0, P, (n // 4, 4), dtype=np.uint32)
        y = rng.integers(0, P, (n // 4 - 1, 4), dtype=np.uint32)
        idx = sorted(rng.choice(n, 7, replace=False).tolist())
        claim = azk.CirclePCSFRIPolynomialClaimV1(log_n, x, y, idx)
        assert np.array_equal(claim.values, O.evaluate_at(x, y, log_n, idx))
        pts = O.domain_points(log_n).reshape(-1, 2)  # storage order
        assert np.array_equal(claim.points, pts[idx])


def test_contract_accepts_the_corpus_valid_case():
    proof = bytes.fromhex(CORPUS["validCase"]["proofHex"])
    claim = corpus_claim()
    assert azk.circle_verify_contract(proof, claim, **corpus_params())
    assert azk.circle_verify_polynomial(proof, claim)


@pytest.mark.parametrize("vector", CORPUS["tamperVectors"], ids=lambda v: v["id"])
def test_contract_rejects_every_tamper_vector(vector):
    import hashlib
    proof = bytes.fromhex(vector["proofHex"])
    assert len(proof) == vector["proofByteCount"]
    assert hashlib.sha3_256(proof).hexdigest() == vector["proofDigestHex"]
    claim = corpus_claim()
    assert vector["expectedVerifierAccepted"] is False
    assert not azk.circle_verify_contract(proof, claim, **corpus_params())
    if vector["id"] == "developer-parameter-proof-rejected-by-contract":
        assert azk.circle_verify_polynomial(proof, claim)


def test_flips_inside_claimed_openings_are_refused():
    """Every byte of the claimed-openings section: a flip is a layout error (non-canonical
    value, wrong index shape) or a refusal, never an acceptance.  The plain proof verifier,
    like CirclePCSFRIProofVerifierV1.verify, does not read that section."""
    proof = bytes.fromhex(CORPUS["validCase"]["proofHex"])
    claim = corpus_claim()
    log_n, count = 6, len(CORPUS["validCase"]["claimedStorageIndices"])
    section = 4 + count * (8 + 16 + 4 + 32 * log_n)  # u32 count, then openings; no grinding nonce
    start = len(proof) - section
    assert struct.unpack_from("<I", proof, start)[0] == count
    digest = bytes.fromhex(CORPUS["validCase"]["publicInputDigestHex"])
    refused = 0
    for off in range(start, len(proof)):
        bad = bytearray(proof)
        bad[off] ^= 0x01
        bad = bytes(bad)
        try:
            ok = azk.circle_verify_contract(bad, claim, **corpus_params())
        except azk.AZKError as e:
            assert e.code == 6
            ok = False
        assert not ok, off
        refused += 1
        rel = off - start - 4
        if rel >= 0 and rel % (8 + 16 + 4 + 32 * log_n) >= 8 + 16 + 4:  # a sibling byte
            assert azk.circle_verify(bad, digest)  # the plain verifier does not look
    assert refused == section


def test_statement_validation_errors():
    vc = CORPUS["validCase"]
    proof = bytes.fromhex(vc["proofHex"])
    claim = corpus_claim()
    params = corpus_params()
    for bad in (dict(params, target_soundness_bits=145),       # above the nominal 4 x 36
                dict(params, log_blowup=6),                     # domain not above the blowup
                dict(params, queries=0)):
        with pytest.raises(azk.AZKError) as e:
            azk.circle_verify_contract(proof, claim, **bad)
        assert e.value.code == 6
    big = azk.CirclePCSFRIPolynomialClaimV1(6, np.ones((3, 4), np.uint32), np.ones((2, 4), np.uint32), [1])
    with pytest.raises(azk.AZKError) as e:  # 5 coefficients over the 2^6 >> 4 = 4 capacity
        azk.circle_verify_contract(proof, big, **params)
    assert e.value.code == 6
    with pytest.raises(azk.AZKError):  # storage indices must increase
        azk.CirclePCSFRIPolynomialClaimV1(6, np.ones((1, 4), np.uint32), [], [5, 5])


def test_claim_with_a_wrong_value_is_refused():
    proof = bytes.fromhex(CORPUS["validCase"]["proofHex"])
    c = corpus_claim()
    values = c.values.copy()
    values[1, 2] = (int(values[1, 2]) + 1) % P
    wrong = azk.CirclePCSFRIPolynomialClaimV1(6, c.x, c.y, c.indices, c.points, values)
    assert not azk.circle_verify_polynomial(proof, wrong)  # the digest moves with the value
    assert not azk.circle_verify_contract(proof, wrong, **corpus_params())


def application_pcs_proof(hexproof: str) -> bytes:
    """The Circle PCS proof inside an ApplicationProofCodecV1 encoding
    (ApplicationProofFormat.swift:411-425): magic, u32 version, 32-byte statement digest,
    u32-length-prefixed sum-check proof, u32-length-prefixed PCS proof."""
    b = bytes.fromhex(hexproof)
    at = 8 + 4 + 32
    (n1,) = struct.unpack_from("<I", b, at)
    at += 4 + n1
    (n2,) = struct.unpack_from("<I", b, at)
    assert at + 4 + n2 == len(b)
    return b[at + 4:]


def app_claim_and_params():
    st = APP["statement"]
    ps = st["parameterSet"]
    claim = azk.CirclePCSFRIPolynomialClaimV1(st["domainLogSize"], qm31_list(st["xCoefficientHex"]),
                                              qm31_list(st["yCoefficientHex"]), st["claimedStorageIndices"])
    return claim, dict(log_blowup=ps["logBlowupFactor"], queries=ps["queryCount"], grinding_bits=ps["grindingBits"],
                       target_soundness_bits=ps["targetSoundnessBits"])


def test_application_corpus_pcs_statement():
    claim, params = app_claim_and_params()
    assert azk.circle_verify_contract(application_pcs_proof(APP["validCase"]["proofHex"]), claim, **params)
    for v in APP["tamperVectors"]:
        pcs = application_pcs_proof(v["proofHex"])
        # only this vector touches the PCS proof; the other two tamper with the statement
        # digest and the sum-check, outside the Circle PCS contract
        assert azk.circle_verify_contract(pcs, claim, **params) == (v["id"] != "pcs-final-layer-nonconstant"), v["id"]


def test_oracle_proof_of_a_claim_verifies_under_the_contract():
    """Prover side: the oracle's proof of a random claim (digest from the product) passes."""
    rng = np.random.default_rng(11)
    log_n, log_blowup = 7, 2
    x = rng.integers(0, P, (16, 4), dtype=np.uint32)
    y = rng.integers(0, P, (16, 4), dtype=np.uint32)
    claim = azk.CirclePCSFRIPolynomialClaimV1(log_n, x, y, [3, 64, 127])
    cw = O.evaluate_direct(x, y, log_n)
    proof, _ = O.fri_prove(cw.reshape(-1, 4), log_n, (log_blowup, 8, 1, 0), claim.digest(), log_n - log_blowup,
                           [3, 64, 127])
    params = dict(log_blowup=log_blowup, queries=8, grinding_bits=0, target_soundness_bits=16)
    assert azk.circle_verify_contract(proof, claim, **params)
    assert not azk.circle_verify_contract(proof, claim, **dict(params, queries=9, target_soundness_bits=16))


# ---------------------------------------------------------------- grinding target
@pytest.mark.parametrize("case", V["grinding_target"]["cases"], ids=str)
def test_grinding_target_reference_vectors(case):
    digest, bits, want = bytes.fromhex(case[0]), case[1], case[2]
    assert azk.grinding_digest_meets_target(digest, bits) is want
    assert O.digest_meets_target(digest, bits) is want


def test_grinding_target_exhaustive_small():
    for digest in [bytes([a, b]) for a in range(256) for b in (0x00, 0x01, 0x7f, 0x80, 0xff)]:
        lead = 16 - int.from_bytes(digest, "big").bit_length()
        for bits in range(0, 18):
            want = bits <= 16 and lead >= bits
            assert azk.grinding_digest_meets_target(digest, bits) is want, (digest.hex(), bits)


# ---------------------------------------------------------------- decode bounds
def header(log_n, rounds, queries, log_blowup=1):
    b = bytearray(b"AZKPFV1\x00") + struct.pack("<II", 1, 1)
    b += struct.pack("<I", 44) + O.domain_descriptor(log_n)
    b += struct.pack("<IIII", log_blowup, queries, 1, 0)
    b += struct.pack("<I", 32) + bytes(32)
    b += struct.pack("<I", rounds) + bytes(32 * rounds)
    return b


def test_untrusted_counts_are_layout_errors():
    # 2^29 final-layer values announced, none present: refused before anything is sized
    crafted = header(30, 1, 0xFFFFFFFF) + struct.pack("<Q", 1 << 29)
    with pytest.raises(azk.AZKError) as e:
        azk.circle_verify(bytes(crafted), bytes(32))
    assert e.value.code == 6
    # 2^32 - 1 queries announced after a complete final layer
    crafted = header(2, 1, 0xFFFFFFFF) + struct.pack("<Q", 2) + bytes(32) + struct.pack("<I", 0xFFFFFFFF)
    for fn in (lambda b: azk.circle_verify(b, bytes(32)),
               lambda b: azk.circle_verify_polynomial(b, corpus_claim())):
        with pytest.raises(azk.AZKError) as e:
            fn(bytes(crafted))
        assert e.value.code == 6
