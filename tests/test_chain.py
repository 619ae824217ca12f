"""The Merkle-transcript FRI fold chain plan (azk_chain_*, CircleFRIMerkleTranscriptFoldChainPlan
CircleFRIFoldPlan.swift:301-649) against the reference's 2^14 / 3-round final-layer digest and
against the oracle's commitments, challenges and folds; the verified form runs its host gate."""
import hashlib
import json
import os

import numpy as np
import pytest

import pyoracle as O
from apple_amd import azk

P = O.P
V = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.json")))


def _m31vec(count, salt):  # makeDeterministicM31Vector, main.swift:1336-1350
    i = np.arange(count, dtype=np.uint64)
    v = (i + 1) * np.uint64(1048573) + np.uint64(salt) * np.uint64(65537) + i * np.uint64(17) * np.uint64(salt | 1)
    return (v % np.uint64(P)).astype(np.uint32)


def _qm31vec(count, a, b, c, d):
    return np.stack([_m31vec(count, a), _m31vec(count, b), _m31vec(count, c), _m31vec(count, d)], 1)


@pytest.fixture(scope="module")
def ctx():
    c = azk.Context(0)
    yield c
    c.close()


@pytest.mark.gpu
def test_chain_reference_final_layer(ctx):
    ev = _qm31vec(1 << 14, 0xCC1, 0xCC7, 0xCCD, 0xCD3)
    digest = bytes(0x90 + i for i in range(32))
    sec = (2, 4, 1, 0)
    plan = azk.CircleFRIMerkleTranscriptFoldChainPlan(ctx, 14, 3, sec, digest)
    got = plan.execute(ev, verified=True)
    fb = V["fold_baselines"]
    assert hashlib.sha3_256(got["values"].astype("<u4").tobytes()).hexdigest() == \
        fb["merkle_transcript_chain_2^14_r3_final_layer"]
    _, meta = O.fri_prove(ev, 14, sec, digest, 3)
    assert got["commitments"] == meta["commitments"]
    assert [tuple(int(v) for v in c) for c in got["challenges"]] == meta["challenges"]


@pytest.mark.gpu
@pytest.mark.parametrize("log_n,rounds,sec", [(3, 2, (1, 3, 1, 0)), (10, 6, (2, 8, 1, 0)), (16, 10, (2, 32, 1, 0)),
                                              (19, 9, (3, 20, 1, 4))])
def test_chain_matches_oracle(ctx, log_n, rounds, sec):
    print(f"chain 2^{log_n} r{rounds} {sec}", flush=True)
    rng = np.random.default_rng(log_n)
    ev = rng.integers(0, P, size=(1 << log_n, 4), dtype=np.uint32)
    digest = bytes(rng.integers(0, 256, 32, dtype=np.uint8))
    plan = azk.CircleFRIMerkleTranscriptFoldChainPlan(ctx, log_n, rounds, sec, digest)
    got = plan.execute(ev, verified=log_n <= 16)
    _, meta = O.fri_prove(ev, log_n, sec, digest, rounds)
    assert got["commitments"] == meta["commitments"]
    assert [tuple(int(v) for v in c) for c in got["challenges"]] == meta["challenges"]
    assert np.array_equal(got["values"], O.fold_chain(ev, log_n, meta["challenges"]))
    buf = ctx.upload(np.vstack([np.zeros((1, 4), np.uint32), ev]))
    out = azk.Buffer(ctx, 16 << (log_n - rounds))
    res = plan.execute_resident(buf, out, evals_offset=16)
    assert res["commitments"] == got["commitments"]
    assert np.array_equal(out.download().view(np.uint32), got["values"])
