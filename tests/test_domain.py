"""Circle domain materialization plan (azk_domain_*) against the reference's stable vector
digests and the oracle.

Mirrors testCircleDomainAndTwiddleStableVectorDigests (Tests/AppleZKProverTests/
CircleDomainTests.swift:129-148): SHA3-256 of the packed log N = 5 domain points, FFT twiddles
and four inverse-domain layers; beyond it, every output equals the oracle's for log N 1..20,
points lie on the circle, and the layout and error cases of the plan
(CircleDomainMaterializationPlan.swift:54-216)."""
import hashlib
import json
import os

import numpy as np
import pytest

import pyoracle as O
from apple_amd import azk

V = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.json")))["circle_domain"]
P = O.P


@pytest.fixture(scope="module")
def ctx():
    c = azk.Context(0)
    yield c
    c.close()


def sha3(b: bytes) -> str:
    return hashlib.sha3_256(b).hexdigest()


@pytest.mark.gpu
def test_reference_digests_log5(ctx):
    print("domain materialization 2^5", flush=True)
    plan = azk.CircleDomainMaterializationPlan(ctx, 5, True, True, 4)
    assert sha3(plan.read_domain_points().astype("<u4").tobytes()) == V["log5_points_sha3"]
    assert sha3(plan.read_codeword_twiddles().astype("<u4").tobytes()) == V["log5_fft_twiddles_sha3"]
    assert sha3(plan.read_flat_inverse_domain().astype("<u4").tobytes()) == V["log5_inverse_layers_4rounds_sha3"]
    counts, offsets = plan.inverse_domain_layout()
    assert counts == [16, 8, 4, 2] and offsets == [0, 16, 24, 28]
    layers = plan.read_inverse_domain_layers()
    assert [len(x) for x in layers] == counts
    plan.close()


@pytest.mark.gpu
@pytest.mark.parametrize("log_n", [1, 2, 3, 7, 12, 16, 20])
def test_outputs_match_oracle(ctx, log_n):
    print(f"domain materialization 2^{log_n}", flush=True)
    rounds = min(log_n, 10)
    plan = azk.CircleDomainMaterializationPlan(ctx, log_n, True, True, rounds)
    pts = plan.read_domain_points()
    assert np.array_equal(pts.reshape(-1), O.domain_points(log_n))
    x, y = pts[:, 0].astype(np.uint64), pts[:, 1].astype(np.uint64)
    assert np.all((x * x + y * y) % P == 1)
    assert np.array_equal(plan.read_codeword_twiddles(), O.fft_twiddles(log_n))
    inv = plan.read_flat_inverse_domain()
    assert np.array_equal(inv[:, 0], O.fold_twiddles(log_n, rounds)) and not inv[:, 1:].any()
    # each layer value times its FFT twiddle is one
    tw = O.fft_twiddles(log_n).astype(np.uint64)
    assert np.all(inv[:, 0].astype(np.uint64) * tw[: len(inv)] % P == 1)
    assert plan.materialization_stats()["gpu_ns"] > 0
    plan.close()


@pytest.mark.gpu
def test_partial_outputs_and_resident_copy(ctx):
    plan = azk.CircleDomainMaterializationPlan(ctx, 9, False, False, 3)
    with pytest.raises(azk.AZKError) as e:
        plan.read_domain_points()
    assert e.value.code == 6
    with pytest.raises(azk.AZKError):
        plan.read_codeword_twiddles()
    n_inv = 512 - 64
    buf = ctx.buffer(16 * n_inv + 32)
    plan.copy_resident(plan.INVERSE_LAYERS, buf, 32)
    got = buf.download(16 * n_inv, 32).view(np.uint32).reshape(-1, 4)
    assert np.array_equal(got, plan.read_flat_inverse_domain())
    with pytest.raises(azk.AZKError) as e:  # range past the buffer
        plan.copy_resident(plan.INVERSE_LAYERS, buf, 48)
    assert e.value.code == 6
    buf.free()
    plan.close()


@pytest.mark.gpu
@pytest.mark.parametrize("log_n,points,tw,rounds", [(0, 1, 0, 0), (31, 1, 0, 0), (5, 0, 0, 0), (5, 1, 0, 6)])
def test_invalid_plans(ctx, log_n, points, tw, rounds):
    with pytest.raises(azk.AZKError) as e:
        azk.CircleDomainMaterializationPlan(ctx, log_n, bool(points), bool(tw), rounds)
    assert e.value.code == 6
