"""Execution plans and the autotuner: every tile shape gives the same proof bytes, the tuner
gates candidates on those bytes, persists its race and reads the winner back, and drift
is tracked against the tuned median (MetalProofPlanner / PlanDatabase semantics)."""
import json

import numpy as np
import pytest

import pyoracle as O
from apple_amd import azk, inputs

pytestmark = pytest.mark.gpu
SEC = (2, 8, 1, 0)
DIGEST = bytes(0x30 + i for i in range(32))

PLANS = [
    dict(log_tile_commit=9, log_tile_fold=9),
    dict(log_tile_commit=10, log_tile_fold=10, log_tree_chunk=6),
    dict(log_tile_commit=6, log_tile_fold=6, log_tile_leafwise=5, log_tree_chunk=4),
    dict(log_tree_chunk=3),
    dict(log_leafwise_below=21, log_tile_leafwise=6),
    dict(log_leafwise_below=0, log_tile_fold=8),      # every committed layer in the big-tile form
    dict(log_leafwise_below=31, log_tile_leafwise=7),  # every committed layer one leaf per thread
    dict(log_leafwise_grid=11),                        # leafwise tiles shrink with the layer
    dict(log_leafwise_grid=8, log_tile_leafwise=3),    # 8-leaf tiles throughout
    dict(log_leafwise_below=31, log_leafwise_grid=10, log_tree_chunk=3),
    dict(tree_waves=8, log_tree_chunk=4),               # 8-wave upper-tree blocks
    dict(tree_waves=16, log_tree_chunk=5, log_leafwise_grid=6),
]


@pytest.fixture(scope="module")
def ctx():
    c = azk.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("log_n,rounds", [(14, 6), (17, 10)])
def test_every_plan_gives_the_oracle_proof(ctx, log_n, rounds):
    coeffs = inputs.circle_fft_coefficients(seed=log_n, log_n=log_n)
    want, _ = O.fri_prove(O.circle_fft(coeffs, log_n).reshape(-1, 4), log_n, SEC, DIGEST, rounds)
    buf = ctx.upload(coeffs)
    p = azk.CircleCodewordPCSFRIProverV1(ctx, log_n, rounds, SEC, DIGEST)
    assert p.prove_fft_coefficients_resident(buf) == want
    for plan in PLANS:
        print(f"2^{log_n} plan {plan}", flush=True)
        p.set_exec_plan(log_tile_commit=11, log_tile_fold=11, log_leafwise_below=19, log_tile_leafwise=8,
                        log_tree_chunk=10, log_leafwise_grid=0, tree_waves=4)
        p.set_exec_plan(**plan)
        assert p.prove_fft_coefficients_resident(buf) == want, plan
        p.set_graph(False)
        assert p.prove_fft_coefficients_resident(buf) == want, plan
        p.set_graph(True)


def test_invalid_plan_rejected(ctx):
    p = azk.CircleCodewordPCSFRIProverV1(ctx, 10, 3, SEC, DIGEST)
    for bad in (dict(log_tile_commit=12), dict(log_tile_fold=5), dict(log_tile_leafwise=9), dict(log_tile_leafwise=2),
                dict(log_tree_chunk=11), dict(log_leafwise_grid=3), dict(log_leafwise_grid=21), dict(tree_waves=5)):
        with pytest.raises(azk.AZKError) as e:
            p.set_exec_plan(**bad)
        assert e.value.code == 10  # AZK_ERR_INVALID_KERNEL_CONFIG
    assert p.exec_plan()["log_tile_commit"] == 11


def test_batch_mixes_plans(ctx):
    jobs = [(13, 5, 9), (15, 8, 10), (12, 4, 11)]
    provers, bufs, want = [], [], []
    for lg, rounds, tile in jobs:
        c = inputs.circle_fft_coefficients(seed=100 + lg, log_n=lg)
        p = azk.CircleCodewordPCSFRIProverV1(ctx, lg, rounds, SEC, DIGEST)
        b = ctx.upload(c)
        want.append(p.prove_fft_coefficients_resident(b))
        p.set_exec_plan(log_tile_commit=tile, log_tile_fold=tile, log_leafwise_below=14)
        provers.append(p)
        bufs.append(b)
    got, _ = azk.prove_batch(provers, bufs)
    assert got == want


def test_tune_persist_and_reload(ctx, tmp_path):
    lg, rounds = 16, 8
    c = inputs.circle_fft_coefficients(seed=5, log_n=lg)
    buf = ctx.upload(c)
    p = azk.CircleCodewordPCSFRIProverV1(ctx, lg, rounds, SEC, DIGEST)
    want = p.prove_fft_coefficients_resident(buf)
    db = str(tmp_path / "plans.jsonl")
    print("tuning 2^16", flush=True)
    r = p.tune(buf, database_path=db, max_candidates=8, measured_runs=3, warmup_runs=1, validation_runs=2)
    assert r["from_database"] == 0
    assert r["candidates_measured"] >= 2
    assert r["median_gpu_ns"] <= r["default_median_gpu_ns"] * 1.10
    assert p.exec_plan() == r["winner"]
    assert p.prove_fft_coefficients_resident(buf) == want  # tuned plan, same bytes
    rows = [json.loads(line) for line in open(db)]
    assert len(rows) == r["candidates_measured"] - r["candidates_rejected"]
    assert sum(row["is_winner"] for row in rows) >= 1
    assert all(row["workload"]["log_n"] == lg and row["protocol_hash"] == "AppleZKProver.CircleFRI.PCS.V1" for row in rows)

    q = azk.CircleCodewordPCSFRIProverV1(ctx, lg, rounds, SEC, DIGEST)
    r2 = q.tune(buf, database_path=db)
    assert r2["from_database"] == 1
    assert r2["winner"] == r["winner"]
    for _ in range(9):
        assert q.prove_fft_coefficients_resident(buf) == want
    d = q.drift()
    assert d["samples"] == 9 and d["baseline_gpu_ns"] > 0
    assert d["stale"] == (d["relative_drift"] > 0.25)
