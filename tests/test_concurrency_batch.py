"""Shared plans under concurrent callers, batches that share provers, the batched form at its
benchmark shape, and the autotuner on a batch program.

- testGPURawLeafMerklePlanConcurrentReuseIsSerialized (Tests/AppleZKProverTests/MerkleTests.swift:
  538-585): 8 tasks commit different leaves through one Merkle plan, each root equal to the oracle's;
  here also 8 threads proving through one Circle PCS prover.
- Two threads run batches over the same provers listed in opposite orders (the provers' locks are
  taken in one order, so neither can wait on the other forever).
- 64 commits with log N drawn from 14..21 (seed 4242, the benchmark's batched shape): every batched
  proof equals its plain proof, and the proofs of 2^17 and smaller equal the oracle's.
"""
import threading

import numpy as np
import pytest

import pyoracle as O
from apple_amd import azk, inputs

pytestmark = pytest.mark.gpu
SEC = (2, 32, 1, 0)
DIGEST = bytes(0xA0 + i for i in range(32))


@pytest.fixture(scope="module")
def ctx():
    c = azk.Context(0)
    yield c
    c.close()


def run_threads(fn, n):
    errors = []

    def wrap(k):
        try:
            fn(k)
        except Exception as e:  # noqa: BLE001 -- reported below
            errors.append((k, repr(e)))

    ts = [threading.Thread(target=wrap, args=(k,)) for k in range(n)]
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout=90)
    assert not any(t.is_alive() for t in ts), "a thread did not finish"
    assert not errors, errors


def test_merkle_plan_shared_by_eight_threads(ctx):
    print("merkle plan, 8 threads", flush=True)
    count, length = 256, 128
    plan = azk.SHA3RawLeavesMerkleCommitPlan(ctx, count, length, length)
    leaves = [inputs.random_bytes(53 * k + 1, count * length) for k in range(8)]
    want = [O.merkle_root(lv, count, length, length) for lv in leaves]
    bufs = [ctx.upload(lv) for lv in leaves]

    def task(k):
        for _ in range(6):
            assert plan.commit(leaves[k]) == want[k]
            assert plan.commit_resident(bufs[k]) == want[k]

    run_threads(task, 8)
    plan.close()


def test_prover_shared_by_eight_threads(ctx):
    print("circle prover, 8 threads", flush=True)
    log_n = 13
    prover = azk.CircleCodewordPCSFRIProverV1(ctx, log_n, 8, SEC, DIGEST)
    coeffs = [inputs.circle_fft_coefficients(seed=60 + k, log_n=log_n) for k in range(8)]
    bufs = [ctx.upload(c) for c in coeffs]
    want = [prover.prove_fft_coefficients_resident(b) for b in bufs]
    assert want[0] == O.fri_prove(O.circle_fft(coeffs[0], log_n).reshape(-1, 4), log_n, SEC, DIGEST, 8)[0]

    def task(k):
        for r in range(4):
            if r % 2:
                assert prover.prove_fft_coefficients(coeffs[k]) == want[k]
            else:
                assert prover.prove_fft_coefficients_resident(bufs[k]) == want[k]

    run_threads(task, 8)
    prover.close()


def test_batches_sharing_provers_in_opposite_orders(ctx):
    print("two batches, opposite prover orders", flush=True)
    logs = [10, 12, 11, 13]
    provers = [azk.CircleCodewordPCSFRIProverV1(ctx, lg, lg - 3, SEC, DIGEST) for lg in logs]
    bufs = [ctx.upload(inputs.circle_fft_coefficients(seed=80 + k, log_n=lg)) for k, lg in enumerate(logs)]
    want = [p.prove_fft_coefficients_resident(b) for p, b in zip(provers, bufs)]
    fwd = azk.CircleBatchPlan(provers, bufs)
    rev = azk.CircleBatchPlan(provers[::-1], bufs[::-1])

    def task(k):
        for _ in range(10):
            if k == 0:
                assert fwd.prove() == want
            else:
                assert rev.prove() == want[::-1]

    run_threads(task, 2)
    fwd.close()
    rev.close()


def test_batch_of_64_jobs_log_14_to_21(ctx):
    rng = np.random.default_rng(4242)
    logs = [int(x) for x in rng.integers(14, 22, size=64)]
    print("batch of 64 commits, log N", min(logs), "..", max(logs), flush=True)
    provers, bufs, coeffs = [], [], []
    for j, ln in enumerate(logs):
        provers.append(azk.CircleCodewordPCSFRIProverV1(ctx, ln, 10, SEC, DIGEST))
        coeffs.append(inputs.circle_fft_coefficients(seed=4242 + j, log_n=ln))
        bufs.append(ctx.upload(coeffs[-1]))
    plain = [p.prove_fft_coefficients_resident(b) for p, b in zip(provers, bufs)]
    plan = azk.CircleBatchPlan(provers, bufs)
    assert plan.prove() == plain
    assert plan.prove() == plain  # graph replay
    # the same batch from the M31 words: the same proofs
    m31 = [ctx.upload(np.ascontiguousarray(c[:, 0])) for c in coeffs]
    plan_m31 = azk.CircleBatchPlan(provers, m31, m31=True)
    assert plan_m31.prove() == plain
    assert plan_m31.prove() == plain
    plan_m31.close()
    for b in m31:
        b.free()
    checked = 0
    for j, ln in enumerate(logs):
        if ln <= 17:
            cw = O.circle_fft(coeffs[j], ln)
            assert plain[j] == O.fri_prove(cw.reshape(-1, 4), ln, SEC, DIGEST, 10)[0], j
            checked += 1
    assert checked >= 16
    plan.close()
    for p, b in zip(provers, bufs):
        p.close()
        b.free()


def test_batch_program_tuning(ctx, tmp_path):
    print("batch tune", flush=True)
    logs = [10, 13, 11, 12, 10]
    provers = [azk.CircleCodewordPCSFRIProverV1(ctx, lg, lg - 2, (2, 8, 1, 0), DIGEST) for lg in logs]
    bufs = [ctx.upload(inputs.circle_fft_coefficients(seed=90 + k, log_n=lg)) for k, lg in enumerate(logs)]
    want = [p.prove_fft_coefficients_resident(b) for p, b in zip(provers, bufs)]
    plan = azk.CircleBatchPlan(provers, bufs)
    db = str(tmp_path / "plans.jsonl")
    r = plan.tune(database_path=db, warmup_runs=1, measured_runs=2, max_candidates=8)
    assert r["candidates_measured"] >= 2 and r["from_database"] == 0
    assert r["median_gpu_ns"] <= r["default_median_gpu_ns"]
    assert plan.prove() == want                       # the winner's program, same proofs
    assert [p.exec_plan() for p in provers] == [r["winner"]] * len(provers)
    assert [p.prove_fft_coefficients_resident(b) for p, b in zip(provers, bufs)] == want
    again = azk.CircleBatchPlan(provers, bufs).tune(database_path=db)
    assert again["from_database"] == 1 and again["winner"] == r["winner"]
    assert "circle_pcs_batch" in open(db).read()
    plan.close()
