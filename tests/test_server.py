"""Serving form: admission by HBM footprint, continuous batching, and determinism (each
job's result alone equals its result in a mix of concurrent jobs)."""
import numpy as np
import pytest

import pyoracle as O
from apple_amd import azk, inputs, serving


def test_trace_is_seeded_and_poisson():
    a = serving.make_trace(5, 400, rate=100.0)
    assert a == serving.make_trace(5, 400, rate=100.0)
    gaps = np.diff([0.0] + [t for t, _, _ in a])
    assert 0.007 < gaps.mean() < 0.013          # mean inter-arrival 1 / rate
    kinds = [k for _, k, _ in a]
    assert 160 < kinds.count("commit") < 240     # half commits, half sum-checks
    assert kinds.count("commit") + kinds.count("sumcheck") == 400
    assert all(12 <= lg <= 22 for _, k, lg in a if k == "commit")
    assert all(16 <= lg <= 22 for _, k, lg in a if k == "sumcheck")
    assert all(t == 0.0 for t, _, _ in serving.make_trace(5, 10, rate=0))
    b = serving.make_trace(6, 400, rate=0, commit_share=0.4, fold_share=0.2, chunk_share=0.1)
    kb = [k for _, k, _ in b]
    assert 50 < kb.count("fold") < 110 and 15 < kb.count("chunk") < 65


def test_words_digest_matches_oracle():
    for n in (1, 1000, 4096):
        w = inputs.random_m31(9 + n, n)
        ref = np.zeros(32, np.uint8)
        O.lib().ora_sumcheck_words_digest(O._p32(np.ascontiguousarray(w, np.uint32)), w.size, O._p8(ref))
        assert serving.words_digest(w) == ref.tobytes()


@pytest.fixture(scope="module")
def ctx():
    c = azk.Context(0)
    yield c
    c.close()


def small_trace():
    return [(0.0, "commit", 10), (0.0, "sumcheck", 12), (0.0, "commit", 13), (0.0, "fold", 12),
            (0.0, "commit", 10), (0.0, "sumcheck", 9), (0.0, "commit", 11), (0.0, "sumcheck", 14),
            (0.0, "fold", 14), (0.0, "commit", 12), (0.0, "commit", 14), (0.0, "sumcheck", 12),
            (0.0, "fold", 12), (0.0, "commit", 9), (0.0, "commit", 13)]


def alone(ctx, kind, lg, data):
    if kind == "commit":
        p = azk.CircleCodewordPCSFRIProverV1(ctx, lg, serving.commit_rounds(lg), serving.SECURITY, serving.DIGEST)
        return p.prove_fft_coefficients(data.get(kind, lg))
    if kind == "fold":
        f = azk.CircleFRIFoldChainPlan(ctx, lg, serving.FOLD_ROUNDS)
        return f.execute(data.get(kind, lg), serving.fold_challenges()).tobytes()
    if kind == "chunk":
        return azk.MetalSumcheckChunkPlan(ctx, 1 << lg, serving.CHUNK_ROUNDS).prove(data.get(kind, lg))
    return azk.M31MultilinearSumcheckPlan(ctx, lg).prove(data.get(kind, lg))


@pytest.mark.gpu
@pytest.mark.parametrize("supply_digests", [False, True])
def test_server_timeline_overlap_matches_plain_runs(ctx, supply_digests):
    """Jobs arrive on a timeline while earlier ones are still on the device (sum-checks of
    2^16..2^18 words, chunks, commits and folds); every result equals the same job run alone
    by the plain runner, and the server ran jobs concurrently."""
    trace = serving.make_trace(21, 40, rate=800.0, commit_logs=(12, 17), sumcheck_logs=(16, 18), fold_logs=(14, 16),
                               commit_share=0.4, fold_share=0.15, chunk_share=0.15, chunk_logs=(12, 16))
    assert {k for _, k, _ in trace} == {"commit", "sumcheck", "fold", "chunk"}
    data = serving.Inputs(trace, seed=11)
    plain = serving.PlainRunner(ctx, trace, data)
    want = plain.run([(0.0, k, lg) for _, k, lg in trace], supply_digests=supply_digests)["results"]
    plain.close()
    srv = azk.Server(0, max_batch=64, aux_lanes=3)
    print("server timeline mix, digests supplied:", supply_digests, flush=True)
    r = serving.run_server(srv, trace, data, collect=True, supply_digests=supply_digests)
    srv.close()
    assert r["failed"] == 0
    for (kind, lg, info, proof), (k2, lg2, ref) in zip(r["results"], want):
        assert (kind, lg) == (k2, lg2)
        assert proof == ref, (kind, lg)
        assert info["kind"] == {"commit": 0, "sumcheck": 1, "fold": 2, "chunk": 3}[kind]
    spans = sorted((i["start_ns"], i["done_ns"]) for _, _, i, _ in r["results"])
    assert any(b[0] < a[1] for a, b in zip(spans, spans[1:])), "no two jobs overlapped on the device"
    for kind, lg, _, proof in r["results"]:
        if kind == "sumcheck":
            assert azk.sumcheck_verify(proof)
        elif kind == "chunk":
            assert azk.sumcheck_chunk_verify(proof)


@pytest.mark.gpu
def test_server_mix_matches_plain_calls(ctx):
    trace = small_trace()
    data = serving.Inputs(trace, seed=3)
    srv = azk.Server(0, max_batch=64, aux_lanes=2)
    print("server mix of", len(trace), "jobs", flush=True)
    r = serving.run_server(srv, trace, data, collect=True)
    srv.close()
    assert r["failed"] == 0
    want = {(k, lg): alone(ctx, k, lg, data) for _, k, lg in trace}
    for kind, lg, info, proof in r["results"]:
        assert proof == want[(kind, lg)], (kind, lg)
        assert info["kind"] == {"commit": 0, "sumcheck": 1, "fold": 2}[kind]
    # one commit and one sum-check also against the oracle
    c = data.get("commit", 10)
    cw = O.circle_fft(c, 10)
    ref, _ = O.fri_prove(cw.reshape(-1, 4), 10, serving.SECURITY, serving.DIGEST, serving.commit_rounds(10))
    assert want[("commit", 10)] == ref
    assert want[("sumcheck", 12)] == O.sumcheck_multilinear_proof(data.get("sumcheck", 12))
    folded = O.fold_chain(data.get("fold", 12).reshape(-1, 4), 12, serving.fold_challenges())
    assert want[("fold", 12)] == np.asarray(folded, np.uint32).tobytes()
    assert r["mean_commit_batch"] > 1  # jobs submitted together share device programs


@pytest.mark.gpu
def test_server_admission_limits_concurrency(ctx):
    trace = [(0.0, "commit", 14)] * 4 + [(0.0, "sumcheck", 14)] * 2
    data = serving.Inputs(trace, seed=4)
    # room for one 2^14 commit's state (2.9 MiB) and the sum-check lanes, not two commits
    budget = 4 << 20
    srv = azk.Server(0, max_batch=64, aux_lanes=2, hbm_budget=budget)
    print("server tight budget", flush=True)
    r = serving.run_server(srv, trace, data, collect=True)
    srv.close()
    assert r["failed"] == 0
    want_c = alone(ctx, "commit", 14, data)
    want_s = alone(ctx, "sumcheck", 14, data)
    for kind, _, info, proof in r["results"]:
        assert proof == (want_c if kind == "commit" else want_s)
        if kind == "commit":
            assert info["batch_jobs"] == 1


@pytest.mark.gpu
def test_server_rejects_and_recovers(ctx):
    srv = azk.Server(0, hbm_budget=4 << 20)
    big = inputs.circle_fft_coefficients(1, 16)
    t = srv.submit_commit(azk.prover_desc(16, 10, serving.SECURITY, serving.DIGEST), big)
    info = srv.wait(t)
    assert info["status"] == 2  # AZK_ERR_ALLOCATION: larger than the whole budget
    with pytest.raises(azk.AZKError):
        srv.take(t)
    with pytest.raises(azk.AZKError):
        srv.submit_sumcheck(np.zeros(12, np.uint32))  # not a power of two
    with pytest.raises(azk.AZKError):
        srv.submit_commit(azk.prover_desc(10, 11, serving.SECURITY, serving.DIGEST), big)  # rounds > log_n
    small = inputs.circle_fft_coefficients(2, 8)
    t2 = srv.submit_commit(azk.prover_desc(8, 3, serving.SECURITY, serving.DIGEST), small)
    got = srv.take(t2)
    srv.close()
    p = azk.CircleCodewordPCSFRIProverV1(ctx, 8, 3, serving.SECURITY, serving.DIGEST)
    assert got == p.prove_fft_coefficients(small)


@pytest.mark.gpu
def test_server_grinding_job(ctx):
    """A commit with grinding bits through the server equals the plain prover's proof."""
    sec = (2, 9, 1, 10)
    coeffs = inputs.circle_fft_coefficients(5, 11)
    srv = azk.Server(0)
    t = srv.submit_commit(azk.prover_desc(11, 5, sec, serving.DIGEST), coeffs)
    got = srv.take(t)
    srv.close()
    p = azk.CircleCodewordPCSFRIProverV1(ctx, 11, 5, sec, serving.DIGEST)
    assert got == p.prove_fft_coefficients(coeffs)
    assert azk.circle_verify(got, serving.DIGEST)


@pytest.mark.gpu
def test_server_m31_commits_match_lifted_commits(ctx):
    """Commits from M31 words and from their QM31 lift, submitted together (one device
    program may hold both forms), give the plain QM31 prover's proof bytes."""
    trace = [(0.0, "commit", lg) for lg in (10, 12, 11, 13, 9)]
    q = serving.Inputs(trace, seed=8)
    m = serving.Inputs(trace, seed=8, m31_commits=True)
    srv = azk.Server(0, max_batch=64, aux_lanes=1)
    print("server M31 and QM31 commits", flush=True)
    tickets = []
    for _, _, lg in trace:
        d = azk.prover_desc(lg, serving.commit_rounds(lg), serving.SECURITY, serving.DIGEST)
        assert m.get("commit", lg).shape == (1 << lg,)
        tickets.append((lg, srv.submit_commit_m31(d, m.get("commit", lg)), srv.submit_commit(d, q.get("commit", lg))))
    infos, got = [], []
    for lg, t1, t2 in tickets:
        infos += [srv.wait(t1), srv.wait(t2)]
        got.append((lg, srv.take(t1), srv.take(t2)))
    with pytest.raises(azk.AZKError):  # fewer words than the codeword has coefficients
        srv.submit_commit_m31(azk.prover_desc(10, 3, serving.SECURITY, serving.DIGEST), np.zeros(512, np.uint32))
    srv.close()
    assert all(i["status"] == 0 for i in infos)
    assert max(i["batch_jobs"] for i in infos) > 1
    for lg, a, b in got:
        want = alone(ctx, "commit", lg, q)
        assert a == want and b == want, lg


@pytest.mark.gpu
@pytest.mark.parametrize("lanes", [1, 2])
def test_server_sumcheck_shapes_under_tight_budget(ctx, lanes):
    """Sum-check lanes with room for about three plans run a cycle of shapes: each lane keeps
    the plans of the shapes it ran and gives back the least recently used ones when the next
    shape does not fit; every proof equals the plain call's."""
    trace = [(0.0, "sumcheck", lg) for lg in (12, 13, 14, 12, 14, 13, 15, 12, 15, 14)]
    data = serving.Inputs(trace, seed=6)
    srv = azk.Server(0, max_batch=64, aux_lanes=lanes, hbm_budget=600 << 10)
    print("server sum-check shapes, lanes", lanes, flush=True)
    r = serving.run_server(srv, trace, data, collect=True, supply_digests=True)
    srv.close()
    assert r["failed"] == 0
    want = {lg: alone(ctx, "sumcheck", lg, data) for lg in {lg for _, _, lg in trace}}
    for kind, lg, info, proof in r["results"]:
        assert proof == want[lg], lg
