"""Serving harness: a seeded Poisson stream of Circle PCS commits and M31 multilinear
sum-checks (and, optionally, FRI fold chains and sum-check chunks) driven through the C ABI
server (azk_server_*), and the same stream run as plain calls one job at a time on its own
timeline, for comparison.

Workload (per job, seeded SplitMix64): commits with log N uniform in [12, 22] (log-uniform
sizes), 10 FRI rounds (fewer when log N is smaller), security (log blowup 2, 32 queries,
folding 1, no grinding), public-input digest 0xa0 + i; sum-checks with n = 2^k, k uniform in
[16, 22], canonical M31 words; fold chains of 2^k QM31 evaluations, k uniform in [16, 22],
10 rounds with fixed challenges; chunks of 2^k words, 3 rounds.  Default shares: half
commits, half sum-checks.

A sum-check's initial digest is a serial SHA3 sponge over its 4n bytes on the host.  With
`supply_digests`, every sum-check arrives with that digest already computed (as a caller that
produced the words would hold it), so the run measures the device work and its scheduling
rather than host hashing.
"""
from __future__ import annotations

import hashlib
import struct
import threading
import time

import numpy as np

from . import azk, inputs

SECURITY = (2, 32, 1, 0)
DIGEST = bytes(0xA0 + i for i in range(32))
FOLD_ROUNDS = 10
CHUNK_ROUNDS = 3


def commit_rounds(log_n: int) -> int:
    return min(10, log_n - 1)


def fold_challenges(rounds: int = FOLD_ROUNDS) -> np.ndarray:
    """Canonical QM31 challenges, one per round (fixed; a fold job carries its own)."""
    return inputs.random_m31(0xF01D, 4 * rounds).reshape(rounds, 4)


def words_digest(words: np.ndarray) -> bytes:
    """M31SumcheckEncodingV1.digestWords (SumcheckOracle.swift:1288-1299)."""
    dom = b"AppleZKProver.M31Sumcheck.Words.V1"
    h = hashlib.sha3_256(struct.pack("<I", len(dom)) + dom + struct.pack("<IQ", 1, words.size))
    h.update(np.ascontiguousarray(words, "<u4").tobytes())
    return h.digest()


def make_trace(seed: int, jobs: int, rate: float, commit_logs=(12, 22), sumcheck_logs=(16, 22), fold_logs=(16, 22),
               commit_share: float = 0.5, fold_share: float = 0.0, chunk_share: float = 0.0,
               chunk_logs=(16, 22)) -> list[tuple[float, str, int]]:
    """(arrival seconds, kind, log size) for `jobs` Poisson arrivals at `rate` jobs/s
    (rate <= 0: every job arrives at time 0).  Sum-checks take the share left over."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    for _ in range(jobs):
        if rate > 0:
            t += float(rng.exponential(1.0 / rate))
        u = rng.random()
        if u < commit_share:
            out.append((t, "commit", int(rng.integers(commit_logs[0], commit_logs[1] + 1))))
        elif u < commit_share + fold_share:
            out.append((t, "fold", int(rng.integers(fold_logs[0], fold_logs[1] + 1))))
        elif u < commit_share + fold_share + chunk_share:
            out.append((t, "chunk", int(rng.integers(chunk_logs[0], chunk_logs[1] + 1))))
        else:
            out.append((t, "sumcheck", int(rng.integers(sumcheck_logs[0], sumcheck_logs[1] + 1))))
    return out


class Inputs:
    """One pinned input per (kind, log size) seen in a trace, and the sum-checks' digests.
    With `m31_commits` a commit's input is its 2^k M31 line-basis coefficients (the words
    whose QM31 lift the default form holds: the same proof from a quarter of the bytes)."""

    def __init__(self, trace, seed: int = 1, m31_commits: bool = False):
        self.arrays: dict = {}
        self.digests: dict = {}
        self._pins = []
        for _, kind, lg in trace:
            if (kind, lg) in self.arrays:
                continue
            if kind == "commit":
                host = inputs.circle_fft_coefficients(seed + lg, lg)
                if m31_commits:
                    host = np.ascontiguousarray(host[:, 0])
            elif kind == "fold":
                host = inputs.random_m31(seed + 200 + lg, 4 << lg)
            elif kind == "chunk":
                host = inputs.random_m31(seed + 300 + lg, 1 << lg)
            else:
                host = inputs.random_m31(seed + 100 + lg, 1 << lg)
            pin = azk.PinnedArray(host.shape, np.uint32)
            pin.array[...] = host
            self._pins.append(pin)
            self.arrays[(kind, lg)] = pin.array
            if kind == "sumcheck":
                self.digests[lg] = words_digest(host)

    def get(self, kind: str, lg: int) -> np.ndarray:
        return self.arrays[(kind, lg)]


def run_server(server: azk.Server, trace, data: Inputs, collect: bool = False, supply_digests: bool = False) -> dict:
    """Submits every job at its arrival time (relative to the start) and collects the results.
    Latency = done - submit on the server's clock."""
    descs = {}
    tickets: list[tuple[int, str, int]] = []
    lock = threading.Lock()
    done = threading.Event()
    results: dict = {}

    def collector():
        k = 0
        while True:
            with lock:
                have = len(tickets)
            if k >= have:
                if done.is_set():
                    with lock:
                        if k >= len(tickets):
                            return
                time.sleep(0.0002)
                continue
            tk, kind, lg = tickets[k]
            info = server.wait(tk)
            proof = server.take(tk) if info["status"] == 0 else None
            results[tk] = (kind, lg, info, proof if collect else (len(proof) if proof else 0))
            k += 1

    th = threading.Thread(target=collector)
    th.start()
    t0 = time.perf_counter()
    s0 = azk.now_ns()
    for at, kind, lg in trace:
        wait = at - (time.perf_counter() - t0)
        if wait > 0:
            time.sleep(wait)
        if kind == "commit":
            d = descs.get(lg)
            if d is None:
                d = descs[lg] = azk.prover_desc(lg, commit_rounds(lg), SECURITY, DIGEST)
            a = data.get(kind, lg)
            tk = server.submit_commit_m31(d, a) if a.ndim == 1 else server.submit_commit(d, a)
        elif kind == "fold":
            tk = server.submit_fold(lg, FOLD_ROUNDS, data.get(kind, lg), fold_challenges())
        elif kind == "chunk":
            tk = server.submit_sumcheck_chunk(data.get(kind, lg), CHUNK_ROUNDS)
        else:
            tk = server.submit_sumcheck(data.get(kind, lg), data.digests[lg] if supply_digests else None)
        with lock:
            tickets.append((tk, kind, lg))
    done.set()
    th.join()
    s1 = max(r[2]["done_ns"] for r in results.values())
    lat = np.array([(r[2]["done_ns"] - r[2]["submit_ns"]) * 1e-6 for r in results.values()])
    failed = sum(1 for r in results.values() if r[2]["status"] != 0)
    batches = [r[2]["batch_jobs"] for r in results.values() if r[0] == "commit"]
    return {
        "jobs": len(trace), "failed": failed, "elapsed_s": (s1 - s0) * 1e-9,
        "jobs_per_s": len(trace) / ((s1 - s0) * 1e-9),
        "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
        "mean_commit_batch": float(np.mean(batches)) if batches else 0.0,
        "results": [results[tk] for tk, _, _ in tickets],
    }


class PlainRunner:
    """Plain calls, one job at a time, on one context: the provers and plans of every shape
    are built (and warmed) before a run, as a caller with a steady workload holds them."""

    def __init__(self, ctx: azk.Context, trace, data: Inputs):
        self.ctx, self.data = ctx, data
        self.fns: dict = {}
        self._objs = []
        for _, kind, lg in trace:
            if (kind, lg) not in self.fns:
                self.fns[(kind, lg)] = self._build(kind, lg)

    def _build(self, kind: str, lg: int):
        a = self.data.get(kind, lg)
        if kind == "commit":
            p = azk.CircleCodewordPCSFRIProverV1(self.ctx, lg, commit_rounds(lg), SECURITY, DIGEST)
            self._objs.append(p)
            if a.ndim == 1:  # M31 words
                f = (lambda sd, p=p, a=a: p.prove_fft_coefficients_m31(a))
            else:
                f = (lambda sd, p=p, a=a: p.prove_fft_coefficients(a))
        elif kind == "fold":
            fp = azk.CircleFRIFoldChainPlan(self.ctx, lg, FOLD_ROUNDS)
            self._objs.append(fp)
            f = (lambda sd, fp=fp, a=a: fp.execute(a, fold_challenges()).tobytes())
        elif kind == "chunk":
            cp = azk.MetalSumcheckChunkPlan(self.ctx, 1 << lg, CHUNK_ROUNDS)
            self._objs.append(cp)
            f = (lambda sd, cp=cp, a=a: cp.prove(a))
        else:
            q = azk.M31MultilinearSumcheckPlan(self.ctx, lg)
            self._objs.append(q)
            d = self.data.digests[lg]
            f = (lambda sd, q=q, a=a, d=d: q.prove(a, initial_digest=d if sd else None))
        f(False)  # warm: builds programs and plans
        return f

    def run(self, trace, supply_digests: bool = False) -> dict:
        """Each job starts at max(arrival, previous finish) on a real clock; latency = finish -
        arrival.  Returns the same fields as run_server, results included."""
        t0 = time.perf_counter()
        lat, results = [], []
        for at, kind, lg in trace:
            now = time.perf_counter() - t0
            if at > now:
                time.sleep(at - now)
            out = self.fns[(kind, lg)](supply_digests)
            fin = time.perf_counter() - t0
            lat.append((fin - at) * 1e3)
            results.append((kind, lg, out))
        span = time.perf_counter() - t0
        lat = np.array(lat)
        return {"jobs": len(trace), "elapsed_s": span, "jobs_per_s": len(trace) / span,
                "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                "results": results}

    def close(self) -> None:
        for o in self._objs:
            o.close()
        self._objs = []
