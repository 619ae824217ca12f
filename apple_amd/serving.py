"""Serving harness: a seeded Poisson stream of Circle PCS commits, M31 multilinear
sum-checks and FRI fold chains driven through the C ABI server (azk_server_*), and the
same stream run as plain sequential calls for comparison.

Workload (per job, seeded SplitMix64): commits with log N uniform in [12, 22], 10 FRI
rounds (fewer when log N is smaller), security (log blowup 2, 32 queries, folding 1, no
grinding), public-input digest 0xa0 + i; sum-checks with n = 2^k, k uniform in [16, 22],
canonical M31 words; fold chains of 2^k QM31 evaluations, k uniform in [16, 22], 10 rounds
with fixed challenges.  The kind shares are parameters (default 40% commits, 40%
sum-checks, 20% folds).
"""
from __future__ import annotations

import threading
import time

import numpy as np

from . import azk, inputs

SECURITY = (2, 32, 1, 0)
DIGEST = bytes(0xA0 + i for i in range(32))


def commit_rounds(log_n: int) -> int:
    return min(10, log_n - 1)


FOLD_ROUNDS = 10


def fold_challenges(rounds: int = FOLD_ROUNDS) -> np.ndarray:
    """Canonical QM31 challenges, one per round (fixed; a fold job carries its own)."""
    return inputs.random_m31(0xF01D, 4 * rounds).reshape(rounds, 4)


def make_trace(seed: int, jobs: int, rate: float, commit_logs=(12, 22), sumcheck_logs=(16, 22), fold_logs=(16, 22),
               commit_share: float = 0.4, fold_share: float = 0.2) -> list[tuple[float, str, int]]:
    """(arrival seconds, kind, log size) for `jobs` Poisson arrivals at `rate` jobs/s
    (rate <= 0: every job arrives at time 0)."""
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
        else:
            out.append((t, "sumcheck", int(rng.integers(sumcheck_logs[0], sumcheck_logs[1] + 1))))
    return out


class Inputs:
    """One pinned input per (kind, log size) seen in a trace."""

    def __init__(self, trace, seed: int = 1):
        self.arrays: dict = {}
        self._pins = []
        for _, kind, lg in trace:
            if (kind, lg) in self.arrays:
                continue
            if kind == "commit":
                host = inputs.circle_fft_coefficients(seed + lg, lg)
            elif kind == "fold":
                host = inputs.random_m31(seed + 200 + lg, 4 << lg)
            else:
                host = inputs.random_m31(seed + 100 + lg, 1 << lg)
            pin = azk.PinnedArray(host.shape, np.uint32)
            pin.array[...] = host
            self._pins.append(pin)
            self.arrays[(kind, lg)] = pin.array

    def get(self, kind: str, lg: int) -> np.ndarray:
        return self.arrays[(kind, lg)]


def run_server(server: azk.Server, trace, data: Inputs, collect: bool = False) -> dict:
    """Submits every job at its arrival time (relative to the start) and collects
    the results.  Latency = done - submit on the server's clock."""
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
            tk = server.submit_commit(d, data.get(kind, lg))
        elif kind == "fold":
            tk = server.submit_fold(lg, FOLD_ROUNDS, data.get(kind, lg), fold_challenges())
        else:
            tk = server.submit_sumcheck(data.get(kind, lg))
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


def service_times(ctx: azk.Context, trace, data: Inputs, repeats: int = 1) -> list[float]:
    """Seconds each job takes as a plain call on an idle device (host input, upload
    included, as the server's jobs)."""
    provers, plans, cache = {}, {}, {}
    out = []
    for _, kind, lg in trace:
        key = (kind, lg)
        if key not in cache:
            if kind == "commit":
                p = provers[lg] = azk.CircleCodewordPCSFRIProverV1(ctx, lg, commit_rounds(lg), SECURITY, DIGEST)
                p.prove_fft_coefficients(data.get(kind, lg))  # warm: builds its program
                f = (lambda p=p, a=data.get(kind, lg): p.prove_fft_coefficients(a))
            elif kind == "fold":
                fp = azk.CircleFRIFoldChainPlan(ctx, lg, FOLD_ROUNDS)
                plans[("fold", lg)] = fp
                fp.execute(data.get(kind, lg), fold_challenges())
                f = (lambda fp=fp, a=data.get(kind, lg): fp.execute(a, fold_challenges()))
            else:
                q = plans[lg] = azk.M31MultilinearSumcheckPlan(ctx, lg)
                q.prove(data.get(kind, lg))
                f = (lambda q=q, a=data.get(kind, lg): q.prove(a))
            ts = []
            for _ in range(repeats):
                a = time.perf_counter()
                f()
                ts.append(time.perf_counter() - a)
            cache[key] = float(np.median(ts))
        out.append(cache[key])
    return out


def sequential_latencies(trace, service: list[float]) -> dict:
    """One job at a time in arrival order: start = max(arrival, previous finish)."""
    fin, lat = 0.0, []
    for (at, _, _), s in zip(trace, service):
        start = max(at, fin)
        fin = start + s
        lat.append((fin - at) * 1e3)
    lat = np.array(lat)
    span = fin - trace[0][0]
    return {"jobs_per_s": len(trace) / span, "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)), "elapsed_s": span}
