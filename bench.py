"""Benchmark: resident Circle-PCS commit superstep on one MI355X.

Workload (one step = one commit): P(x), Q(x) of 2^19 M31 coefficients each (SplitMix64,
seeded), laid out as interleaved line-basis coefficients (the reference lifts them to QM31
(a, 0, 0, 0)), Circle FFT to a 2^21 codeword on the canonical bit-reversed domain, SHA3-256
Merkle commitment of the 16-byte QM31 leaves, 10 Circle FRI fold rounds with on-device
Fiat-Shamir challenges, 32 query openings; only the proof bytes (roots, final layer,
openings) are read back.  The timed entry takes the M31 words (its FFT runs on M31 words);
the QM31-input entry on the lifted array gives the same proof and is timed beside it.
Inputs are resident in HBM before the timed region.  Every proof of the timed region is
byte-identical to the first, which is checked against the oracle (--check).

Prints one JSON line (rank 0, N = 1).  A second, batched form (64 independent commits of
log sizes 14..21 in one captured graph) is reported beside the plain form.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from apple_amd import azk, inputs  # noqa: E402

METRIC = "Merkle+FRI commits/sec and GB/s (bit-exact) batched vs plain; p99 latency"
LOG_N, ROUNDS, SECURITY = 21, 10, (2, 32, 1, 0)  # log blowup 2, 32 queries, folding 1, no grinding
DIGEST = bytes(0xA0 + i for i in range(32))
HBM_PEAK_GBS = 8000.0
# VALU roof of the hashing kernels, in Keccak-f[1600] permutations per second.  gfx950 issues a
# wave64 v_xor_b32 / v_bitop3_b32 in 2 SIMD cycles (a v_bitop3_b32 whose three VGPR sources
# share one parity in 4) and v_alignbit_b32 in 4 (measured live by azk_calibrate_valu).  The
# one-lane permutation kf_v runs its middle 22 rounds on hand-placed registers (keccak_round_asm.hpp):
# 120 v_bitop3_b32 + 2 v_xor_b32 + 58 v_alignbit_b32 per round, no same-parity three-source op;
# its first and last rounds are compiled (70 v_bitop3_b32 + 62 v_xor_b32 + 58 v_alignbit_b32).
# 22 x (122 x 2 + 58 x 4) + 2 x (132 x 2 + 58 x 4) = 11,464 issue cycles per wave of 64
# permutations.  At the 2.4 GHz peak clock on 256 CUs x 4 SIMDs:
SIMDS, PEAK_CLOCK_HZ = 256 * 4, 2.4e9
KECCAK_WAVE_CYCLES = 22 * (2 * 122 + 4 * 58) + 2 * (2 * 132 + 4 * 58)
KECCAK_PEAK_GPERMS = SIMDS * PEAK_CLOCK_HZ * 64 / KECCAK_WAVE_CYCLES / 1e9   # 13.7 G permutations/s
CALIBRATION: dict = {}


def keccak_roofline(perms: float, seconds: float, **extra) -> dict:
    """Permutations per second against the issue roof, and against the rate the same
    permutation reaches on register-resident states on this chip (no memory traffic)."""
    achieved = perms / seconds / 1e9
    r = {"bound": "valu", "achieved": achieved, "peak": KECCAK_PEAK_GPERMS, "unit": "Gperm/s",
         "frac": achieved / KECCAK_PEAK_GPERMS,
         "peak_model": f"{KECCAK_WAVE_CYCLES} issue cycles per wave64 Keccak-f at 2.4 GHz on 1024 SIMDs"}
    if CALIBRATION:
        ceil = CALIBRATION["keccak_perms_per_s"] / 1e9
        r["measured_ceiling"] = ceil
        r["frac_of_measured_ceiling"] = achieved / ceil
    r.update(extra)
    return r


def commit_bytes(log_n: int, rounds: int) -> int:
    """Algorithmic bytes of one commit: coefficients in + every committed layer's leaves."""
    n = 1 << log_n
    return 16 * n + sum(16 * (n >> r) for r in range(rounds))


def commit_perms(log_n: int, rounds: int) -> int:
    """Keccak-f permutations of one commit: per committed layer N_r leaves + N_r - 1 parents."""
    n = 1 << log_n
    return sum(2 * (n >> r) - 1 for r in range(rounds))


def pmc_traffic(kernel: str):
    """HBM bytes per launch of `kernel` from the committed FETCH_SIZE/WRITE_SIZE passes
    (profiles/pmc_bytes_per_dispatch_2^21.txt, counters scaled by a known-size copy)."""
    path = os.path.join(ROOT, "profiles", "pmc_bytes_per_dispatch_2^21.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split()
        if len(parts) == 3 and parts[0] == kernel:
            return int(parts[1]) + int(parts[2])
    return None


def pct(v, q):
    return float(np.percentile(np.asarray(v, dtype=np.float64), q))


def cpu_baseline(coeffs, sample_commits: int) -> dict:
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O  # the oracle is the CPU baseline (single-threaded C restatement)

    t0 = time.perf_counter()
    for _ in range(sample_commits):
        cw = O.circle_fft(coeffs, LOG_N)
        O.fri_prove(cw.reshape(-1, 4), LOG_N, SECURITY, DIGEST, ROUNDS)
    dt = time.perf_counter() - t0
    return {"value": sample_commits / dt, "unit": "commits/s", "cores": 1, "kind": "port",
            "sample": f"{sample_commits} full 2^{LOG_N} commits (FFT + prove) with oracle/liboracle.so, "
                      f"{dt:.1f} s on one host core"}


def host_threads() -> dict:
    """Host cores: what the OS reports, what this process may use, and the threads the CPU
    legs run (the GPU box gives one GPU's jobs a share of 16 cores)."""
    aff = len(os.sched_getaffinity(0))
    return {"nproc": os.cpu_count(), "affinity": aff, "threads": min(16, aff)}


def cpu_baselines_by_config(seed: int = 4242) -> dict:
    """The oracle (a single-threaded C restatement) on bounded samples of the other configs, on
    one core and on every core of the share at once.  Reported baselines only."""
    from concurrent.futures import ThreadPoolExecutor
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O  # the oracle is the CPU baseline
    from apple_amd import serving
    th = host_threads()
    n_th = th["threads"]
    out = {"host": th}

    # config 2: Keccak-f[1600] on 25 x u64 states, SHA3-256 of 136-byte messages (2 perms each)
    count = 1 << 17
    rng = np.random.default_rng(2)
    states = rng.integers(0, 256, size=200 * count, dtype=np.uint8)
    msgs = rng.integers(0, 256, size=136 * count, dtype=np.uint8)

    def perm(_):
        O.keccakf_batch(states, count=count)

    def sha(_):
        O.hash_oneblock_batch(msgs, 136, 136, count, 0x06)

    res = {}
    for name, fn, unit_per_call in (("keccakf_perms_per_s", perm, count), ("sha3_256_messages_per_s", sha, count)):
        t0 = time.perf_counter()
        fn(0)
        one = unit_per_call / (time.perf_counter() - t0)
        t0 = time.perf_counter()
        with ThreadPoolExecutor(n_th) as ex:
            list(ex.map(fn, range(n_th)))
        res[name] = {"one_core": one, "all_cores": n_th * unit_per_call / (time.perf_counter() - t0)}
    out["keccak_batch"] = {"sample": f"2^17 states / 2^17 x 136-byte messages per call, one call per thread", **res}

    # config 4: the batched shape (64 commits, log N 14..21, seed 4242): one job per task on every core
    logs = [int(x) for x in np.random.default_rng(seed).integers(14, 22, size=64)]

    def commit(j):
        ln = logs[j]
        c = inputs.circle_fft_coefficients(seed=seed + j, log_n=ln)
        O.fri_prove(O.circle_fft(c, ln).reshape(-1, 4), ln, SECURITY, DIGEST, ROUNDS)

    t0 = time.perf_counter()
    with ThreadPoolExecutor(n_th) as ex:
        list(ex.map(commit, range(len(logs))))
    dt = time.perf_counter() - t0
    out["batched"] = {"commits_per_s_all_cores": len(logs) / dt, "seconds": dt,
                      "sample": "the whole 64-job batch, one commit per task"}

    # config 5: the serving mix run one job at a time on one core (the first 12 jobs of the trace)
    trace = serving.make_trace(77, 12, rate=0)
    data = {}
    t0 = time.perf_counter()
    for _, kind, lg in trace:
        if kind == "commit":
            c = data.setdefault((kind, lg), inputs.circle_fft_coefficients(seed + lg, lg))
            O.fri_prove(O.circle_fft(c, lg).reshape(-1, 4), lg, serving.SECURITY, serving.DIGEST,
                        serving.commit_rounds(lg))
        else:
            w = data.setdefault((kind, lg), inputs.random_m31(seed + 100 + lg, 1 << lg))
            O.sumcheck_multilinear_proof(w)
    dt = time.perf_counter() - t0
    out["serving"] = {"jobs_per_s_one_core": len(trace) / dt, "seconds": dt,
                      "sample": "the first 12 jobs of the seeded mix (half commits, half sum-checks), sequential"}
    return out


def cpu_all_cores(coeffs, threads: int | None = None) -> dict:
    """The same oracle commit on every host core at once: one independent 2^21 commit per
    thread (the oracle's C calls release the interpreter lock), a reported baseline only."""
    from concurrent.futures import ThreadPoolExecutor
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O  # the oracle is the CPU baseline

    n = threads or host_threads()["threads"]

    def one(_):
        cw = O.circle_fft(coeffs, LOG_N)
        O.fri_prove(cw.reshape(-1, 4), LOG_N, SECURITY, DIGEST, ROUNDS)

    t0 = time.perf_counter()
    with ThreadPoolExecutor(n) as ex:
        list(ex.map(one, range(n)))
    dt = time.perf_counter() - t0
    return {"value": n / dt, "unit": "commits/s", "cores": n, "kind": "port", **host_threads(),
            "sample": f"{n} full 2^{LOG_N} commits, one per thread, oracle/liboracle.so, {dt:.1f} s"}


def batched_form(ctx, seed: int = 4242, jobs: int = 64, reps: int = 3) -> dict:
    """64 independent commits, log sizes uniform in 14..21 (seeded), one batch program vs
    the same jobs proved one after another; every batched proof equals its plain proof.  Both
    sides take each job's M31 coefficients (the QM31-input batch is timed beside)."""
    rng = np.random.default_rng(seed)
    logs = [int(x) for x in rng.integers(14, 22, size=jobs)]
    provers, bufs, bufs_qm = [], [], []
    for j, ln in enumerate(logs):
        pr = azk.CircleCodewordPCSFRIProverV1(ctx, ln, ROUNDS, SECURITY, DIGEST)
        provers.append(pr)
        c = inputs.circle_fft_coefficients(seed=seed + j, log_n=ln)
        bufs.append(ctx.upload(np.ascontiguousarray(c[:, 0])))
        bufs_qm.append(ctx.upload(c))
    plain = [pr.prove_fft_coefficients_m31_resident(b) for pr, b in zip(provers, bufs)]  # warm + reference
    t_plain = []
    for _ in range(reps):
        t0 = time.perf_counter()
        for pr, b in zip(provers, bufs):
            pr.prove_fft_coefficients_m31_resident(b)
        t_plain.append(time.perf_counter() - t0)
    plan = azk.CircleBatchPlan(provers, bufs, m31=True)
    if plan.prove() != plain:
        raise SystemExit("FATAL: batched proofs differ from plain single-job proofs")
    t_batch = []
    for _ in range(reps):
        t0 = time.perf_counter()
        plan.prove()
        t_batch.append(time.perf_counter() - t0)
    plan.close()
    plan_qm = azk.CircleBatchPlan(provers, bufs_qm)
    if plan_qm.prove() != plain:
        raise SystemExit("FATAL: QM31-input batched proofs differ from the M31-input proofs")
    t_qm = []
    for _ in range(reps):
        t0 = time.perf_counter()
        plan_qm.prove()
        t_qm.append(time.perf_counter() - t0)
    plan_qm.close()
    for pr in provers:
        pr.close()
    plain_rate = jobs / float(np.median(t_plain))
    batch_rate = jobs / float(np.median(t_batch))
    nbytes = sum(commit_bytes(ln, ROUNDS) for ln in logs)
    return {"workload": "64 circle_pcs_commit jobs, log_n ~ U{14..21}", "jobs": jobs,
            "entry": "azk_circle_batch_create_m31 (plain: azk_circle_prove_fft_coefficients_m31_resident)",
            "plain_commits_per_s": plain_rate, "batched_commits_per_s": batch_rate,
            "speedup": batch_rate / plain_rate, "batched_gbps": nbytes / float(np.median(t_batch)) / 1e9,
            "batch_ms": 1e3 * float(np.median(t_batch)), "bit_exact_vs_plain": True,
            "qm31_input_batched_commits_per_s": jobs / float(np.median(t_qm))}


def serving_form(ctx, seed: int = 77, jobs: int = 200, sat_jobs: int = 300) -> dict:
    """Poisson stream of Circle PCS commits (log N uniform in 12..22) and M31 multilinear
    sum-checks (2^16..2^22 words), half each, through the C ABI server at 0.8 x its saturated
    throughput, against the same stream run as plain calls one job at a time on a real
    timeline (same inputs, same host arrays, provers and plans warm).  Every served result is
    compared byte for byte with the plain call's.  Two variants: sum-check digests hashed on
    the host inside each job, and digests supplied by the caller (device work only)."""
    from apple_amd import serving
    shapes = [(0.0, "commit", lg) for lg in range(12, 23)] + [(0.0, "sumcheck", lg) for lg in range(16, 23)]
    data = serving.Inputs(shapes, seed=seed, m31_commits=True)
    plain = serving.PlainRunner(ctx, shapes, data)
    # 12 sum-check lanes: each hashes its job's words on one host core (the box gives 16)
    srv = azk.Server(0, max_batch=64, aux_lanes=12)
    serving.run_server(srv, shapes + shapes, data)  # warm: provers and plans of every shape
    # steady state: the server holds the provers and plans a saturated batch needs (several of
    # one shape in one device program), as a serving process does after its first batches
    serving.run_server(srv, serving.make_trace(seed, sat_jobs, rate=0), data)
    out = {"workload": f"poisson mix: 50% Circle PCS commits log N U{{12..22}} from M31 coefficients (10 rounds, 32 "
                       f"queries), 50% multilinear "
                       f"sum-checks 2^U{{16..22}} words; seed {seed}; rate 0.8 x saturated",
           "jobs": jobs}
    for variant, supply in (("host_digests", False), ("supplied_digests", True)):
        sat_trace = serving.make_trace(seed, sat_jobs, rate=0)
        sat = serving.run_server(srv, sat_trace, data, supply_digests=supply)
        seq_sat = plain.run(sat_trace, supply_digests=supply)
        rate = 0.8 * sat["jobs_per_s"]
        trace = serving.make_trace(seed + 1, jobs, rate=rate)
        run = serving.run_server(srv, trace, data, collect=True, supply_digests=supply)
        seq = plain.run(trace, supply_digests=supply)
        same = run["failed"] == 0 and all(info["status"] == 0 and proof == ref for (_, _, info, proof), (_, _, ref)
                                          in zip(run["results"], seq["results"]))
        out[variant] = {
            "rate_jobs_per_s": rate, "saturated_jobs_per_s": sat["jobs_per_s"],
            "jobs_per_s": run["jobs_per_s"], "p50_ms": run["p50_ms"], "p99_ms": run["p99_ms"],
            "mean_commit_batch": run["mean_commit_batch"], "failed": run["failed"],
            "results_match_plain_calls": bool(same),
            "plain_sequential": {"saturated_jobs_per_s": seq_sat["jobs_per_s"], "jobs_per_s": seq["jobs_per_s"],
                                 "p50_ms": seq["p50_ms"], "p99_ms": seq["p99_ms"]},
            "speedup_saturated": sat["jobs_per_s"] / seq_sat["jobs_per_s"],
        }
    srv.close()
    plain.close()
    return out


# VALU instructions per QM31 butterfly in this build's tile loops (gfx950 assembly): stages run in
# radix-4 pairs (205 instructions per group of four butterflies); an odd stage count per pass leaves
# one radix-2 stage (58 per butterfly in the column passes, 53 in the final contiguous pass).
FFT_PAIR_BUTTERFLY_OPS = 205 / 4
FFT_SINGLE_BUTTERFLY_OPS = (58, 53)


def fft_stage_split(log_n: int, log_t_final: int = 10, log_t_pass: int = 13) -> list:
    """Stage counts of the passes the codeword plan runs: passes of at most log_t_pass - 3
    stages, balanced, then the final pass of min(log_t_final, log_n) stages (2^10-element
    final tiles when the codeword is not hashed)."""
    lt = min(log_t_final, log_n)
    s, kmax = log_n - lt, log_t_pass - 3
    p = (s + kmax - 1) // kmax if s else 0
    return [s // p + (1 if i < s % p else 0) for i in range(p)] + [lt]


def fft_valu_ops(log_n: int) -> float:
    ks = fft_stage_split(log_n)
    per_element_half = sum((k // 2) * 2 * FFT_PAIR_BUTTERFLY_OPS + (k % 2) * FFT_SINGLE_BUTTERFLY_OPS[i == len(ks) - 1]
                           for i, k in enumerate(ks))
    return (1 << (log_n - 1)) * per_element_half


def standalone_form(ctx, log_n: int = 26, reps: int = 5) -> dict:
    """The codeword and fold-chain plans on their own at a size far beyond the 256 MB
    MALL: Circle FFT of 2^26 QM31 coefficients (1 GiB) and one fold 2^26 -> 2^25.
    Inputs are the 2^22-coefficient seeded block repeated (the kernels' work does not
    depend on the values); parity of both plans is covered by the tests at <= 2^20."""
    n = 1 << log_n
    block = inputs.circle_fft_coefficients(seed=3, log_n=22)
    cin, cout = ctx.buffer(16 * n), ctx.buffer(16 * n)
    for k in range(n >> 22):
        cin.upload(block, offset=k * block.nbytes)
    fft = azk.CircleCodewordPlan(ctx, log_n)
    fft.execute_resident(cin, cout)
    t_fft = float(np.median([fft.execute_resident(cin, cout)["gpu_ns"] for _ in range(reps)])) * 1e-9
    fold = azk.CircleFRIFoldChainPlan(ctx, log_n, 1)
    fout = ctx.buffer(16 * n // 2)
    ch = np.array([[123456789, 987654321, 55555, 7]], np.uint32)
    fold.execute_resident(cout, ch, fout)
    t_fold = float(np.median([fold.execute_resident(cout, ch, fout)["gpu_ns"] for _ in range(reps)])) * 1e-9
    fold_bytes = 16 * n + 4 * (n // 2) + 16 * (n // 2)  # pairs in, M31 twiddles in, outputs
    fft_ops = fft_valu_ops(log_n)
    for b in (cin, cout, fout):
        b.free()
    # the codeword of M31 coefficients (N words in, N words out)
    block = np.ascontiguousarray(inputs.circle_fft_coefficients(seed=3, log_n=22)[:, 0])
    min_, mout = ctx.buffer(4 * n), ctx.buffer(4 * n)
    for k in range(n >> 22):
        min_.upload(block, offset=k * block.nbytes)
    fft.execute_m31_resident(min_, mout)
    t_m31 = float(np.median([fft.execute_m31_resident(min_, mout)["gpu_ns"] for _ in range(reps)])) * 1e-9
    for b in (min_, mout):
        b.free()
    return {
        "fold": {"shape": f"2^{log_n} -> 2^{log_n - 1} QM31", "ms": t_fold * 1e3,
                 "roofline": {"bound": "hbm", "achieved": fold_bytes / t_fold / 1e9, "peak": HBM_PEAK_GBS,
                              "unit": "GB/s", "frac": fold_bytes / t_fold / 1e9 / HBM_PEAK_GBS,
                              "algorithmic_bytes": fold_bytes}},
        "fft": {"shape": f"2^{log_n} QM31 codeword", "ms": t_fft * 1e3,
                # algorithmic bytes: the codeword read once and written once (32 B per element)
                "roofline": {"bound": "hbm", "achieved": 32 * n / t_fft / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                             "frac": 32 * n / t_fft / 1e9 / HBM_PEAK_GBS, "algorithmic_bytes": 32 * n,
                             "stages_per_pass": fft_stage_split(log_n),
                             # VALU issue at 4 SIMD cycles per wave64 instruction (an upper bound: some
                             # of the butterfly's adds and moves issue in 2)
                             "valu_instructions_per_butterfly": fft_ops / (log_n * (n // 2)),
                             "valu_issue_frac_4_cycle_bound": fft_ops / 64 * 4 / (SIMDS * PEAK_CLOCK_HZ * t_fft)}},
        "fft_m31": {"shape": f"2^{log_n} M31 codeword (M31 coefficients)", "ms": t_m31 * 1e3,
                    "roofline": {"bound": "hbm", "achieved": 8 * n / t_m31 / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                 "frac": 8 * n / t_m31 / 1e9 / HBM_PEAK_GBS, "algorithmic_bytes": 8 * n}},
    }


def witness_basis_form(ctx, log_n: int = 21, reps: int = 5, small_log_n: int = 13) -> dict:
    """Monomial witness columns -> Circle-FFT basis (azk_witness_basis_*), resident: two full
    columns of 2^(log_n - 1) QM31 monomial coefficients -> 2^log_n interleaved basis
    coefficients.  The reference's route is O(n^2) on the CPU (the resident plan's dense matrix
    stops at 4096 per column); at 2^small_log_n the oracle's C restatement of that route is
    timed beside the device for scale.  Parity is covered by the tests (oracle to 2^15,
    FFT-of-basis vs Horner evaluation at 2^17 .. 2^23)."""
    half = 1 << (log_n - 1)
    rng = np.random.default_rng(5)
    x = ctx.upload(rng.integers(0, 2**31 - 1, size=(half, 4), dtype=np.uint32))
    y = ctx.upload(rng.integers(0, 2**31 - 1, size=(half, 4), dtype=np.uint32))
    out = ctx.buffer(32 * half)
    plan = azk.CircleWitnessToFFTBasisPlanV1(ctx, log_n)
    plan.execute_resident(x, half, y, half, out)
    t = float(np.median([plan.execute_resident(x, half, y, half, out)["gpu_ns"] for _ in range(reps)])) * 1e-9
    small = 1 << (small_log_n - 1)
    xs = rng.integers(0, 2**31 - 1, size=(small, 4), dtype=np.uint32)
    ys = rng.integers(0, 2**31 - 1, size=(small, 4), dtype=np.uint32)
    sp = azk.CircleWitnessToFFTBasisPlanV1(ctx, small_log_n)
    sp.execute(xs, ys)
    ts = []
    for _ in range(reps):
        sp.execute(xs, ys)
        ts.append(sp.last_stats["gpu_ns"])  # host arrays: upload + conversion + download
    t_small = float(np.median(ts)) * 1e-9
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O  # the CPU baseline for this row (C restatement of the reference's O(n^2) route)
    c0 = time.perf_counter()
    O.basis_from_monomial(xs, ys, small_log_n)
    t_cpu = time.perf_counter() - c0
    for b in (x, y, out):
        b.free()
    return {"shape": f"2 x 2^{log_n - 1} monomial QM31 -> 2^{log_n} Circle-FFT basis", "ms": t * 1e3,
            "coefficients_per_s": 2 * half / t,
            "small": {"shape": f"2 x 2^{small_log_n - 1} -> 2^{small_log_n}", "gpu_ms": t_small * 1e3,
                      "cpu_oracle_ms": t_cpu * 1e3, "cpu_cores": 1}}


def monomial_input_form(ctx, prover, steps: int = 20) -> dict:
    """The commit workload with P and Q given as monomial coefficients (2^19 M31 each, lifted to
    QM31): the device witness -> Circle-FFT-basis conversion, then the same commit, both on
    resident buffers (wall time per commit).  The reference's monomial route is O(n^2) on the
    CPU, so its own benchmark times the basis-coefficient input."""
    h = 1 << (LOG_N - 2)
    pq = inputs.random_m31(1, 2 * h)
    x = np.zeros((h, 4), np.uint32)
    y = np.zeros((h, 4), np.uint32)
    x[:, 0], y[:, 0] = pq[:h], pq[h:]
    bx, by = ctx.upload(x), ctx.upload(y)
    basis = ctx.buffer(16 << LOG_N)
    plan = azk.CircleWitnessToFFTBasisPlanV1(ctx, LOG_N)
    conv = []

    def step():
        conv.append(plan.execute_resident(bx, h, by, h, basis)["gpu_ns"])
        return prover.prove_fft_coefficients_resident(basis)

    first = step()
    t0 = time.perf_counter()
    for _ in range(steps):
        last = step()
    dt = (time.perf_counter() - t0) / steps
    for b in (bx, by, basis):
        b.free()
    return {"shape": f"P, Q of 2^{LOG_N - 2} monomial M31 coefficients -> 2^{LOG_N} commit", "steps": steps,
            "ms_per_commit": dt * 1e3, "commits_per_s": 1.0 / dt,
            "conversion_ms": float(np.median(conv[1:])) * 1e-6, "deterministic": last == first}


def keccak_form(ctx, log_count: int = 22, reps: int = 10, warm: int = 30) -> dict:
    """Keccak-f[1600] on 2^22 independent states (200 B each) and SHA3-256 / Keccak-256 of 2^22
    one-block messages of 136 bytes (padding makes two permutations per message at the full
    rate block), resident, uniform random bytes from a seed.  SHA3 digests are spot-checked
    against hashlib."""
    count = 1 << log_count
    rng = np.random.default_rng(2)
    states = ctx.upload(rng.integers(0, 256, size=200 * count, dtype=np.uint8))
    sout = ctx.buffer(200 * count)
    kp = azk.KeccakF1600PermutationBatchPlan(ctx, count)
    for _ in range(warm):  # the clock ramps over the first calls after the uploads
        kp.permute_resident(states, sout)
    t_perm = float(np.median([kp.permute_resident(states, sout)["gpu_ns"] for _ in range(reps)])) * 1e-9
    states.free()
    sout.free()
    msgs_host = rng.integers(0, 256, size=136 * count, dtype=np.uint8)
    msgs = ctx.upload(msgs_host)
    dout = ctx.buffer(32 * count)
    res = {}
    for name, cls in (("sha3_256", azk.SHA3FixedOneBlockHashPlan), ("keccak_256", azk.Keccak256FixedOneBlockHashPlan)):
        hp = cls(ctx, count, 136, 136, 32)
        for _ in range(warm):
            hp.hash_resident(msgs, dout)
        t = float(np.median([hp.hash_resident(msgs, dout)["gpu_ns"] for _ in range(reps)])) * 1e-9
        # a 136-byte message is a full rate block plus a padding block: two permutations
        res[name] = {"ms": t * 1e3, "messages_per_s": count / t, "gbps_in": 136 * count / t / 1e9,
                     "roofline": keccak_roofline(2 * count, t, permutations_per_message=2)}
        # every digest of the timed launch: against the product's host gate of the same plan
        # (hash_execute_verified recomputes all 2^22 on the host), then 1024 against hashlib
        got = dout.download().reshape(count, 32)
        gate = hp.hash(msgs_host, verified=True).reshape(count, 32)
        res[name]["all_digests_vs_host_gate"] = bool(np.array_equal(got, gate))
        if name == "sha3_256":
            res[name]["spot_check_1024_vs_hashlib"] = all(
                hashlib.sha3_256(msgs_host[136 * i:136 * (i + 1)].tobytes()).digest() == got[i].tobytes()
                for i in range(1024))
    msgs.free()
    dout.free()
    return {"states": count, "permute_ms": t_perm * 1e3, "perms_per_s": count / t_perm,
            "roofline": keccak_roofline(count, t_perm), **res}


def merkle_plain_form(ctx, log_count: int = 16, leaf_bytes: int = 64, openings: int = 16, reps: int = 20,
                      check: bool = True) -> dict:
    """The plain SHA3-256 Merkle commitment of 2^16 leaves of 64 bytes (seeded uniform bytes):
    resident commit (root) and 16 opening paths gathered from the resident tree (one call each, and
    all 16 in one azk_merkle_open_batch call), against the oracle on one host core; root and
    openings bit-exact."""
    count = 1 << log_count
    raw = inputs.random_bytes(16, count * leaf_bytes)
    plan = azk.SHA3RawLeavesMerkleCommitPlan(ctx, count, leaf_bytes, leaf_bytes)
    buf = ctx.upload(raw)
    for _ in range(reps):
        root = plan.commit_resident(buf)
    t_gpu, t_wall = [], []
    for _ in range(reps):
        plan.commit_resident(buf)
        t_gpu.append(plan.last_stats["gpu_ns"] * 1e-9)
        t_wall.append(plan.last_stats["wall_ns"] * 1e-9)
    idx = [int(x) for x in inputs.splitmix64(count, openings) % np.uint64(count)]
    t0 = time.perf_counter()
    paths = [plan.open(i) for i in idx]
    t_open = time.perf_counter() - t0
    plan.open_batch(idx)
    t_batch = []
    for _ in range(reps):
        t0 = time.perf_counter()
        batch, broot = plan.open_batch(idx)
        t_batch.append(time.perf_counter() - t0)
    if batch != [p for p, _ in paths] or broot != root:
        raise SystemExit("FATAL: batched Merkle openings differ from the one-at-a-time form")
    t_batch = float(np.median(t_batch))
    res = {"workload": f"sha3_merkle_2^{log_count}x{leaf_bytes}B", "leaves": count, "openings": openings,
           "commit_gpu_ms": 1e3 * float(np.median(t_gpu)), "commit_wall_ms": 1e3 * float(np.median(t_wall)),
           "commits_per_s": 1.0 / float(np.median(t_wall)), "openings_ms": 1e3 * t_open,
           "openings_batch_ms": 1e3 * t_batch,
           "commit_and_openings_per_s": 1.0 / (float(np.median(t_wall)) + t_batch)}
    perms = 2 * count - 1  # one permutation per 64-byte leaf and per parent
    res["roofline"] = keccak_roofline(perms, float(np.median(t_gpu)), permutations=perms)
    if check:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as O  # checker and CPU baseline only
        t0 = time.perf_counter()
        want = O.merkle_root(raw, count, leaf_bytes, leaf_bytes)
        res["cpu_oracle_ms"] = 1e3 * (time.perf_counter() - t0)
        ok = want == root and all(O.merkle_opening(raw, count, leaf_bytes, leaf_bytes, i) == pth
                                  for i, pth in zip(idx, paths))
        if not ok:
            raise SystemExit("FATAL: Merkle root or openings differ from the oracle")
        res["bit_exact_vs_oracle"] = True
    buf.free()
    plan.close()
    return res


def calibrate_copy(ctx, log_count: int = 23) -> None:
    """A plain copy of known size for scaling the PMC byte counters: k_copy moves
    2^23 x 32 bytes (256 MiB) in and out (the prehashed commit's first kernel)."""
    count = 1 << log_count
    hashes = ctx.buffer(32 * count)
    plan = azk.SHA3PrehashedLeavesMerkleCommitPlan(ctx, count)
    plan.commit_resident(hashes)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--check", type=int, default=1, help="compare the first proof with the oracle")
    ap.add_argument("--cpu-sample", type=int, default=2, help="commits timed on the CPU oracle (0 = skip)")
    ap.add_argument("--batched", type=int, default=1)
    ap.add_argument("--serving", type=int, default=1)
    ap.add_argument("--serving-only", type=int, default=0, help="only the serving form")
    ap.add_argument("--standalone", type=int, default=1, help="FFT / fold plans at 2^26")
    ap.add_argument("--standalone-only", type=int, default=0,
                    help="profiling aid: only the 2^26 FFT / fold plans (one call each) and the calibration copy")
    ap.add_argument("--calibrate", type=int, default=0, help="run the known-size copy kernel (PMC scaling)")
    ap.add_argument("--monomial", type=int, default=1, help="the commit from monomial coefficients")
    ap.add_argument("--keccak", type=int, default=1, help="Keccak-f / one-block hash batches of 2^22")
    ap.add_argument("--keccak-only", type=int, default=0, help="profiling aid: only the Keccak batches")
    ap.add_argument("--basis-only", type=int, default=0,
                    help="profiling aid: only the witness -> Circle-FFT-basis conversion at 2^21")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (jobs shard whole per GPU; "
              "see the sharding design). Run with --gpus 1.", file=sys.stderr)
        return 2

    ctx = azk.Context(0)
    caps = ctx.capabilities()
    CALIBRATION.update(azk.calibrate_valu(ctx))
    if args.keccak_only:
        print(json.dumps({"keccak_batch": keccak_form(ctx), "valu_calibration": CALIBRATION}), flush=True)
        return 0
    if args.serving_only:
        print(json.dumps({"serving": serving_form(ctx)}), flush=True)
        return 0
    if args.basis_only:
        print(json.dumps({"witness_basis": witness_basis_form(ctx)}), flush=True)
        return 0
    if args.standalone_only:
        calibrate_copy(ctx)
        print(json.dumps({"standalone": standalone_form(ctx, reps=1)}), flush=True)
        return 0
    if args.calibrate:
        calibrate_copy(ctx)
    coeffs = inputs.circle_fft_coefficients(seed=1, log_n=LOG_N)
    # the configuration's coefficients are M31 (P, Q lifted to (a, 0, 0, 0)): the headline
    # proves from the M31 words (the FFT on M31 words); the QM31-input entry is timed beside it
    m31 = np.ascontiguousarray(coeffs[:, 0])
    buf = ctx.upload(m31)
    buf_qm = ctx.upload(coeffs)
    prover = azk.CircleCodewordPCSFRIProverV1(ctx, LOG_N, ROUNDS, SECURITY, DIGEST)

    first = prover.prove_fft_coefficients_m31_resident(buf)
    if prover.prove_fft_coefficients_resident(buf_qm) != first:
        print("FATAL: M31-input and QM31-input proofs differ", file=sys.stderr)
        return 3
    checked = False
    if args.check:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as O  # checker only
        want, _ = O.fri_prove(O.circle_fft(coeffs, LOG_N).reshape(-1, 4), LOG_N, SECURITY, DIGEST, ROUNDS)
        if first != want:
            print("FATAL: GPU proof differs from the oracle", file=sys.stderr)
            return 3
        checked = True
    # the QM31-input entry on the lifted array (16 bytes per coefficient, QM31 FFT)
    for _ in range(max(args.warmup, 2)):
        prover.prove_fft_coefficients_resident(buf_qm)
    lat_qm = []
    for _ in range(min(args.steps, 20)):
        s0 = time.perf_counter()
        prover.prove_fft_coefficients_resident(buf_qm)
        lat_qm.append(time.perf_counter() - s0)
    for _ in range(args.warmup):
        prover.prove_fft_coefficients_m31_resident(buf)

    ctx.synchronize()
    lat = []
    digest0 = hashlib.sha3_256(first).hexdigest()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s0 = time.perf_counter()
        p = prover.prove_fft_coefficients_m31_resident(buf)
        lat.append(time.perf_counter() - s0)
        if len(p) != len(first):
            return 4
    ctx.synchronize()
    total = time.perf_counter() - t0
    if prover.prove_fft_coefficients_m31_resident(buf) != first:
        print("FATAL: proof changed across replays", file=sys.stderr)
        return 5

    commits_per_s = args.steps / total
    nbytes = commit_bytes(LOG_N, ROUNDS)

    # the same commit when the caller hands over host coefficients (pinned): H2D included
    pin = azk.PinnedArray(m31.shape, np.uint32)
    pin.array[...] = m31
    prover.prove_fft_coefficients_m31(pin.array)
    t_h = time.perf_counter()
    for _ in range(10):
        prover.prove_fft_coefficients_m31(pin.array)
    host_input_per_s = 10 / (time.perf_counter() - t_h)
    pin.close()

    # per-kernel device times (eager launches bracketed by HIP events on the prover's stream)
    prover.set_timing(True)
    ph = []
    for _ in range(5):
        prover.prove_fft_coefficients_m31_resident(buf)
        ph.append(prover.phase_times())
    prover.set_timing(False)
    counts = {k: ph[0].pop(k) for k in ("tree_launches", "tree_levels")}
    for x in ph[1:]:
        for k in counts:
            x.pop(k)
    phases = {k.replace("_ns", "_ms"): float(np.median([x[k] for x in ph])) / 1e6 for k in ph[0]}
    # the upper trees are a serial chain: one dependent one-state-per-wave permutation per level,
    # one graph kernel per launch, and the round transcripts in the last launch of each round (two
    # dependent permutations each: the root, then the challenge frame), all at latencies measured on
    # this chip in the same run
    chain_lat = azk.calibrate_latency(ctx)
    counts["transcript_perms"] = 2 * ROUNDS
    floor_ms = ((counts["tree_levels"] + counts["transcript_perms"]) * chain_lat["wave_perm_us"]
                + counts["tree_launches"] * chain_lat["graph_kernel_us"]) / 1e3
    serial_chain = {"bound": "latency", "kernel": "k_tree_reduce1 (upper Merkle levels, round transcripts)",
                    "latency_calibration": chain_lat, **counts, "tree_ms": phases["tree_ms"], "floor_ms": floor_ms,
                    "frac": floor_ms / phases["tree_ms"] if phases["tree_ms"] else None}
    # dominant kernel: first-layer commitment (final FFT stages fused with 2^21 leaf hashes and
    # 2^10 tile subtrees of 2^11 leaves); algorithmic work = 2^21 + (2^21 - 2^10) permutations
    n = 1 << LOG_N
    perms0 = n + (n - (n >> 11))
    t0_s = phases["commit0_ms"] / 1e3
    # the launch's 11 FFT stages are not counted: a lower bound on its VALU use
    roofline = keccak_roofline(perms0, t0_s, kernel="k_fft_final_m31<true> (final M31 FFT stages, QM31 leaves (a, 0, 0, 0), "
                                                    "leaf hashes, tile subtrees)",
                               traffic=pmc_traffic("k_fft_final_m31<true>"),
                               traffic_source="profiles/pmc_bytes_per_dispatch_2^21.txt (rocprofv3 --pmc FETCH_SIZE, "
                                              "WRITE_SIZE)",
                               algorithmic_permutations_per_launch=perms0, avg_launch_ms=t0_s * 1e3)
    # the same launch against HBM: 2^21 M31 words read, the 2^21 x 16 B QM31 layer written,
    # 2^21 x 4 B twiddles, every tree node below the tile roots written (32 B each) -- far below
    # the HBM roof
    hbm_bytes = n * 4 + n * 16 + n * 4 + perms0 * 32
    roofline["hbm_view"] = {"bound": "hbm", "achieved": hbm_bytes / t0_s / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": hbm_bytes / t0_s / 1e9 / HBM_PEAK_GBS, "algorithmic_bytes": hbm_bytes}

    line = {
        "metric": METRIC, "value": commits_per_s, "unit": "commits/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * total / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u32", "data": "synthetic (SplitMix64 seed 1)",
        "config": {"workload": "circle_pcs_commit", "log_n": LOG_N, "coefficients": "2 x 2^19 M31 (P, Q)",
                   "entry": "azk_circle_prove_fft_coefficients_m31_resident",
                   "rounds": ROUNDS, "queries": SECURITY[1], "log_blowup": SECURITY[0], "leaf_bytes": 16,
                   "hash": "SHA3-256", "proof_bytes": len(first), "proof_sha3": digest0},
        "gbps": nbytes * commits_per_s / 1e9,
        "latency_ms": {"p50": 1e3 * pct(lat, 50), "p99": 1e3 * pct(lat, 99), "min": 1e3 * min(lat)},
        "keccak_perms_per_s": commit_perms(LOG_N, ROUNDS) * commits_per_s,
        "host_input_commits_per_s": host_input_per_s,
        "qm31_input": {"entry": "azk_circle_prove_fft_coefficients_resident (lifted QM31 array, same proof)",
                       "commits_per_s": len(lat_qm) / sum(lat_qm), "p50_ms": 1e3 * pct(lat_qm, 50)},
        "phases_ms": phases, "bit_exact_vs_oracle": checked, "device": caps["name"] or caps["arch"],
        "roofline": roofline,
        "serial_chain": serial_chain,
        "valu_calibration": CALIBRATION,
    }
    if args.batched:
        line["batched"] = batched_form(ctx)
    if args.serving:
        line["serving"] = serving_form(ctx)
    if args.keccak:
        line["keccak_batch"] = keccak_form(ctx)
        line["merkle_plain"] = merkle_plain_form(ctx, check=bool(args.check))
    if args.monomial:
        line["monomial_input"] = monomial_input_form(ctx, prover)
    if args.standalone:
        line["standalone"] = standalone_form(ctx)
        line["standalone"]["witness_basis"] = witness_basis_form(ctx)
    if args.cpu_sample:
        line["cpu_baseline"] = cpu_baseline(coeffs, args.cpu_sample)
        line["cpu_baseline_all_cores"] = cpu_all_cores(coeffs)
        line["cpu_baseline_other_configs"] = cpu_baselines_by_config()
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
