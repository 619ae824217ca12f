"""ctypes binding of libazk.so (include/azk.h).

Python-side mirror of the reference's reusable-plan API (AppleZKProver, Swift): the
class names are the Swift plan names, methods take numpy arrays for host data and
``Buffer`` objects for resident HBM data.  Errors raise ``AZKError`` carrying the
status code that maps one-to-one onto ``AppleZKProverError`` (MetalTypes.swift:3-23).

There is no CPU fallback: if libazk.so or a GPU is missing, constructors raise.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
import re
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libazk.so")
HEADER = os.path.join(os.path.dirname(_HERE), "include", "azk.h")

STATUS = {
    0: "ok", 1: "noDevice", 2: "allocationFailed", 3: "unsupportedOneBlockLength", 4: "invalidLeafCount",
    5: "invalidPrehashedLeafLength", 6: "invalidInputLayout", 7: "correctnessValidationFailed",
    8: "commandExecutionFailed", 9: "unavailableOnThisPlatform", 10: "invalidKernelConfiguration",
    11: "invalidArgument",
}


class AZKError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"{STATUS.get(code, code)}: {message}")
        self.code = code
        self.status = STATUS.get(code, str(code))


class Stats(C.Structure):
    _fields_ = [("wall_ns", C.c_uint64), ("gpu_ns", C.c_uint64), ("submit_ns", C.c_uint64)]


class Capabilities(C.Structure):
    _fields_ = [("name", C.c_char * 64), ("arch", C.c_char * 32), ("compute_units", C.c_int),
                ("wavefront_size", C.c_int), ("lds_bytes_per_workgroup", C.c_size_t), ("hbm_bytes", C.c_size_t),
                ("clock_khz", C.c_int)]


class KeccakfDesc(C.Structure):
    _fields_ = [("count", C.c_uint32), ("input_stride", C.c_uint32), ("output_stride", C.c_uint32)]


class OneBlockDesc(C.Structure):
    _fields_ = [("count", C.c_uint32), ("input_stride", C.c_uint32), ("input_length", C.c_uint32),
                ("output_stride", C.c_uint32)]


class MerkleDesc(C.Structure):
    _fields_ = [("leaf_count", C.c_uint32), ("leaf_stride", C.c_uint32), ("leaf_length", C.c_uint32),
                ("subtree_log", C.c_uint32)]


class ProverDesc(C.Structure):
    _fields_ = [("log_n", C.c_uint32), ("rounds", C.c_uint32), ("log_blowup", C.c_uint32),
                ("query_count", C.c_uint32), ("folding_step", C.c_uint32), ("grinding_bits", C.c_uint32),
                ("public_input_digest", C.c_uint8 * 32)]


class SumcheckResult(C.Structure):
    _fields_ = [("claimed_sum", C.c_uint32), ("final_evaluation", C.c_uint32), ("initial_digest", C.c_uint8 * 32),
                ("round_polys", C.POINTER(C.c_uint32)), ("challenges", C.POINTER(C.c_uint32))]


class ExecPlan(C.Structure):
    _fields_ = [("log_tile_commit", C.c_uint32), ("log_tile_fold", C.c_uint32), ("log_leafwise_below", C.c_uint32),
                ("log_tile_leafwise", C.c_uint32), ("log_tree_chunk", C.c_uint32),
                ("log_leafwise_grid", C.c_uint32), ("tree_waves", C.c_uint32)]


class TuningConfig(C.Structure):
    _fields_ = [("warmup_runs", C.c_uint32), ("measured_runs", C.c_uint32), ("max_candidates", C.c_uint32),
                ("validation_runs", C.c_uint32), ("database_path", C.c_char_p), ("force", C.c_int)]


class TuningResult(C.Structure):
    _fields_ = [("winner", ExecPlan), ("median_gpu_ns", C.c_double), ("p95_gpu_ns", C.c_double),
                ("median_submit_ns", C.c_double), ("default_median_gpu_ns", C.c_double),
                ("candidates_measured", C.c_uint32), ("candidates_rejected", C.c_uint32), ("from_database", C.c_int)]


class DriftStatus(C.Structure):
    _fields_ = [("samples", C.c_uint32), ("stale", C.c_int), ("ema_gpu_ns", C.c_double),
                ("baseline_gpu_ns", C.c_double), ("relative_drift", C.c_double)]


def _plan_dict(x: ExecPlan) -> dict:
    return {k: getattr(x, k) for k, _ in ExecPlan._fields_}


class ServerDesc(C.Structure):
    _fields_ = [("device", C.c_int), ("max_batch", C.c_uint32), ("aux_lanes", C.c_uint32),
                ("reserved", C.c_uint32), ("hbm_budget", C.c_uint64)]


class JobInfo(C.Structure):
    _fields_ = [("status", C.c_int), ("kind", C.c_uint32), ("batch_jobs", C.c_uint32), ("result_bytes", C.c_size_t),
                ("submit_ns", C.c_uint64), ("admit_ns", C.c_uint64), ("start_ns", C.c_uint64), ("done_ns", C.c_uint64)]


class PolynomialClaim(C.Structure):
    _fields_ = [("log_n", C.c_uint32), ("x_count", C.c_uint32), ("y_count", C.c_uint32), ("claim_count", C.c_uint32),
                ("x_coeffs", C.c_void_p), ("y_coeffs", C.c_void_p), ("storage_indices", C.c_void_p),
                ("points", C.c_void_p), ("values", C.c_void_p)]


class ParameterSet(C.Structure):
    _fields_ = [("log_blowup", C.c_uint32), ("queries", C.c_uint32), ("grinding_bits", C.c_uint32),
                ("target_soundness_bits", C.c_uint32)]


class ValuCalibration(C.Structure):
    _fields_ = [("keccak_perms_per_s", C.c_double), ("keccak_clock_mhz", C.c_double),
                ("keccak_cycles_per_wave_perm", C.c_double), ("cycles_per_instr", C.c_double * 6),
                ("clock_mhz", C.c_double * 6)]


VALU_CALIBRATION_INSTRUCTIONS = ("v_xor_b32", "v_bitop3_b32", "v_alignbit_b32", "v_add_u32", "v_min_u32", "v_mad_u64_u32")


class PhaseTimes(C.Structure):
    _fields_ = [("fft_ns", C.c_uint64), ("commit0_ns", C.c_uint64), ("fold_ns", C.c_uint64),
                ("tree_ns", C.c_uint64), ("transcript_ns", C.c_uint64), ("emit_ns", C.c_uint64),
                ("host_ns", C.c_uint64), ("total_ns", C.c_uint64), ("tree_launches", C.c_uint32),
                ("tree_levels", C.c_uint32)]


class LatencyCalibration(C.Structure):
    _fields_ = [("wave_perm_us", C.c_double), ("pair_perm_us", C.c_double), ("graph_kernel_us", C.c_double)]


_lib = None
vp = C.c_void_p
sz = C.c_size_t
u32 = C.c_uint32


def declared_symbols() -> list[str]:
    """Every function name declared in include/azk.h."""
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(azk_[a-z0-9_]+)\s*\(", text)))


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise AZKError(9, f"{LIB_PATH} is not built (run __graft_entry__.build())")
    lib = C.CDLL(LIB_PATH)
    st = C.c_int
    sp = C.POINTER(Stats)
    sigs = {
        "azk_last_error": (C.c_char_p, []), "azk_version": (C.c_char_p, []),
        "azk_context_create": (st, [C.c_int, C.POINTER(vp)]), "azk_context_destroy": (None, [vp]),
        "azk_context_capabilities": (st, [vp, C.POINTER(Capabilities)]), "azk_context_synchronize": (st, [vp]),
        "azk_buffer_alloc": (st, [vp, sz, C.POINTER(vp)]), "azk_buffer_wrap": (st, [vp, vp, sz, C.POINTER(vp)]),
        "azk_buffer_free": (None, [vp]), "azk_buffer_size": (sz, [vp]), "azk_buffer_device_ptr": (vp, [vp]),
        "azk_buffer_upload": (st, [vp, sz, vp, sz]), "azk_buffer_download": (st, [vp, sz, vp, sz]),
        "azk_keccakf_plan_create": (st, [vp, C.POINTER(KeccakfDesc), C.POINTER(vp)]),
        "azk_keccakf_plan_destroy": (None, [vp]),
        "azk_keccakf_permute": (st, [vp, vp, sz, vp, sz, sp]),
        "azk_keccakf_permute_resident": (st, [vp, vp, sz, vp, sz, sp]),
        "azk_keccakf_permute_verified": (st, [vp, vp, sz, vp, sz, sp]),
        "azk_keccakf_plan_clear": (st, [vp]),
        "azk_sha3_256_plan_create": (st, [vp, C.POINTER(OneBlockDesc), C.POINTER(vp)]),
        "azk_keccak_256_plan_create": (st, [vp, C.POINTER(OneBlockDesc), C.POINTER(vp)]),
        "azk_hash_plan_destroy": (None, [vp]),
        "azk_hash_execute": (st, [vp, vp, sz, vp, sz, sp]),
        "azk_hash_execute_resident": (st, [vp, vp, sz, vp, sz, sp]),
        "azk_hash_execute_verified": (st, [vp, vp, sz, vp, sz, sp]),
        "azk_hash_plan_clear": (st, [vp]),
        "azk_merkle_plan_create": (st, [vp, C.POINTER(MerkleDesc), C.POINTER(vp)]),
        "azk_merkle_plan_destroy": (None, [vp]),
        "azk_merkle_commit": (st, [vp, vp, sz, vp, sp]),
        "azk_merkle_commit_resident": (st, [vp, vp, sz, vp, sp]),
        "azk_merkle_commit_verified": (st, [vp, vp, sz, vp, sp]),
        "azk_merkle_open": (st, [vp, u32, vp, sz, vp]),
        "azk_merkle_open_batch": (st, [vp, vp, sz, vp, sz, vp]),
        "azk_merkle_plan_clear": (st, [vp]),
        "azk_merkle_commit_noalloc": (st, [vp, vp, sz, u32, u32, u32, vp, sz, vp]),
        "azk_prehashed_plan_create": (st, [vp, u32, C.POINTER(vp)]),
        "azk_prehashed_plan_destroy": (None, [vp]),
        "azk_prehashed_commit_resident": (st, [vp, vp, sz, vp, sp]),
        "azk_codeword_plan_create": (st, [vp, u32, C.POINTER(vp)]),
        "azk_codeword_plan_destroy": (None, [vp]),
        "azk_codeword_execute_resident": (st, [vp, vp, sz, vp, sz, sp]),
        "azk_codeword_execute": (st, [vp, vp, sz, vp, sz, sp]),
        "azk_codeword_execute_m31_resident": (st, [vp, vp, sz, vp, sz, sp]),
        "azk_codeword_execute_m31": (st, [vp, vp, sz, vp, sz, sp]),
        "azk_codeword_read_twiddles": (st, [vp, vp, sz]),
        "azk_witness_basis_plan_create": (st, [vp, u32, C.POINTER(vp)]),
        "azk_witness_basis_plan_destroy": (None, [vp]),
        "azk_witness_basis_execute_resident": (st, [vp, vp, sz, u32, vp, sz, u32, vp, sz, sp]),
        "azk_witness_basis_execute": (st, [vp, vp, u32, vp, u32, vp, sz, sp]),
        "azk_fold_plan_create": (st, [vp, u32, u32, C.POINTER(vp)]),
        "azk_fold_plan_destroy": (None, [vp]),
        "azk_fold_execute_resident": (st, [vp, vp, sz, vp, vp, sz, sp]),
        "azk_fold_execute": (st, [vp, vp, sz, vp, vp, sz, sp]),
        "azk_circle_prover_create": (st, [vp, C.POINTER(ProverDesc), C.POINTER(vp)]),
        "azk_circle_prover_destroy": (None, [vp]),
        "azk_circle_prover_proof_size": (sz, [vp]),
        "azk_circle_prove_fft_coefficients_resident": (st, [vp, vp, sz, vp, sz, C.POINTER(sz), sp]),
        "azk_circle_prove_evaluations_resident": (st, [vp, vp, sz, vp, sz, C.POINTER(sz), sp]),
        "azk_circle_prove_fft_coefficients": (st, [vp, vp, sz, vp, sz, C.POINTER(sz), sp]),
        "azk_circle_prove_fft_coefficients_m31_resident": (st, [vp, vp, sz, vp, sz, C.POINTER(sz), sp]),
        "azk_circle_prove_fft_coefficients_m31": (st, [vp, vp, sz, vp, sz, C.POINTER(sz), sp]),
        "azk_circle_prover_clear": (st, [vp]),
        "azk_circle_prover_set_graph": (st, [vp, C.c_int]),
        "azk_circle_prover_set_timing": (st, [vp, C.c_int]),
        "azk_circle_prover_phase_times": (st, [vp, C.POINTER(PhaseTimes)]),
        "azk_circle_batch_create": (st, [C.POINTER(vp), C.POINTER(vp), C.POINTER(sz), u32, C.POINTER(vp)]),
        "azk_circle_batch_create_m31": (st, [C.POINTER(vp), C.POINTER(vp), C.POINTER(sz), u32, C.POINTER(vp)]),
        "azk_circle_batch_destroy": (None, [vp]),
        "azk_circle_batch_tune": (st, [vp, C.POINTER(TuningConfig), C.POINTER(TuningResult)]),
        "azk_circle_batch_prove": (st, [vp, C.POINTER(vp), C.POINTER(sz), C.POINTER(sz), sp]),
        "azk_circle_prove_batch": (st, [C.POINTER(vp), C.POINTER(vp), C.POINTER(sz), u32, C.POINTER(vp),
                                        C.POINTER(sz), C.POINTER(sz), sp]),
        "azk_sumcheck_plan_create": (st, [vp, u32, C.POINTER(vp)]),
        "azk_sumcheck_plan_destroy": (None, [vp]),
        "azk_sumcheck_proof_size": (sz, [u32]),
        "azk_sumcheck_prove": (st, [vp, vp, sz, vp, vp, sz, C.POINTER(sz), sp]),
        "azk_sumcheck_prove_resident": (st, [vp, vp, sz, vp, vp, C.POINTER(SumcheckResult), sp]),
        "azk_sumcheck_verify": (st, [vp, sz, C.POINTER(C.c_int)]),
        "azk_circle_verify": (st, [vp, sz, vp, C.POINTER(C.c_int)]),
        "azk_circle_verify_with_coefficients": (st, [vp, sz, vp, vp, sz, C.POINTER(C.c_int)]),
        "azk_circle_prove_fft_coefficients_verified": (st, [vp, vp, sz, vp, sz, C.POINTER(sz), sp]),
        "azk_circle_prove_fft_coefficients_resident_verified": (st, [vp, vp, sz, vp, sz, C.POINTER(sz), sp]),
        "azk_codeword_execute_verified": (st, [vp, vp, sz, vp, sz, sp]),
        "azk_fold_execute_verified": (st, [vp, vp, sz, vp, vp, sz, sp]),
        "azk_witness_basis_execute_verified": (st, [vp, vp, u32, vp, u32, vp, sz, sp]),
        "azk_host_circle_fft": (st, [vp, u32, vp]),
        "azk_host_circle_prove": (st, [vp, C.POINTER(ProverDesc), vp, sz, C.POINTER(sz)]),
        "azk_circle_grinding_digest_meets_target": (C.c_int, [vp, sz, u32]),
        "azk_circle_polynomial_claim_make": (st, [C.POINTER(PolynomialClaim), vp, vp]),
        "azk_circle_polynomial_claim_digest": (st, [C.POINTER(PolynomialClaim), vp]),
        "azk_circle_verify_polynomial": (st, [vp, sz, C.POINTER(PolynomialClaim), C.POINTER(C.c_int)]),
        "azk_circle_verify_contract": (st, [vp, sz, C.POINTER(ParameterSet), C.POINTER(PolynomialClaim),
                                            C.POINTER(C.c_int)]),
        "azk_chain_plan_create": (st, [vp, C.POINTER(ProverDesc), C.POINTER(vp)]),
        "azk_chain_plan_destroy": (None, [vp]),
        "azk_chain_execute_resident": (st, [vp, vp, sz, vp, vp, vp, sz, sp]),
        "azk_chain_execute": (st, [vp, vp, sz, vp, vp, vp, sz, sp]),
        "azk_chain_execute_verified": (st, [vp, vp, sz, vp, vp, vp, sz, sp]),
        "azk_host_fold_chain": (st, [vp, u32, vp, u32, vp]),
        "azk_witness_basis_check": (st, [vp, u32, vp, u32, u32, vp, C.POINTER(C.c_int)]),
        "azk_circle_prover_set_exec_plan": (st, [vp, C.POINTER(ExecPlan)]),
        "azk_circle_prover_get_exec_plan": (st, [vp, C.POINTER(ExecPlan)]),
        "azk_circle_prover_tune": (st, [vp, vp, sz, C.POINTER(TuningConfig), C.POINTER(TuningResult)]),
        "azk_circle_prover_drift": (st, [vp, C.POINTER(DriftStatus)]),
        "azk_server_create": (st, [C.POINTER(ServerDesc), C.POINTER(vp)]),
        "azk_server_destroy": (None, [vp]),
        "azk_server_submit_commit": (st, [vp, C.POINTER(ProverDesc), vp, sz, C.POINTER(C.c_uint64)]),
        "azk_server_submit_commit_m31": (st, [vp, C.POINTER(ProverDesc), vp, sz, C.POINTER(C.c_uint64)]),
        "azk_server_submit_sumcheck": (st, [vp, vp, sz, C.POINTER(C.c_uint64)]),
        "azk_server_submit_sumcheck_digest": (st, [vp, vp, sz, vp, C.POINTER(C.c_uint64)]),
        "azk_server_submit_sumcheck_chunk": (st, [vp, vp, sz, u32, C.POINTER(C.c_uint64)]),
        "azk_sumcheck_prove_with_digest": (st, [vp, vp, sz, vp, vp, vp, sz, C.POINTER(sz), sp]),
        "azk_server_submit_fold": (st, [vp, u32, u32, vp, sz, vp, C.POINTER(C.c_uint64)]),
        "azk_server_wait": (st, [vp, C.c_uint64, C.POINTER(JobInfo)]),
        "azk_server_take": (st, [vp, C.c_uint64, vp, sz, C.POINTER(sz)]),
        "azk_now_ns": (C.c_uint64, []),
        "azk_host_alloc": (st, [sz, C.POINTER(vp)]),
        "azk_calibrate_valu": (st, [vp, C.POINTER(ValuCalibration)]),
        "azk_calibrate_latency": (st, [vp, C.POINTER(LatencyCalibration)]),
        "azk_sumcheck_chunk_plan_create": (st, [vp, C.c_uint64, u32, C.POINTER(vp)]),
        "azk_sumcheck_chunk_plan_destroy": (None, [vp]),
        "azk_sumcheck_chunk_coefficient_words": (sz, [C.c_uint64, u32]),
        "azk_sumcheck_chunk_execute": (st, [vp, vp, sz, vp, vp, vp, sp]),
        "azk_sumcheck_chunk_execute_resident": (st, [vp, vp, sz, vp, vp, vp, sp]),
        "azk_sumcheck_chunk_execute_verified": (st, [vp, vp, sz, vp, vp, vp, sp]),
        "azk_sumcheck_chunk_plan_clear": (st, [vp]),
        "azk_sumcheck_chunk_proof_size": (sz, [C.c_uint64, u32]),
        "azk_sumcheck_chunk_prove": (st, [vp, vp, sz, vp, sz, C.POINTER(sz), sp]),
        "azk_sumcheck_chunk_verify": (st, [vp, sz, C.POINTER(C.c_int)]),
        "azk_domain_plan_create": (st, [vp, u32, C.c_int, C.c_int, u32, C.POINTER(vp)]),
        "azk_domain_plan_destroy": (None, [vp]),
        "azk_domain_plan_stats": (st, [vp, sp]),
        "azk_domain_inverse_layout": (st, [vp, vp, vp, u32]),
        "azk_domain_read": (st, [vp, C.c_int, vp, sz]),
        "azk_domain_copy_resident": (st, [vp, C.c_int, vp, sz]),
        "azk_host_free": (None, [vp]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


_live: "weakref.WeakSet" = weakref.WeakSet()


def _track(obj):
    _live.add(obj)
    return obj


@atexit.register
def _close_all() -> None:
    """Frees what is still alive at interpreter exit while the HIP runtime is up:
    servers and batches, then plans, buffers and pinned memory, then contexts (a
    context is freed by the library after the last object built on it)."""
    for cls in (Server, CircleBatchPlan, _Plan, Buffer, PinnedArray, Context):
        for obj in list(_live):
            if isinstance(obj, cls):
                try:
                    (obj.free if isinstance(obj, Buffer) else obj.close)()
                except Exception:
                    pass


def _ck(rc: int) -> None:
    if rc != 0:
        raise AZKError(rc, load().azk_last_error().decode())


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def _host(a) -> np.ndarray:
    if isinstance(a, (bytes, bytearray)):
        a = np.frombuffer(bytes(a), np.uint8)
    a = np.ascontiguousarray(a)
    return a.view(np.uint8).reshape(-1) if a.size else np.zeros(1, np.uint8)


def _stats_dict(s: Stats) -> dict:
    return {"wall_ns": s.wall_ns, "gpu_ns": s.gpu_ns, "submit_ns": s.submit_ns}


class Context:
    """MetalContext (MetalContext.swift:76-112): one device, one stream."""

    def __init__(self, device: int = 0):
        lib = load()
        h = vp()
        _ck(lib.azk_context_create(device, C.byref(h)))
        self.h = h
        _track(self)
        self._lib = lib

    def capabilities(self) -> dict:
        c = Capabilities()
        _ck(self._lib.azk_context_capabilities(self.h, C.byref(c)))
        return {"name": c.name.decode(), "arch": c.arch.decode(), "compute_units": c.compute_units,
                "wavefront_size": c.wavefront_size, "lds_bytes": c.lds_bytes_per_workgroup,
                "hbm_bytes": c.hbm_bytes, "clock_khz": c.clock_khz}

    def synchronize(self) -> None:
        _ck(self._lib.azk_context_synchronize(self.h))

    def buffer(self, nbytes: int) -> "Buffer":
        return Buffer(self, nbytes)

    def upload(self, data) -> "Buffer":
        a = _host(data)
        b = Buffer(self, a.size)
        b.upload(a)
        return b

    def close(self) -> None:
        if self.h:
            self._lib.azk_context_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Buffer:
    """Resident HBM buffer (MTLBuffer + offset pairs in the reference)."""

    def __init__(self, ctx: Context, nbytes: int, device_ptr: int | None = None):
        self.ctx = ctx
        h = vp()
        if device_ptr is None:
            _ck(ctx._lib.azk_buffer_alloc(ctx.h, nbytes, C.byref(h)))
        else:
            _ck(ctx._lib.azk_buffer_wrap(ctx.h, device_ptr, nbytes, C.byref(h)))
        self.h = h
        _track(self)
        self.nbytes = nbytes

    @property
    def device_ptr(self) -> int:
        return self.ctx._lib.azk_buffer_device_ptr(self.h)

    def upload(self, data, offset: int = 0) -> None:
        a = _host(data)
        _ck(self.ctx._lib.azk_buffer_upload(self.h, offset, _ptr(a), a.size if len(data) else 0))

    def download(self, nbytes: int | None = None, offset: int = 0) -> np.ndarray:
        n = self.nbytes - offset if nbytes is None else nbytes
        out = np.zeros(max(n, 1), np.uint8)
        _ck(self.ctx._lib.azk_buffer_download(self.h, offset, _ptr(out), n))
        return out[:n]

    def free(self) -> None:
        if self.h:
            self.ctx._lib.azk_buffer_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class _Plan:
    _destroy = ""

    def close(self) -> None:
        if getattr(self, "h", None):
            getattr(self.ctx._lib, self._destroy)(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class KeccakF1600PermutationBatchPlan(_Plan):
    """KeccakF1600PermutationBatcher.swift:73-294."""
    _destroy = "azk_keccakf_plan_destroy"

    def __init__(self, ctx: Context, count: int, input_stride: int = 200, output_stride: int = 200):
        self.ctx, self.count, self.input_stride, self.output_stride = ctx, count, input_stride, output_stride
        h = vp()
        _ck(ctx._lib.azk_keccakf_plan_create(ctx.h, C.byref(KeccakfDesc(count, input_stride, output_stride)), C.byref(h)))
        self.h = h
        _track(self)
        self.last_stats: dict = {}

    def permute(self, states, verified: bool = False) -> np.ndarray:
        a = _host(states)
        out = np.zeros(self.count * self.output_stride, np.uint8)
        s = Stats()
        fn = self.ctx._lib.azk_keccakf_permute_verified if verified else self.ctx._lib.azk_keccakf_permute
        _ck(fn(self.h, _ptr(a), a.size, _ptr(out), out.size, C.byref(s)))
        self.last_stats = _stats_dict(s)
        return out

    def permute_resident(self, inp: Buffer, out: Buffer, in_offset: int = 0, out_offset: int = 0) -> dict:
        s = Stats()
        _ck(self.ctx._lib.azk_keccakf_permute_resident(self.h, inp.h, in_offset, out.h, out_offset, C.byref(s)))
        return _stats_dict(s)

    def clear(self) -> None:
        _ck(self.ctx._lib.azk_keccakf_plan_clear(self.h))


class _OneBlockPlan(_Plan):
    _destroy = "azk_hash_plan_destroy"
    _create = ""

    def __init__(self, ctx: Context, count: int, input_stride: int, input_length: int, output_stride: int = 32):
        self.ctx, self.count = ctx, count
        self.input_stride, self.input_length, self.output_stride = input_stride, input_length, output_stride
        h = vp()
        d = OneBlockDesc(count, input_stride, input_length, output_stride)
        _ck(getattr(ctx._lib, self._create)(ctx.h, C.byref(d), C.byref(h)))
        self.h = h
        _track(self)
        self.last_stats: dict = {}

    def hash(self, messages, verified: bool = False) -> np.ndarray:
        a = _host(messages)
        out = np.zeros(max(self.count * self.output_stride, 1), np.uint8)
        s = Stats()
        fn = self.ctx._lib.azk_hash_execute_verified if verified else self.ctx._lib.azk_hash_execute
        _ck(fn(self.h, _ptr(a), a.size if self.count * self.input_stride else 0, _ptr(out),
               self.count * self.output_stride, C.byref(s)))
        self.last_stats = _stats_dict(s)
        return out[: self.count * self.output_stride]

    def hash_resident(self, inp: Buffer, out: Buffer, in_offset: int = 0, out_offset: int = 0) -> dict:
        s = Stats()
        _ck(self.ctx._lib.azk_hash_execute_resident(self.h, inp.h, in_offset, out.h, out_offset, C.byref(s)))
        return _stats_dict(s)

    def clear(self) -> None:
        _ck(self.ctx._lib.azk_hash_plan_clear(self.h))


class SHA3FixedOneBlockHashPlan(_OneBlockPlan):
    """SHA3BatchHasher.swift:81-264."""
    _create = "azk_sha3_256_plan_create"


class Keccak256FixedOneBlockHashPlan(_OneBlockPlan):
    """Keccak256BatchHasher.swift:73-257."""
    _create = "azk_keccak_256_plan_create"


class SHA3RawLeavesMerkleCommitPlan(_Plan):
    """MerkleCommitter.swift:526-1182 (tree kept resident; openings gathered from it)."""
    _destroy = "azk_merkle_plan_destroy"

    def __init__(self, ctx: Context, leaf_count: int, leaf_stride: int, leaf_length: int, subtree_log: int = 0):
        self.ctx, self.leaf_count, self.leaf_stride, self.leaf_length = ctx, leaf_count, leaf_stride, leaf_length
        h = vp()
        _ck(ctx._lib.azk_merkle_plan_create(ctx.h, C.byref(MerkleDesc(leaf_count, leaf_stride, leaf_length, subtree_log)),
                                            C.byref(h)))
        self.h = h
        _track(self)
        self.depth = leaf_count.bit_length() - 1
        self.last_stats: dict = {}

    def commit(self, leaves, verified: bool = False) -> bytes:
        a = _host(leaves)
        root = np.zeros(32, np.uint8)
        s = Stats()
        fn = self.ctx._lib.azk_merkle_commit_verified if verified else self.ctx._lib.azk_merkle_commit
        _ck(fn(self.h, _ptr(a), a.size if self.leaf_count * self.leaf_stride else 0, _ptr(root), C.byref(s)))
        self.last_stats = _stats_dict(s)
        return root.tobytes()

    def commit_resident(self, leaves: Buffer, offset: int = 0) -> bytes:
        root = np.zeros(32, np.uint8)
        s = Stats()
        _ck(self.ctx._lib.azk_merkle_commit_resident(self.h, leaves.h, offset, _ptr(root), C.byref(s)))
        self.last_stats = _stats_dict(s)
        return root.tobytes()

    def open(self, index: int) -> tuple[bytes, bytes]:
        sib = np.zeros(max(32 * self.depth, 1), np.uint8)
        root = np.zeros(32, np.uint8)
        _ck(self.ctx._lib.azk_merkle_open(self.h, index, _ptr(sib), 32 * self.depth, _ptr(root)))
        return sib[: 32 * self.depth].tobytes(), root.tobytes()

    def open_batch(self, indices) -> tuple[list[bytes], bytes]:
        """Openings of every index in one call: (sibling paths, root)."""
        idx = np.ascontiguousarray(indices, dtype=np.uint32)
        per = 32 * self.depth
        sib = np.zeros(max(per * idx.size, 1), np.uint8)
        root = np.zeros(32, np.uint8)
        _ck(self.ctx._lib.azk_merkle_open_batch(self.h, _ptr(idx) if idx.size else None, idx.size, _ptr(sib),
                                                 per * idx.size, _ptr(root)))
        return [sib[per * k: per * (k + 1)].tobytes() for k in range(idx.size)], root.tobytes()

    def clear(self) -> None:
        _ck(self.ctx._lib.azk_merkle_plan_clear(self.h))


def merkle_commit_noalloc(ctx: "Context", leaves: "Buffer", leaf_count: int, leaf_stride: int, leaf_length: int,
                          root: "Buffer", leaves_offset: int = 0, root_offset: int = 0) -> None:
    """Allocation-free raw-leaf Merkle root into device memory, enqueued on the context stream
    (synchronize before reading `root`)."""
    _ck(load().azk_merkle_commit_noalloc(ctx.h, leaves.h, leaves_offset, leaf_count, leaf_stride, leaf_length, root.h,
                                         root_offset, None))


class SHA3PrehashedLeavesMerkleCommitPlan(_Plan):
    """MerkleCommitter.swift:314-524."""
    _destroy = "azk_prehashed_plan_destroy"

    def __init__(self, ctx: Context, leaf_count: int):
        self.ctx = ctx
        h = vp()
        _ck(ctx._lib.azk_prehashed_plan_create(ctx.h, leaf_count, C.byref(h)))
        self.h = h
        _track(self)

    def commit_resident(self, leaf_hashes: Buffer, offset: int = 0) -> bytes:
        root = np.zeros(32, np.uint8)
        s = Stats()
        _ck(self.ctx._lib.azk_prehashed_commit_resident(self.h, leaf_hashes.h, offset, _ptr(root), C.byref(s)))
        return root.tobytes()


class CircleCodewordPlan(_Plan):
    """CircleCodeword.swift:411-887 (input: interleaved line-basis coefficients)."""
    _destroy = "azk_codeword_plan_destroy"

    def __init__(self, ctx: Context, log_n: int):
        self.ctx, self.log_n = ctx, log_n
        h = vp()
        _ck(ctx._lib.azk_codeword_plan_create(ctx.h, log_n, C.byref(h)))
        self.h = h
        _track(self)
        self.last_stats: dict = {}

    def execute(self, coeffs: np.ndarray, verified: bool = False) -> np.ndarray:
        a = _host(np.ascontiguousarray(coeffs, dtype=np.uint32))
        out = np.zeros(4 << self.log_n, np.uint32)
        s = Stats()
        fn = self.ctx._lib.azk_codeword_execute_verified if verified else self.ctx._lib.azk_codeword_execute
        _ck(fn(self.h, _ptr(a), a.size, _ptr(out), out.nbytes, C.byref(s)))
        self.last_stats = _stats_dict(s)
        return out

    def execute_resident(self, coeffs: Buffer, out: Buffer, coeff_offset: int = 0, out_offset: int = 0) -> dict:
        s = Stats()
        _ck(self.ctx._lib.azk_codeword_execute_resident(self.h, coeffs.h, coeff_offset, out.h, out_offset, C.byref(s)))
        return _stats_dict(s)

    def execute_m31(self, coeffs: np.ndarray) -> np.ndarray:
        """N M31 coefficients (the first components of a QM31 lift) -> N M31 evaluations."""
        a = np.ascontiguousarray(coeffs, dtype=np.uint32)
        out = np.zeros(1 << self.log_n, np.uint32)
        s = Stats()
        _ck(self.ctx._lib.azk_codeword_execute_m31(self.h, _ptr(a), a.size, _ptr(out), out.size, C.byref(s)))
        self.last_stats = _stats_dict(s)
        return out

    def execute_m31_resident(self, coeffs: Buffer, out: Buffer, coeff_offset: int = 0, out_offset: int = 0) -> dict:
        s = Stats()
        _ck(self.ctx._lib.azk_codeword_execute_m31_resident(self.h, coeffs.h, coeff_offset, out.h, out_offset,
                                                              C.byref(s)))
        return _stats_dict(s)

    def read_twiddles(self) -> np.ndarray:
        n = (1 << self.log_n) - 1
        out = np.zeros(max(n, 1), np.uint32)
        _ck(self.ctx._lib.azk_codeword_read_twiddles(self.h, _ptr(out), n))
        return out[:n]


class CircleWitnessToFFTBasisPlanV1(_Plan):
    """CircleWitnessFFTBasis.swift:230-517: x / y monomial witness columns (QM31 rows) ->
    2^log_n interleaved Circle-FFT-basis coefficients (CircleCodeword.swift:100-127)."""
    _destroy = "azk_witness_basis_plan_destroy"

    def __init__(self, ctx: Context, log_n: int):
        self.ctx, self.log_n = ctx, log_n
        h = vp()
        _ck(ctx._lib.azk_witness_basis_plan_create(ctx.h, log_n, C.byref(h)))
        self.h = h
        _track(self)
        self.last_stats: dict = {}

    def execute(self, x: np.ndarray, y: np.ndarray, verified: bool = False) -> np.ndarray:
        xa = np.ascontiguousarray(np.asarray(x, dtype=np.uint32).reshape(-1, 4))
        ya = np.ascontiguousarray(np.asarray(y, dtype=np.uint32).reshape(-1, 4))
        out = np.zeros(4 << self.log_n, np.uint32)
        s = Stats()
        fn = self.ctx._lib.azk_witness_basis_execute_verified if verified else self.ctx._lib.azk_witness_basis_execute
        _ck(fn(self.h, _ptr(xa) if len(xa) else None, len(xa), _ptr(ya) if len(ya) else None, len(ya), _ptr(out),
               out.size, C.byref(s)))
        self.last_stats = _stats_dict(s)
        return out

    def execute_resident(self, x: Buffer | None, x_count: int, y: Buffer | None, y_count: int, out: Buffer,
                         x_offset: int = 0, y_offset: int = 0, out_offset: int = 0) -> dict:
        s = Stats()
        _ck(self.ctx._lib.azk_witness_basis_execute_resident(
            self.h, x.h if x is not None else None, x_offset, x_count, y.h if y is not None else None, y_offset,
            y_count, out.h, out_offset, C.byref(s)))
        return _stats_dict(s)


class CircleDomainMaterializationPlan(_Plan):
    """CircleDomainMaterializationPlan: points, FFT twiddles and QM31 inverse-domain layers of
    the canonical 2^log_n domain, materialized on the device at creation."""
    POINTS, FFT_TWIDDLES, INVERSE_LAYERS = 0, 1, 2
    _destroy = "azk_domain_plan_destroy"

    def __init__(self, ctx: Context, log_n: int, materialize_domain_points: bool = True,
                 materialize_codeword_twiddles: bool = False, inverse_domain_round_count: int = 0):
        self.ctx, self.log_n, self.rounds = ctx, log_n, inverse_domain_round_count
        h = vp()
        _ck(ctx._lib.azk_domain_plan_create(ctx.h, log_n, int(materialize_domain_points),
                                            int(materialize_codeword_twiddles), inverse_domain_round_count, C.byref(h)))
        self.h = h
        _track(self)

    def materialization_stats(self) -> dict:
        st = Stats()
        _ck(self.ctx._lib.azk_domain_plan_stats(self.h, C.byref(st)))
        return _stats_dict(st)

    def inverse_domain_layout(self) -> tuple[list, list]:
        counts = np.zeros(max(self.rounds, 1), np.uint64)
        offsets = np.zeros(max(self.rounds, 1), np.uint64)
        _ck(self.ctx._lib.azk_domain_inverse_layout(self.h, _ptr(counts), _ptr(offsets), self.rounds))
        return counts[: self.rounds].tolist(), offsets[: self.rounds].tolist()

    def _read(self, which: int, words: int) -> np.ndarray:
        out = np.zeros(max(words, 1), np.uint32)
        _ck(self.ctx._lib.azk_domain_read(self.h, which, _ptr(out), 4 * words))
        return out[:words]

    def read_domain_points(self) -> np.ndarray:
        return self._read(self.POINTS, 2 << self.log_n).reshape(-1, 2)

    def read_codeword_twiddles(self) -> np.ndarray:
        return self._read(self.FFT_TWIDDLES, (1 << self.log_n) - 1)

    def read_flat_inverse_domain(self) -> np.ndarray:
        n = 1 << self.log_n
        return self._read(self.INVERSE_LAYERS, 4 * (n - (n >> self.rounds))).reshape(-1, 4)

    def read_inverse_domain_layers(self) -> list:
        flat = self.read_flat_inverse_domain()
        counts, offsets = self.inverse_domain_layout()
        return [flat[o:o + c] for c, o in zip(counts, offsets)]

    def copy_resident(self, which: int, dst: Buffer, offset: int = 0) -> None:
        _ck(self.ctx._lib.azk_domain_copy_resident(self.h, which, dst.h, offset))


class CircleFRIFoldChainPlan(_Plan):
    """CircleFRIFoldPlan.swift:5-280 (explicit challenges)."""
    _destroy = "azk_fold_plan_destroy"

    def __init__(self, ctx: Context, log_n: int, rounds: int):
        self.ctx, self.log_n, self.rounds = ctx, log_n, rounds
        h = vp()
        _ck(ctx._lib.azk_fold_plan_create(ctx.h, log_n, rounds, C.byref(h)))
        self.h = h
        _track(self)

    def execute(self, evals: np.ndarray, challenges, verified: bool = False) -> np.ndarray:
        a = _host(np.ascontiguousarray(evals, dtype=np.uint32))
        ch = np.ascontiguousarray(np.asarray(challenges, dtype=np.uint32).reshape(-1))
        out = np.zeros(4 << (self.log_n - self.rounds), np.uint32)
        s = Stats()
        fn = self.ctx._lib.azk_fold_execute_verified if verified else self.ctx._lib.azk_fold_execute
        _ck(fn(self.h, _ptr(a), a.size, _ptr(ch), _ptr(out), out.nbytes, C.byref(s)))
        return out

    def execute_resident(self, evals: Buffer, challenges, out: Buffer, evals_offset: int = 0, out_offset: int = 0) -> dict:
        ch = np.ascontiguousarray(np.asarray(challenges, dtype=np.uint32).reshape(-1))
        s = Stats()
        _ck(self.ctx._lib.azk_fold_execute_resident(self.h, evals.h, evals_offset, _ptr(ch), out.h, out_offset, C.byref(s)))
        return _stats_dict(s)


class CircleFRIMerkleTranscriptFoldChainPlan(_Plan):
    """CircleFRIFoldPlan.swift:301-649: per round commit, absorb the root, squeeze the
    challenge, fold; returns the roots, the challenges and the final layer."""
    _destroy = "azk_chain_plan_destroy"

    def __init__(self, ctx: Context, log_n: int, rounds: int, security=(2, 32, 1, 0),
                 public_input_digest: bytes = b""):
        self.ctx, self.log_n, self.rounds = ctx, log_n, rounds
        d = prover_desc(log_n, rounds, security, public_input_digest)
        h = vp()
        _ck(ctx._lib.azk_chain_plan_create(ctx.h, C.byref(d), C.byref(h)))
        self.h = h
        _track(self)
        self.last_stats: dict = {}

    def execute(self, evals: np.ndarray, verified: bool = False) -> dict:
        e = np.ascontiguousarray(evals, dtype=np.uint32).reshape(-1)
        roots = np.zeros(32 * self.rounds, np.uint8)
        ch = np.zeros(4 * self.rounds, np.uint32)
        fin = np.zeros(4 << (self.log_n - self.rounds), np.uint32)
        s = Stats()
        fn = self.ctx._lib.azk_chain_execute_verified if verified else self.ctx._lib.azk_chain_execute
        _ck(fn(self.h, _ptr(e), e.size, _ptr(roots), _ptr(ch), _ptr(fin), fin.size, C.byref(s)))
        self.last_stats = _stats_dict(s)
        return {"commitments": [roots[32 * r:32 * (r + 1)].tobytes() for r in range(self.rounds)],
                "challenges": ch.reshape(-1, 4), "values": fin}

    def execute_resident(self, evals: Buffer, out: Buffer, evals_offset: int = 0, out_offset: int = 0) -> dict:
        roots = np.zeros(32 * self.rounds, np.uint8)
        ch = np.zeros(4 * self.rounds, np.uint32)
        s = Stats()
        _ck(self.ctx._lib.azk_chain_execute_resident(self.h, evals.h, evals_offset, _ptr(roots), _ptr(ch), out.h,
                                                     out_offset, C.byref(s)))
        return {"commitments": [roots[32 * r:32 * (r + 1)].tobytes() for r in range(self.rounds)],
                "challenges": ch.reshape(-1, 4), "stats": _stats_dict(s)}


class CircleCodewordPCSFRIProverV1(_Plan):
    """CircleCodeword.swift:920-1419 + CirclePCSFRIResidentProverV1 (CircleFRIFoldPlan.swift:693-1025)."""
    _destroy = "azk_circle_prover_destroy"

    def __init__(self, ctx: Context, log_n: int, rounds: int, security=(2, 32, 1, 0), public_input_digest: bytes = b""):
        self.ctx, self.log_n, self.rounds = ctx, log_n, rounds
        d = ProverDesc()
        d.log_n, d.rounds = log_n, rounds
        d.log_blowup, d.query_count, d.folding_step, d.grinding_bits = security
        dg = public_input_digest or bytes(32)
        if len(dg) != 32:
            raise AZKError(6, "public input digest must be 32 bytes")
        d.public_input_digest[:] = list(dg)
        h = vp()
        _ck(ctx._lib.azk_circle_prover_create(ctx.h, C.byref(d), C.byref(h)))
        self.h = h
        _track(self)
        self.proof_size = int(ctx._lib.azk_circle_prover_proof_size(h))
        self._out = np.zeros(self.proof_size, np.uint8)
        self.last_stats: dict = {}

    def _finish(self, rc, n, s) -> bytes:
        _ck(rc)
        self.last_stats = _stats_dict(s)
        return self._out[: n.value].tobytes()

    def prove_fft_coefficients_resident(self, coeffs: Buffer, offset: int = 0, verified: bool = False) -> bytes:
        n, s = C.c_size_t(0), Stats()
        fn = (self.ctx._lib.azk_circle_prove_fft_coefficients_resident_verified if verified
              else self.ctx._lib.azk_circle_prove_fft_coefficients_resident)
        rc = fn(self.h, coeffs.h, offset, _ptr(self._out), self.proof_size, C.byref(n), C.byref(s))
        return self._finish(rc, n, s)

    def prove_fft_coefficients_m31_resident(self, coeffs: Buffer, offset: int = 0) -> bytes:
        """The same proof from N M31 coefficients (each lifted to (c, 0, 0, 0))."""
        n, s = C.c_size_t(0), Stats()
        rc = self.ctx._lib.azk_circle_prove_fft_coefficients_m31_resident(self.h, coeffs.h, offset, _ptr(self._out),
                                                                         self.proof_size, C.byref(n), C.byref(s))
        return self._finish(rc, n, s)

    def prove_fft_coefficients_m31(self, coeffs: np.ndarray) -> bytes:
        a = np.ascontiguousarray(coeffs, dtype=np.uint32)
        n, s = C.c_size_t(0), Stats()
        rc = self.ctx._lib.azk_circle_prove_fft_coefficients_m31(self.h, _ptr(a), a.size, _ptr(self._out),
                                                                self.proof_size, C.byref(n), C.byref(s))
        return self._finish(rc, n, s)

    def prove_evaluations_resident(self, evals: Buffer, offset: int = 0) -> bytes:
        n, s = C.c_size_t(0), Stats()
        rc = self.ctx._lib.azk_circle_prove_evaluations_resident(self.h, evals.h, offset, _ptr(self._out),
                                                                self.proof_size, C.byref(n), C.byref(s))
        return self._finish(rc, n, s)

    def prove_fft_coefficients(self, coeffs: np.ndarray, verified: bool = False) -> bytes:
        a = _host(np.ascontiguousarray(coeffs, dtype=np.uint32))
        n, s = C.c_size_t(0), Stats()
        fn = (self.ctx._lib.azk_circle_prove_fft_coefficients_verified if verified
              else self.ctx._lib.azk_circle_prove_fft_coefficients)
        rc = fn(self.h, _ptr(a), a.size, _ptr(self._out), self.proof_size, C.byref(n), C.byref(s))
        return self._finish(rc, n, s)

    def set_graph(self, enable: bool) -> None:
        _ck(self.ctx._lib.azk_circle_prover_set_graph(self.h, int(enable)))

    def set_timing(self, enable: bool) -> None:
        _ck(self.ctx._lib.azk_circle_prover_set_timing(self.h, int(enable)))

    def phase_times(self) -> dict:
        t = PhaseTimes()
        _ck(self.ctx._lib.azk_circle_prover_phase_times(self.h, C.byref(t)))
        return {k: getattr(t, k) for k, _ in PhaseTimes._fields_}

    def clear(self) -> None:
        _ck(self.ctx._lib.azk_circle_prover_clear(self.h))

    def set_exec_plan(self, **plan) -> None:
        """Tile shapes (azk_circle_exec_plan); unspecified knobs keep their current value."""
        x = ExecPlan()
        _ck(self.ctx._lib.azk_circle_prover_get_exec_plan(self.h, C.byref(x)))
        for k, v in plan.items():
            setattr(x, k, v)
        _ck(self.ctx._lib.azk_circle_prover_set_exec_plan(self.h, C.byref(x)))

    def exec_plan(self) -> dict:
        x = ExecPlan()
        _ck(self.ctx._lib.azk_circle_prover_get_exec_plan(self.h, C.byref(x)))
        return _plan_dict(x)

    def tune(self, coeffs: Buffer, offset: int = 0, database_path: str | None = None, force: bool = False,
             warmup_runs: int = 0, measured_runs: int = 0, max_candidates: int = 0, validation_runs: int = 0) -> dict:
        """Races tile shapes on these coefficients (azk_circle_prover_tune) and keeps the winner."""
        c = TuningConfig(warmup_runs, measured_runs, max_candidates, validation_runs,
                         database_path.encode() if database_path else None, int(force))
        r = TuningResult()
        _ck(self.ctx._lib.azk_circle_prover_tune(self.h, coeffs.h, offset, C.byref(c), C.byref(r)))
        out = {k: getattr(r, k) for k, _ in TuningResult._fields_ if k != "winner"}
        out["winner"] = _plan_dict(r.winner)
        return out

    def drift(self) -> dict:
        d = DriftStatus()
        _ck(self.ctx._lib.azk_circle_prover_drift(self.h, C.byref(d)))
        return {k: getattr(d, k) for k, _ in DriftStatus._fields_}


def prove_batch(provers: list, coeff_buffers: list, offsets=None) -> tuple[list[bytes], dict]:
    """Many independent commits in one captured graph (azk_circle_prove_batch)."""
    n = len(provers)
    lib = load()
    offs = (sz * n)(*(offsets or [0] * n))
    outs = [np.zeros(p.proof_size, np.uint8) for p in provers]
    ph = (vp * n)(*[p.h for p in provers])
    bh = (vp * n)(*[b.h for b in coeff_buffers])
    op = (vp * n)(*[o.ctypes.data for o in outs])
    caps = (sz * n)(*[p.proof_size for p in provers])
    lens = (sz * n)()
    s = Stats()
    _ck(lib.azk_circle_prove_batch(ph, bh, offs, n, op, caps, lens, C.byref(s)))
    return [outs[i][: lens[i]].tobytes() for i in range(n)], _stats_dict(s)


class CircleBatchPlan:
    """A fixed set of jobs captured once as one graph (azk_circle_batch_*)."""

    def __init__(self, provers: list, coeff_buffers: list, offsets=None, m31: bool = False):
        """m31: the buffers hold each job's N M31 coefficients (azk_circle_batch_create_m31)."""
        self.lib = load()
        self.provers, self.buffers = list(provers), list(coeff_buffers)
        n = len(self.provers)
        offs = (sz * n)(*(offsets or [0] * n))
        ph = (vp * n)(*[p.h for p in self.provers])
        bh = (vp * n)(*[b.h for b in self.buffers])
        h = vp()
        create = self.lib.azk_circle_batch_create_m31 if m31 else self.lib.azk_circle_batch_create
        _ck(create(ph, bh, offs, n, C.byref(h)))
        self.h = h
        _track(self)
        self._outs = [np.zeros(p.proof_size, np.uint8) for p in self.provers]
        self._op = (vp * n)(*[o.ctypes.data for o in self._outs])
        self._caps = (sz * n)(*[p.proof_size for p in self.provers])
        self._lens = (sz * n)()
        self.last_stats: dict = {}

    def prove(self) -> list[bytes]:
        s = Stats()
        _ck(self.lib.azk_circle_batch_prove(self.h, self._op, self._caps, self._lens, C.byref(s)))
        self.last_stats = _stats_dict(s)
        return [self._outs[i][: self._lens[i]].tobytes() for i in range(len(self.provers))]

    def tune(self, database_path: str | None = None, force: bool = False, warmup_runs: int = 0,
             measured_runs: int = 0, max_candidates: int = 0) -> dict:
        """Races one execution plan for the whole batch program (azk_circle_batch_tune)."""
        c = TuningConfig(warmup_runs, measured_runs, max_candidates, 0,
                         database_path.encode() if database_path else None, int(force))
        r = TuningResult()
        _ck(self.lib.azk_circle_batch_tune(self.h, C.byref(c), C.byref(r)))
        out = {k: getattr(r, k) for k, _ in TuningResult._fields_ if k != "winner"}
        out["winner"] = _plan_dict(r.winner)
        return out

    def close(self) -> None:
        if self.h:
            self.lib.azk_circle_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class M31MultilinearSumcheckPlan(_Plan):
    """GPU form of M31MultilinearSumcheckProofBuilderV1 (SumcheckOracle.swift:817-885)."""
    _destroy = "azk_sumcheck_plan_destroy"

    def __init__(self, ctx: Context, log_n: int):
        self.ctx, self.log_n = ctx, log_n
        h = vp()
        _ck(ctx._lib.azk_sumcheck_plan_create(ctx.h, log_n, C.byref(h)))
        self.h = h
        _track(self)
        self.proof_size = int(ctx._lib.azk_sumcheck_proof_size(log_n))
        self._out = np.zeros(self.proof_size, np.uint8)
        self.last_stats: dict = {}

    def prove(self, evaluations, claimed_sum: int | None = None, initial_digest: bytes | None = None) -> bytes:
        """M31MultilinearSumcheckProofCodecV1 bytes of the proof (`initial_digest`: the
        words' digestWords when the caller already holds it)."""
        a = np.ascontiguousarray(evaluations, dtype=np.uint32)
        cs = C.c_uint32(claimed_sum) if claimed_sum is not None else None
        d = np.frombuffer(initial_digest, np.uint8) if initial_digest is not None else None
        n, s = C.c_size_t(0), Stats()
        _ck(self.ctx._lib.azk_sumcheck_prove_with_digest(self.h, _ptr(a), a.size, _ptr(d) if d is not None else None,
                                                         C.byref(cs) if cs is not None else None,
                                                         _ptr(self._out), self.proof_size, C.byref(n), C.byref(s)))
        self.last_stats = _stats_dict(s)
        return self._out[: n.value].tobytes()

    def prove_resident(self, evals: Buffer, offset: int = 0, initial_digest: bytes | None = None,
                       claimed_sum: int | None = None) -> dict:
        """Round polynomials, challenges, final evaluation and statement fields."""
        r = SumcheckResult()
        polys = np.zeros(2 * self.log_n, np.uint32)
        ch = np.zeros(self.log_n, np.uint32)
        r.round_polys = polys.ctypes.data_as(C.POINTER(u32))
        r.challenges = ch.ctypes.data_as(C.POINTER(u32))
        dg = None
        if initial_digest is not None:
            if len(initial_digest) != 32:
                raise AZKError(6, "initial digest must be 32 bytes")
            dg = (C.c_uint8 * 32)(*initial_digest)
        cs = C.c_uint32(claimed_sum) if claimed_sum is not None else None
        s = Stats()
        _ck(self.ctx._lib.azk_sumcheck_prove_resident(self.h, evals.h, offset, dg, C.byref(cs) if cs is not None else None,
                                                      C.byref(r), C.byref(s)))
        self.last_stats = _stats_dict(s)
        return {"claimed_sum": r.claimed_sum, "final_evaluation": r.final_evaluation,
                "initial_digest": bytes(r.initial_digest), "round_polys": polys, "challenges": ch}


class MetalSumcheckChunkPlan(_Plan):
    """The M31 sum-check chunk (MetalSumcheckChunkPlan): per round the whole vector is absorbed,
    one challenge squeezed, and the vector folded v[2i] + r v[2i + 1]."""
    _destroy = "azk_sumcheck_chunk_plan_destroy"

    def __init__(self, ctx: Context, lane_count: int, rounds_per_superstep: int):
        self.ctx, self.lane_count, self.rounds = ctx, lane_count, rounds_per_superstep
        h = vp()
        _ck(ctx._lib.azk_sumcheck_chunk_plan_create(ctx.h, lane_count, rounds_per_superstep, C.byref(h)))
        self.h = h
        _track(self)
        self.last_stats = {}

    def _outputs(self):
        nc = self.ctx._lib.azk_sumcheck_chunk_coefficient_words(self.lane_count, self.rounds)
        return (np.zeros(self.lane_count >> self.rounds, np.uint32), np.zeros(nc, np.uint32),
                np.zeros(self.rounds, np.uint32))

    def execute(self, evaluations, verified: bool = False) -> dict:
        e = np.ascontiguousarray(evaluations, dtype=np.uint32).reshape(-1)
        fin, co, ch = self._outputs()
        s = Stats()
        fn = (self.ctx._lib.azk_sumcheck_chunk_execute_verified if verified
              else self.ctx._lib.azk_sumcheck_chunk_execute)
        _ck(fn(self.h, _ptr(e), e.size, _ptr(fin), _ptr(co), _ptr(ch), C.byref(s)))
        self.last_stats = _stats_dict(s)
        return {"final_vector": fin, "coefficients": co, "challenges": ch}

    def execute_uploaded_vector(self, buf: Buffer, offset: int = 0) -> dict:
        fin, co, ch = self._outputs()
        s = Stats()
        _ck(self.ctx._lib.azk_sumcheck_chunk_execute_resident(self.h, buf.h, offset, _ptr(fin), _ptr(co), _ptr(ch),
                                                              C.byref(s)))
        self.last_stats = _stats_dict(s)
        return {"final_vector": fin, "coefficients": co, "challenges": ch}

    def prove(self, evaluations) -> bytes:
        """M31SumcheckProofBuilderV1.prove, codec bytes ("AZKSCV1")."""
        e = np.ascontiguousarray(evaluations, dtype=np.uint32).reshape(-1)
        cap = self.ctx._lib.azk_sumcheck_chunk_proof_size(self.lane_count, self.rounds)
        out = np.zeros(max(cap, 1), np.uint8)
        n = C.c_size_t(0)
        s = Stats()
        _ck(self.ctx._lib.azk_sumcheck_chunk_prove(self.h, _ptr(e), e.size, _ptr(out), cap, C.byref(n), C.byref(s)))
        self.last_stats = _stats_dict(s)
        return out[: n.value].tobytes()

    def clear(self) -> None:
        _ck(self.ctx._lib.azk_sumcheck_chunk_plan_clear(self.h))


def sumcheck_chunk_verify(proof: bytes) -> bool:
    """M31SumcheckProofCodecV1.decode + M31SumcheckVerifierV1.verify (revealed folding trace)."""
    b = np.frombuffer(proof, np.uint8) if len(proof) else np.zeros(1, np.uint8)
    ok = C.c_int(0)
    _ck(load().azk_sumcheck_chunk_verify(b.ctypes.data, len(proof), C.byref(ok)))
    return bool(ok.value)


def circle_verify(proof: bytes, public_input_digest: bytes) -> bool:
    """CirclePCSFRIProofCodecV1.decode + CirclePCSFRIProofVerifierV1.verify on the host."""
    lib = load()
    if len(public_input_digest) != 32:
        raise AZKError(6, "public input digest must be 32 bytes")
    b = np.frombuffer(proof, np.uint8) if len(proof) else np.zeros(1, np.uint8)
    d = np.frombuffer(public_input_digest, np.uint8)
    ok = C.c_int(0)
    _ck(lib.azk_circle_verify(b.ctypes.data, len(proof), d.ctypes.data, C.byref(ok)))
    return bool(ok.value)


def circle_verify_with_coefficients(proof: bytes, public_input_digest: bytes, coeffs: np.ndarray) -> bool:
    """circle_verify, and every opened first-layer value equals the host Circle FFT of the
    line-basis coefficients (the prover's CPU gate)."""
    lib = load()
    if len(public_input_digest) != 32:
        raise AZKError(6, "public input digest must be 32 bytes")
    b = np.frombuffer(proof, np.uint8) if len(proof) else np.zeros(1, np.uint8)
    d = np.frombuffer(public_input_digest, np.uint8)
    c = np.ascontiguousarray(coeffs, dtype=np.uint32).reshape(-1)
    ok = C.c_int(0)
    _ck(lib.azk_circle_verify_with_coefficients(b.ctypes.data, len(proof), d.ctypes.data, _ptr(c), c.size, C.byref(ok)))
    return bool(ok.value)


def calibrate_valu(ctx: "Context") -> dict:
    """VALU issue rates and the register-resident Keccak-f rate of this chip (measurement aid)."""
    c = ValuCalibration()
    _ck(load().azk_calibrate_valu(ctx.h, C.byref(c)))
    return {"keccak_perms_per_s": c.keccak_perms_per_s, "keccak_clock_mhz": c.keccak_clock_mhz,
            "keccak_cycles_per_wave_perm": c.keccak_cycles_per_wave_perm,
            "cycles_per_wave64_instruction": {n: round(c.cycles_per_instr[k], 3)
                                              for k, n in enumerate(VALU_CALIBRATION_INSTRUCTIONS)},
            "clock_mhz": {n: round(c.clock_mhz[k], 1) for k, n in enumerate(VALU_CALIBRATION_INSTRUCTIONS)}}


def calibrate_latency(ctx: "Context") -> dict:
    """Dependent-permutation latency of the wave and two-lane Keccak forms, and the cost of one
    kernel of a captured graph (measurement aid)."""
    c = LatencyCalibration()
    _ck(load().azk_calibrate_latency(ctx.h, C.byref(c)))
    return {"wave_perm_us": c.wave_perm_us, "pair_perm_us": c.pair_perm_us, "graph_kernel_us": c.graph_kernel_us}


def grinding_digest_meets_target(digest: bytes, bits: int) -> bool:
    """CircleFRIGrindingV1.digestMeetsTarget."""
    d = np.frombuffer(digest, np.uint8) if len(digest) else np.zeros(1, np.uint8)
    return bool(load().azk_circle_grinding_digest_meets_target(d.ctypes.data, len(digest), bits))


class CirclePCSFRIPolynomialClaimV1:
    """A polynomial (monomial x / y QM31 coefficients) on the canonical 2^log_n domain with
    evaluation claims at storage indices; `points` / `values` default to what
    CirclePCSFRIPolynomialClaimV1.make computes."""

    def __init__(self, log_n: int, x_coeffs, y_coeffs, storage_indices, points=None, values=None):
        self.log_n = log_n
        self.x = np.ascontiguousarray(np.asarray(x_coeffs, dtype=np.uint32).reshape(-1, 4))
        self.y = np.ascontiguousarray(np.asarray(y_coeffs, dtype=np.uint32).reshape(-1, 4))
        self.indices = np.ascontiguousarray(np.asarray(storage_indices, dtype=np.uint64).reshape(-1))
        k = self.indices.size
        if points is None or values is None:
            self.points = np.zeros((k, 2), np.uint32)
            self.values = np.zeros((k, 4), np.uint32)
            _ck(load().azk_circle_polynomial_claim_make(C.byref(self._struct()), _ptr(self.points), _ptr(self.values)))
        if points is not None:
            self.points = np.ascontiguousarray(np.asarray(points, dtype=np.uint32).reshape(k, 2))
        if values is not None:
            self.values = np.ascontiguousarray(np.asarray(values, dtype=np.uint32).reshape(k, 4))

    def _struct(self) -> PolynomialClaim:
        def p(a):
            return a.ctypes.data if a.size else None
        pts = getattr(self, "points", None)
        vals = getattr(self, "values", None)
        return PolynomialClaim(self.log_n, len(self.x), len(self.y), self.indices.size, p(self.x), p(self.y),
                               p(self.indices), p(pts) if pts is not None else None,
                               p(vals) if vals is not None else None)

    def digest(self) -> bytes:
        """CirclePCSFRIPolynomialClaimDigestV1.digest: the public-input digest of the claim."""
        out = np.zeros(32, np.uint8)
        _ck(load().azk_circle_polynomial_claim_digest(C.byref(self._struct()), _ptr(out)))
        return out.tobytes()


def circle_verify_polynomial(proof: bytes, claim: CirclePCSFRIPolynomialClaimV1) -> bool:
    """CirclePCSFRIPolynomialVerifierV1.verify."""
    b = np.frombuffer(proof, np.uint8) if len(proof) else np.zeros(1, np.uint8)
    ok = C.c_int(0)
    _ck(load().azk_circle_verify_polynomial(b.ctypes.data, len(proof), C.byref(claim._struct()), C.byref(ok)))
    return bool(ok.value)


def circle_verify_contract(proof: bytes, claim: CirclePCSFRIPolynomialClaimV1, log_blowup: int = 4, queries: int = 36,
                           grinding_bits: int = 0, target_soundness_bits: int = 128) -> bool:
    """CirclePCSFRIContractVerifierV1.verify(encodedProof:statement:); the defaults are the
    conservative-128 profile (CircleProofFormat.swift:238-250)."""
    b = np.frombuffer(proof, np.uint8) if len(proof) else np.zeros(1, np.uint8)
    ps = ParameterSet(log_blowup, queries, grinding_bits, target_soundness_bits)
    ok = C.c_int(0)
    _ck(load().azk_circle_verify_contract(b.ctypes.data, len(proof), C.byref(ps), C.byref(claim._struct()), C.byref(ok)))
    return bool(ok.value)


def host_circle_fft(coeffs: np.ndarray, log_n: int) -> np.ndarray:
    """CircleCodewordOracle.evaluateWithCircleFFT on the host (the codeword plan's CPU gate)."""
    c = np.ascontiguousarray(coeffs, dtype=np.uint32).reshape(-1)
    if c.size != 4 << log_n:
        raise AZKError(6, "need 4 x 2^log_n coefficient words")
    out = np.zeros(4 << log_n, np.uint32)
    _ck(load().azk_host_circle_fft(_ptr(c), log_n, _ptr(out)))
    return out


def host_circle_prove(coeffs: np.ndarray, log_n: int, rounds: int, security=(2, 32, 1, 0),
                      public_input_digest: bytes = b"") -> bytes:
    """CircleFRIProofBuilderV1.prove on the host (the verified prover's CPU gate)."""
    c = np.ascontiguousarray(coeffs, dtype=np.uint32).reshape(-1)
    d = prover_desc(log_n, rounds, security, public_input_digest)
    cap = 64 << 20
    out = np.zeros(cap, np.uint8)
    n = C.c_size_t(0)
    _ck(load().azk_host_circle_prove(_ptr(c), C.byref(d), _ptr(out), cap, C.byref(n)))
    return out[: n.value].tobytes()


def host_fold_chain(evals: np.ndarray, log_n: int, challenges) -> np.ndarray:
    """CircleFRILayerOracleV1.fold on the host (the fold chain plan's CPU gate)."""
    e = np.ascontiguousarray(evals, dtype=np.uint32).reshape(-1)
    ch = np.ascontiguousarray(np.asarray(challenges, dtype=np.uint32).reshape(-1))
    rounds = ch.size // 4
    if e.size != 4 << log_n or not 1 <= rounds <= log_n:
        raise AZKError(6, "need 4 x 2^log_n evaluation words and 1..log_n challenges")
    out = np.zeros(4 << (log_n - rounds), np.uint32)
    _ck(load().azk_host_fold_chain(_ptr(e), log_n, _ptr(ch), rounds, _ptr(out)))
    return out


def witness_basis_check(x: np.ndarray, y: np.ndarray, log_n: int, basis: np.ndarray) -> bool:
    """Whether the basis and the monomial columns agree at pseudo-random QM31 points."""
    xa = np.ascontiguousarray(np.asarray(x, dtype=np.uint32).reshape(-1, 4))
    ya = np.ascontiguousarray(np.asarray(y, dtype=np.uint32).reshape(-1, 4))
    b = np.ascontiguousarray(basis, dtype=np.uint32).reshape(-1)
    if b.size != 4 << log_n:
        raise AZKError(6, "need 4 x 2^log_n basis words")
    ok = C.c_int(0)
    _ck(load().azk_witness_basis_check(_ptr(xa) if len(xa) else None, len(xa), _ptr(ya) if len(ya) else None, len(ya),
                                       log_n, _ptr(b), C.byref(ok)))
    return bool(ok.value)


def sumcheck_verify(proof: bytes) -> bool:
    """M31MultilinearSumcheckProofCodecV1.decode + M31MultilinearSumcheckVerifierV1.verify."""
    lib = load()
    b = np.frombuffer(proof, np.uint8)
    ok = C.c_int(0)
    _ck(lib.azk_sumcheck_verify(b.ctypes.data, b.size, C.byref(ok)))
    return bool(ok.value)


def now_ns() -> int:
    return int(load().azk_now_ns())


class PinnedArray:
    """Pinned host memory (azk_host_alloc) viewed as a numpy array; inputs placed here
    upload asynchronously."""

    def __init__(self, shape, dtype=np.uint32):
        self.lib = load()
        dt = np.dtype(dtype)
        n = int(np.prod(shape)) * dt.itemsize
        h = vp()
        _ck(self.lib.azk_host_alloc(n, C.byref(h)))
        self.h = h
        _track(self)
        buf = (C.c_uint8 * max(n, 1)).from_address(h.value)
        self.array = np.frombuffer(buf, np.uint8, count=n).view(dt).reshape(shape)

    def close(self) -> None:
        if self.h:
            self.array = None
            self.lib.azk_host_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def prover_desc(log_n: int, rounds: int, security=(2, 32, 1, 0), public_input_digest: bytes = b"") -> ProverDesc:
    d = ProverDesc()
    d.log_n, d.rounds = log_n, rounds
    d.log_blowup, d.query_count, d.folding_step, d.grinding_bits = security
    dg = public_input_digest or bytes(32)
    if len(dg) != 32:
        raise AZKError(6, "public input digest must be 32 bytes")
    d.public_input_digest[:] = list(dg)
    return d


class Server:
    """Serving form (azk_server_*): commits and sum-checks submitted at any time."""

    def __init__(self, device: int = 0, max_batch: int = 64, aux_lanes: int = 4, hbm_budget: int = 0):
        self.lib = load()
        d = ServerDesc(device, max_batch, aux_lanes, 0, hbm_budget)
        h = vp()
        _ck(self.lib.azk_server_create(C.byref(d), C.byref(h)))
        self.h = h
        _track(self)
        self._keep: dict = {}

    def submit_commit(self, desc: ProverDesc, coeffs: np.ndarray) -> int:
        a = np.ascontiguousarray(coeffs, dtype=np.uint32)
        t = C.c_uint64(0)
        _ck(self.lib.azk_server_submit_commit(self.h, C.byref(desc), a.ctypes.data, a.nbytes, C.byref(t)))
        self._keep[t.value] = a  # read in place until the job is done
        return t.value

    def submit_commit_m31(self, desc: ProverDesc, coeffs: np.ndarray) -> int:
        """A commit from M31 line-basis coefficients (the proof of their QM31 lift)."""
        a = np.ascontiguousarray(coeffs, dtype=np.uint32)
        t = C.c_uint64(0)
        _ck(self.lib.azk_server_submit_commit_m31(self.h, C.byref(desc), a.ctypes.data, a.size, C.byref(t)))
        self._keep[t.value] = a  # read in place until the job is done
        return t.value

    def submit_sumcheck(self, evals: np.ndarray, initial_digest: bytes | None = None) -> int:
        a = np.ascontiguousarray(evals, dtype=np.uint32)
        t = C.c_uint64(0)
        if initial_digest is None:
            _ck(self.lib.azk_server_submit_sumcheck(self.h, a.ctypes.data, a.size, C.byref(t)))
        else:
            d = np.frombuffer(initial_digest, np.uint8)
            _ck(self.lib.azk_server_submit_sumcheck_digest(self.h, a.ctypes.data, a.size, d.ctypes.data, C.byref(t)))
        self._keep[t.value] = a
        return t.value

    def submit_sumcheck_chunk(self, evals: np.ndarray, rounds: int) -> int:
        a = np.ascontiguousarray(evals, dtype=np.uint32)
        t = C.c_uint64(0)
        _ck(self.lib.azk_server_submit_sumcheck_chunk(self.h, a.ctypes.data, a.size, rounds, C.byref(t)))
        self._keep[t.value] = a
        return t.value

    def submit_fold(self, log_n: int, rounds: int, evals: np.ndarray, challenges) -> int:
        a = np.ascontiguousarray(evals, dtype=np.uint32)
        ch = np.ascontiguousarray(np.asarray(challenges, dtype=np.uint32).reshape(-1))
        t = C.c_uint64(0)
        _ck(self.lib.azk_server_submit_fold(self.h, log_n, rounds, a.ctypes.data, a.nbytes, ch.ctypes.data, C.byref(t)))
        self._keep[t.value] = a
        return t.value

    def wait(self, ticket: int) -> dict:
        i = JobInfo()
        rc = self.lib.azk_server_wait(self.h, ticket, C.byref(i))
        info = {k: getattr(i, k) for k, _ in JobInfo._fields_}
        info["error"] = self.lib.azk_last_error().decode() if rc else ""
        return info

    def take(self, ticket: int) -> bytes:
        i = JobInfo()
        rc = self.lib.azk_server_wait(self.h, ticket, C.byref(i))
        if rc:
            msg = self.lib.azk_last_error().decode()
            self.lib.azk_server_take(self.h, ticket, None, 0, None)
            self._keep.pop(ticket, None)
            raise AZKError(rc, msg)
        out = np.zeros(max(i.result_bytes, 1), np.uint8)
        n = C.c_size_t(0)
        _ck(self.lib.azk_server_take(self.h, ticket, out.ctypes.data, out.size, C.byref(n)))
        self._keep.pop(ticket, None)
        return out[: n.value].tobytes()

    def close(self) -> None:
        if self.h:
            self.lib.azk_server_destroy(self.h)
            self.h = None
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
