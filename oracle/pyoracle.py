"""ctypes view of the CPU oracle (oracle/liboracle.so) -- test infrastructure only.

Used by tests/, the smoke check and the bench's CPU-baseline leg as the checker.
The product library never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
P = 2147483647

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


u8p = C.POINTER(C.c_uint8)
u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)


class Security(C.Structure):
    _fields_ = [("log_blowup", C.c_uint32), ("query_count", C.c_uint32),
                ("folding_step", C.c_uint32), ("grinding_bits", C.c_uint32)]


class ProveInfo(C.Structure):
    _fields_ = [("commitments", C.c_uint8 * (32 * 32)), ("challenges", C.c_uint32 * (32 * 4)),
                ("query_pairs", C.c_uint32 * 4096), ("n_queries", C.c_uint32),
                ("grinding_nonce", C.c_uint64), ("pre_grinding", C.c_uint64 * 25)]


class Transcript(C.Structure):
    _fields_ = [("s", C.c_uint64 * 25)]


def _declare(l: C.CDLL) -> None:
    sz = C.c_size_t
    sigs = {
        "ora_keccak_f1600": (None, [u64p]),
        "ora_sponge": (C.c_int, [u8p, sz, C.c_int, C.c_uint8, u8p, sz]),
        "ora_sha3_256": (None, [u8p, sz, u8p]),
        "ora_keccak_256": (None, [u8p, sz, u8p]),
        "ora_keccakf_batch": (C.c_int, [u8p, sz, u8p, sz, sz]),
        "ora_hash_oneblock_batch": (C.c_int, [u8p, sz, sz, u8p, sz, sz, C.c_uint8]),
        "ora_transcript_init": (None, [C.c_void_p]),
        "ora_transcript_absorb": (None, [C.c_void_p, u8p, sz, C.c_uint8]),
        "ora_transcript_squeeze_u32": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, u32p]),
        "ora_transcript_squeeze_bytes": (None, [C.c_void_p, sz, u8p]),
        "ora_m31_mul": (C.c_uint32, [C.c_uint32, C.c_uint32]),
        "ora_m31_inv": (C.c_uint32, [C.c_uint32]),
        "ora_qm31_mul": (None, [u32p, u32p, u32p]),
        "ora_qm31_square": (None, [u32p, u32p]),
        "ora_qm31_inv": (C.c_int, [u32p, u32p]),
        "ora_circle_point_from_index": (None, [C.c_uint32, u32p]),
        "ora_circle_point_index": (C.c_int, [C.c_uint32, C.c_uint32, u32p]),
        "ora_circle_domain_points": (C.c_int, [C.c_uint32, u32p]),
        "ora_circle_fft_twiddles": (C.c_int, [C.c_uint32, u32p]),
        "ora_circle_fold_twiddles": (C.c_int, [C.c_uint32, C.c_uint32, u32p]),
        "ora_circle_domain_descriptor": (C.c_int, [C.c_uint32, u8p]),
        "ora_circle_fft": (C.c_int, [u32p, C.c_uint32, u32p]),
        "ora_circle_basis_from_monomial": (C.c_int, [u32p, C.c_uint32, u32p, C.c_uint32, C.c_uint32, u32p]),
        "ora_circle_evaluate_direct": (C.c_int, [u32p, C.c_uint32, u32p, C.c_uint32, C.c_uint32, u32p]),
        "ora_circle_evaluate_at": (C.c_int, [u32p, C.c_uint32, u32p, C.c_uint32, C.c_uint32, u32p, sz, u32p]),
        "ora_fold_layer": (C.c_int, [u32p, sz, u32p, u32p, u32p]),
        "ora_circle_fold_chain": (C.c_int, [u32p, C.c_uint32, u32p, C.c_uint32, u32p]),
        "ora_merkle_root": (C.c_int, [u8p, sz, sz, sz, u8p]),
        "ora_merkle_tree": (C.c_int, [u8p, sz, sz, sz, u8p]),
        "ora_merkle_opening": (C.c_int, [u8p, sz, sz, sz, sz, u8p, u8p]),
        "ora_merkle_verify": (C.c_int, [u8p, sz, sz, u8p, C.c_uint32, u8p]),
        "ora_circle_fri_prove": (C.c_int, [u32p, C.c_uint32, C.POINTER(Security), u8p, C.c_uint32, u32p,
                                           C.c_uint32, u8p, sz, C.POINTER(sz), C.POINTER(ProveInfo)]),
        "ora_circle_fri_verify": (C.c_int, [u8p, sz, u8p]),
        "ora_digest_meets_target": (C.c_int, [u8p, sz, C.c_uint32]),
        "ora_circle_grinding_digest": (None, [u64p, C.c_uint32, C.c_uint64, u8p]),
        "ora_circle_transcript_challenges": (C.c_int, [C.c_uint32, C.POINTER(Security), u8p, u8p, C.c_uint32, u32p]),
        "ora_sumcheck_chunk": (C.c_int, [u32p, sz, C.c_uint32, u32p, u32p, u32p]),
        "ora_sumcheck_multilinear": (C.c_int, [u32p, sz, u32p, u32p, u32p, u32p, u8p]),
        "ora_sumcheck_words_digest": (None, [u32p, sz, u8p]),
        "ora_sumcheck_multilinear_proof": (sz, [u32p, sz, u8p, sz]),
        "ora_splitmix_bytes": (None, [C.c_uint64, sz, u8p]),
        "ora_splitmix_m31": (None, [C.c_uint64, sz, u32p]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(l, name)
        fn.restype = res
        fn.argtypes = args


def _p8(a: np.ndarray):
    return a.ctypes.data_as(u8p)


def _p32(a: np.ndarray):
    return a.ctypes.data_as(u32p)


def _bytes(b) -> np.ndarray:
    return np.frombuffer(bytes(b), dtype=np.uint8).copy() if len(b) else np.zeros(1, np.uint8)


def _check(rc: int, what: str) -> None:
    if rc < 0:
        raise ValueError(f"oracle rejected {what} (rc={rc})")


# ---- hashing -------------------------------------------------------------
def keccak_f1600(lanes) -> list[int]:
    st = (C.c_uint64 * 25)(*lanes)
    lib().ora_keccak_f1600(st)
    return list(st)


def sponge(data: bytes, rate: int, suffix: int, outlen: int) -> bytes:
    a = _bytes(data)
    out = np.zeros(max(outlen, 1), np.uint8)
    _check(lib().ora_sponge(_p8(a), len(data), rate, suffix, _p8(out), outlen), "sponge")
    return out[:outlen].tobytes()


def sha3_256(data: bytes) -> bytes:
    return sponge(data, 136, 0x06, 32)


def keccak_256(data: bytes) -> bytes:
    return sponge(data, 136, 0x01, 32)


def keccakf_batch(states: np.ndarray, in_stride: int = 200, out_stride: int = 200, count: int | None = None) -> np.ndarray:
    raw = np.ascontiguousarray(states).view(np.uint8).reshape(-1)
    n = count if count is not None else raw.size // in_stride
    out = np.zeros(n * out_stride, np.uint8)
    _check(lib().ora_keccakf_batch(_p8(raw), in_stride, _p8(out), out_stride, n), "keccakf batch")
    return out


def hash_oneblock_batch(msgs: np.ndarray, stride: int, length: int, count: int, suffix: int, out_stride: int = 32) -> np.ndarray:
    raw = np.ascontiguousarray(msgs).view(np.uint8).reshape(-1)
    if raw.size == 0:
        raw = np.zeros(1, np.uint8)
    out = np.zeros(count * out_stride, np.uint8)
    _check(lib().ora_hash_oneblock_batch(_p8(raw), stride, length, _p8(out), out_stride, count, suffix), "hash batch")
    return out


# ---- transcript ----------------------------------------------------------
class TranscriptState:
    def __init__(self) -> None:
        self.t = Transcript()
        lib().ora_transcript_init(C.byref(self.t))

    def absorb(self, data: bytes, suffix: int = 0x06) -> None:
        a = _bytes(data)
        lib().ora_transcript_absorb(C.byref(self.t), _p8(a), len(data), suffix)

    def squeeze_u32(self, count: int, modulus: int) -> list[int]:
        out = np.zeros(max(count, 1), np.uint32)
        _check(lib().ora_transcript_squeeze_u32(C.byref(self.t), count, modulus, _p32(out)), "squeeze")
        return [int(v) for v in out[:count]]

    def squeeze_bytes(self, count: int) -> bytes:
        out = np.zeros(max(count, 1), np.uint8)
        lib().ora_transcript_squeeze_bytes(C.byref(self.t), count, _p8(out))
        return out[:count].tobytes()


# ---- fields --------------------------------------------------------------
def qm31_mul(a, b) -> tuple:
    x = np.array(a, np.uint32); y = np.array(b, np.uint32); o = np.zeros(4, np.uint32)
    lib().ora_qm31_mul(_p32(x), _p32(y), _p32(o))
    return tuple(int(v) for v in o)


def qm31_square(a) -> tuple:
    x = np.array(a, np.uint32); o = np.zeros(4, np.uint32)
    lib().ora_qm31_square(_p32(x), _p32(o))
    return tuple(int(v) for v in o)


def qm31_inv(a) -> tuple:
    x = np.array(a, np.uint32); o = np.zeros(4, np.uint32)
    _check(lib().ora_qm31_inv(_p32(x), _p32(o)), "qm31 inverse")
    return tuple(int(v) for v in o)


def m31_mul(a: int, b: int) -> int:
    return int(lib().ora_m31_mul(a, b))


def m31_inv(a: int) -> int:
    return int(lib().ora_m31_inv(a))


# ---- circle --------------------------------------------------------------
def point_from_index(i: int) -> tuple:
    o = np.zeros(2, np.uint32)
    lib().ora_circle_point_from_index(i, _p32(o))
    return int(o[0]), int(o[1])


def point_index(log_n: int, natural: int) -> int:
    o = np.zeros(1, np.uint32)
    _check(lib().ora_circle_point_index(log_n, natural, _p32(o)), "point index")
    return int(o[0])


def domain_points(log_n: int) -> np.ndarray:
    o = np.zeros(2 << log_n, np.uint32)
    _check(lib().ora_circle_domain_points(log_n, _p32(o)), "domain")
    return o


def fft_twiddles(log_n: int) -> np.ndarray:
    o = np.zeros(max((1 << log_n) - 1, 1), np.uint32)
    _check(lib().ora_circle_fft_twiddles(log_n, _p32(o)), "fft twiddles")
    return o[: (1 << log_n) - 1]


def fold_twiddles(log_n: int, rounds: int) -> np.ndarray:
    n = 1 << log_n
    total = n - (n >> rounds)
    o = np.zeros(n, np.uint32)
    _check(lib().ora_circle_fold_twiddles(log_n, rounds, _p32(o)), "fold twiddles")
    return o[:total]


def domain_descriptor(log_n: int) -> bytes:
    o = np.zeros(44, np.uint8)
    _check(lib().ora_circle_domain_descriptor(log_n, _p8(o)), "descriptor")
    return o.tobytes()


def circle_fft(coeffs: np.ndarray, log_n: int) -> np.ndarray:
    c = np.ascontiguousarray(coeffs, dtype=np.uint32).reshape(-1)
    o = np.zeros(4 << log_n, np.uint32)
    _check(lib().ora_circle_fft(_p32(c), log_n, _p32(o)), "fft")
    return o


def _qm_arr(vals) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(vals, dtype=np.uint32).reshape(-1))
    return a if a.size else np.zeros(4, np.uint32)


def basis_from_monomial(x, y, log_n: int) -> np.ndarray:
    xa, ya = _qm_arr(x), _qm_arr(y)
    nx, ny = len(x), len(y)
    o = np.zeros(4 << log_n, np.uint32)
    _check(lib().ora_circle_basis_from_monomial(_p32(xa), nx, _p32(ya), ny, log_n, _p32(o)), "monomial basis")
    return o


def evaluate_direct(x, y, log_n: int) -> np.ndarray:
    xa, ya = _qm_arr(x), _qm_arr(y)
    o = np.zeros(4 << log_n, np.uint32)
    _check(lib().ora_circle_evaluate_direct(_p32(xa), len(x), _p32(ya), len(y), log_n, _p32(o)), "direct eval")
    return o


def evaluate_at(x, y, log_n: int, idx) -> np.ndarray:
    """P(x) + y Q(x) at the given storage indices (Horner per point)."""
    xa, ya = _qm_arr(x), _qm_arr(y)
    ia = np.ascontiguousarray(idx, dtype=np.uint32)
    o = np.zeros(4 * len(ia), np.uint32)
    _check(lib().ora_circle_evaluate_at(_p32(xa), len(x), _p32(ya), len(y), log_n, _p32(ia), len(ia), _p32(o)),
           "evaluate at")
    return o.reshape(-1, 4)


def fold_layer(vals: np.ndarray, tw: np.ndarray, ch) -> np.ndarray:
    v = np.ascontiguousarray(vals, dtype=np.uint32).reshape(-1)
    t = np.ascontiguousarray(tw, dtype=np.uint32)
    c = np.array(ch, np.uint32)
    n = v.size // 4
    o = np.zeros(4 * (n // 2), np.uint32)
    _check(lib().ora_fold_layer(_p32(v), n, _p32(t), _p32(c), _p32(o)), "fold")
    return o


def fold_chain(evals: np.ndarray, log_n: int, challenges) -> np.ndarray:
    v = np.ascontiguousarray(evals, dtype=np.uint32).reshape(-1)
    c = np.ascontiguousarray(np.asarray(challenges, dtype=np.uint32).reshape(-1))
    r = c.size // 4
    o = np.zeros(4 * ((1 << log_n) >> r), np.uint32)
    _check(lib().ora_circle_fold_chain(_p32(v), log_n, _p32(c), r, _p32(o)), "fold chain")
    return o


# ---- Merkle --------------------------------------------------------------
def merkle_root(raw: bytes | np.ndarray, count: int, stride: int, length: int) -> bytes:
    a = _bytes(raw) if isinstance(raw, (bytes, bytearray)) else np.ascontiguousarray(raw).view(np.uint8).reshape(-1)
    if a.size == 0:
        a = np.zeros(1, np.uint8)
    o = np.zeros(32, np.uint8)
    _check(lib().ora_merkle_root(_p8(a), count, stride, length, _p8(o)), "merkle root")
    return o.tobytes()


def merkle_tree(raw, count: int, stride: int, length: int) -> np.ndarray:
    a = _bytes(raw) if isinstance(raw, (bytes, bytearray)) else np.ascontiguousarray(raw).view(np.uint8).reshape(-1)
    if a.size == 0:
        a = np.zeros(1, np.uint8)
    o = np.zeros((2 * count - 1) * 32, np.uint8)
    _check(lib().ora_merkle_tree(_p8(a), count, stride, length, _p8(o)), "merkle tree")
    return o


def merkle_opening(raw, count: int, stride: int, length: int, index: int) -> tuple[bytes, bytes]:
    a = _bytes(raw) if isinstance(raw, (bytes, bytearray)) else np.ascontiguousarray(raw).view(np.uint8).reshape(-1)
    if a.size == 0:
        a = np.zeros(1, np.uint8)
    nsib = count.bit_length() - 1
    sib = np.zeros(max(32 * nsib, 1), np.uint8)
    root = np.zeros(32, np.uint8)
    _check(lib().ora_merkle_opening(_p8(a), count, stride, length, index, _p8(sib), _p8(root)), "opening")
    return sib[: 32 * nsib].tobytes(), root.tobytes()


def merkle_verify(leaf: bytes, index: int, siblings: bytes, root: bytes) -> bool:
    lf, sb, rt = _bytes(leaf), _bytes(siblings), _bytes(root)
    rc = lib().ora_merkle_verify(_p8(lf), len(leaf), index, _p8(sb), len(siblings) // 32, _p8(rt))
    _check(rc, "verify")
    return rc == 1


# ---- Circle PCS / FRI ----------------------------------------------------
def fri_prove(evals: np.ndarray, log_n: int, security, digest: bytes, rounds: int, claimed=()):
    v = np.ascontiguousarray(evals, dtype=np.uint32).reshape(-1)
    sec = Security(*security)
    d = _bytes(digest)
    cl = np.array(list(claimed) or [0], np.uint32)
    cap = 64 << 20
    out = np.zeros(cap, np.uint8)
    n = C.c_size_t(0)
    info = ProveInfo()
    _check(lib().ora_circle_fri_prove(_p32(v), log_n, C.byref(sec), _p8(d), rounds, _p32(cl), len(claimed),
                                      _p8(out), cap, C.byref(n), C.byref(info)), "fri prove")
    meta = {
        "commitments": [bytes(info.commitments[32 * r:32 * r + 32]) for r in range(rounds)],
        "challenges": [tuple(info.challenges[4 * r:4 * r + 4]) for r in range(rounds)],
        "query_pairs": list(info.query_pairs[: info.n_queries]),
        "grinding_nonce": int(info.grinding_nonce),
        "pre_grinding": np.array(info.pre_grinding[:], np.uint64),
    }
    return out[: n.value].tobytes(), meta


def fri_verify(proof: bytes, digest: bytes) -> int:
    p, d = _bytes(proof), _bytes(digest)
    return int(lib().ora_circle_fri_verify(_p8(p), len(proof), _p8(d)))


def transcript_challenges(log_n: int, security, digest: bytes, commitments: list[bytes]) -> list[tuple]:
    sec = Security(*security)
    d = _bytes(digest)
    cm = _bytes(b"".join(commitments))
    o = np.zeros(4 * len(commitments), np.uint32)
    lib().ora_circle_transcript_challenges(log_n, C.byref(sec), _p8(d), _p8(cm), len(commitments), _p32(o))
    return [tuple(int(x) for x in o[4 * r:4 * r + 4]) for r in range(len(commitments))]


# ---- sum-check -----------------------------------------------------------
def digest_meets_target(digest: bytes, bits: int) -> bool:
    d = _bytes(digest) if digest else np.zeros(1, np.uint8)
    return bool(lib().ora_digest_meets_target(_p8(d), len(digest), bits))


def grinding_digest(pre_grinding: np.ndarray, bits: int, nonce: int) -> bytes:
    st = np.ascontiguousarray(pre_grinding, dtype=np.uint64)
    out = np.zeros(32, np.uint8)
    lib().ora_circle_grinding_digest(st.ctypes.data_as(u64p), bits, nonce, _p8(out))
    return out.tobytes()


def sumcheck_chunk(evals, rounds: int):
    e = np.ascontiguousarray(evals, dtype=np.uint32)
    n = e.size
    fin = np.zeros(n, np.uint32)
    coeffs = np.zeros(2 * n, np.uint32)
    ch = np.zeros(rounds, np.uint32)
    _check(lib().ora_sumcheck_chunk(_p32(e), n, rounds, _p32(fin), _p32(coeffs), _p32(ch)), "sumcheck chunk")
    ncoef = sum(n >> r for r in range(rounds))
    return fin[: n >> rounds].tolist(), coeffs[:ncoef].tolist(), ch.tolist()


def sumcheck_multilinear(evals):
    e = np.ascontiguousarray(evals, dtype=np.uint32)
    n = e.size
    rounds = n.bit_length() - 1
    rp = np.zeros(2 * rounds, np.uint32)
    ch = np.zeros(rounds, np.uint32)
    fin = np.zeros(1, np.uint32)
    cs = np.zeros(1, np.uint32)
    dg = np.zeros(32, np.uint8)
    _check(lib().ora_sumcheck_multilinear(_p32(e), n, _p32(rp), _p32(ch), _p32(fin), _p32(cs), _p8(dg)), "multilinear")
    return {"round_polys": rp.reshape(-1, 2).tolist(), "challenges": ch.tolist(), "final": int(fin[0]),
            "claimed_sum": int(cs[0]), "digest": dg.tobytes()}


def sumcheck_multilinear_proof(evals) -> bytes:
    e = np.ascontiguousarray(evals, dtype=np.uint32)
    n = e.size
    rounds = n.bit_length() - 1
    cap = 80 + 4 * n + 12 * rounds + 64
    out = np.zeros(cap, np.uint8)
    k = lib().ora_sumcheck_multilinear_proof(_p32(e), n, _p8(out), cap)
    if k == 0:
        raise ValueError("oracle rejected the sum-check layout")
    return out[:k].tobytes()


# ---- seeded inputs -------------------------------------------------------
def splitmix_bytes(seed: int, n: int) -> np.ndarray:
    o = np.zeros(max(n, 1), np.uint8)
    lib().ora_splitmix_bytes(seed, n, _p8(o))
    return o[:n]


def splitmix_m31(seed: int, n: int) -> np.ndarray:
    o = np.zeros(max(n, 1), np.uint32)
    lib().ora_splitmix_m31(seed, n, _p32(o))
    return o[:n]
