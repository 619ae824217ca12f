"""Seeded synthetic inputs (SplitMix64), vectorised with numpy.

The generator is the one the reference's planner uses for its randomized
validation (DeterministicByteGenerator, Sources/AppleZKProver/Planner/
MetalProofPlanner.swift:758-789): state += 0x9e3779b97f4a7c15, then the
SplitMix64 finaliser; bytes are taken little-endian from each 64-bit output;
seed 0 maps to 0x243f6a8885a308d3.  M31 words are the same stream read as LE
u32 words, masked to 31 bits, with the value p rejected.
"""
from __future__ import annotations

import numpy as np

P = 2147483647
_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, count: int) -> np.ndarray:
    s0 = np.uint64(seed if seed else 0x243F6A8885A308D3)
    k = np.arange(1, count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = s0 + k * _GAMMA
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def random_bytes(seed: int, n: int) -> np.ndarray:
    words = splitmix64(seed, (n + 7) // 8)
    return words.view(np.uint8)[:n].copy()


def random_m31(seed: int, n: int) -> np.ndarray:
    out = np.empty(0, np.uint32)
    draw = n + 16
    while out.size < n:
        w = (splitmix64(seed, (draw + 1) // 2).view(np.uint32) & np.uint32(0x7FFFFFFF))
        out = w[w != P]
        draw = draw * 2
    return out[:n].copy()


def circle_fft_coefficients(seed: int, log_n: int, half_len: int | None = None) -> np.ndarray:
    """Interleaved line-basis coefficients of P(x) + y Q(x): P, Q of `half_len` M31 values
    each (default N/4, blowup 2 of the coefficient budget), lifted to QM31 (a, 0, 0, 0),
    coeff[2i] = P_i, coeff[2i+1] = Q_i, zero above (CircleCodeword.swift:100-127)."""
    n = 1 << log_n
    h = n // 4 if half_len is None else half_len
    pq = random_m31(seed, 2 * h)
    c = np.zeros((n, 4), np.uint32)
    c[0:2 * h:2, 0] = pq[:h]
    c[1:2 * h:2, 0] = pq[h:]
    return c
