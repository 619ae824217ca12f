"""MI355X-native Circle-PCS commit engine (gfx950 HIP kernels behind a C ABI).

``apple_amd.azk`` binds ``libazk.so``; ``apple_amd.inputs`` makes the seeded
synthetic inputs used by the bench and the tests.
"""
from . import azk, inputs  # noqa: F401

__all__ = ["azk", "inputs"]
