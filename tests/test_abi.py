"""CPU checks of the drop-in boundary: the library loads, exports every symbol that
include/azk.h declares, and the host-side input generator matches the oracle's."""
import ctypes
import os

import numpy as np
import pytest

from apple_amd import azk, inputs
import pyoracle as O


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(azk.LIB_PATH)
    declared = azk.declared_symbols()
    assert len(declared) > 40
    missing = [s for s in declared if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_declares_every_header_symbol():
    lib = azk.load()
    for s in azk.declared_symbols():
        assert getattr(lib, s).restype is not None or s.endswith(("destroy", "free")), s


def test_no_device_reports_status_not_crash():
    try:
        c = azk.Context(0)
    except azk.AZKError as e:  # no GPU here: a status, not a crash
        assert e.status == "noDevice"
        return
    c.close()
    pytest.skip("a GPU is present")


def test_splitmix_matches_oracle():
    for seed in [0, 1, 12345, 2**63 + 5]:
        assert np.array_equal(inputs.random_bytes(seed, 1001), O.splitmix_bytes(seed, 1001))
        assert np.array_equal(inputs.random_m31(seed, 777), O.splitmix_m31(seed, 777))
    assert inputs.random_m31(3, 100000).max() < inputs.P


def test_coefficient_layout():
    c = inputs.circle_fft_coefficients(5, 6)
    assert c.shape == (64, 4)
    assert not c[32:].any() and not c[:, 1:].any()
    pq = inputs.random_m31(5, 32)
    assert np.array_equal(c[0:32:2, 0], pq[:16]) and np.array_equal(c[1:32:2, 0], pq[16:])
