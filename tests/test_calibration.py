"""Measurement aids behind the bench line's rooflines: VALU issue costs, the serial chain's
latency figures, and the upper-tree counts a timed prove reports."""
import numpy as np
import pytest

from apple_amd import azk, inputs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = azk.Context(0)
    yield c
    c.close()


def test_valu_calibration_is_plausible(ctx):
    v = azk.calibrate_valu(ctx)
    cyc = v["cycles_per_wave64_instruction"]
    assert 1.5 < cyc["v_xor_b32"] < 3.5 and 3.0 < cyc["v_alignbit_b32"] < 6.0
    assert v["keccak_perms_per_s"] > 1e9


def test_latency_calibration_orders_the_forms(ctx):
    lat = azk.calibrate_latency(ctx)
    print("latency calibration", lat, flush=True)
    # one state per wave is the latency form; two lanes per state sits between it and one lane
    assert 0.5 < lat["wave_perm_us"] < lat["pair_perm_us"] < 50.0
    assert 0.1 < lat["graph_kernel_us"] < 100.0


def test_phase_times_count_upper_tree_launches(ctx):
    log_n, rounds = 14, 5
    c = inputs.circle_fft_coefficients(seed=9, log_n=log_n)
    p = azk.CircleCodewordPCSFRIProverV1(ctx, log_n, rounds, (2, 8, 1, 0), bytes(32))
    buf = ctx.upload(c)
    want = p.prove_fft_coefficients_resident(buf)
    p.set_timing(True)
    assert p.prove_fft_coefficients_resident(buf) == want  # eager, phase-timed: the same proof
    ph = p.phase_times()
    p.set_timing(False)
    assert ph["tree_launches"] > 0 and ph["tree_levels"] >= ph["tree_launches"]
    assert ph["tree_ns"] > 0 and ph["total_ns"] >= ph["tree_ns"]
