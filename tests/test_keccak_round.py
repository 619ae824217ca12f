"""The pinned-register Keccak-f round (apple_amd/csrc/keccak_round_asm.hpp, written by
apple_amd/csrc/gen_keccak_round.py): its instruction program, interpreted on the host,
is Keccak-f[1600] (the oracle's, itself pinned to the FIPS 202 / SHA3 vectors in
test_oracle.py), its register layout keeps every three-source op off a single parity, and
the committed header is the generator's current output.
"""
import os
import random
import sys

import pytest

import pyoracle as O

CSRC = os.path.join(os.path.dirname(__file__), "..", "apple_amd", "csrc")
sys.path.insert(0, CSRC)
import gen_keccak_round as G  # noqa: E402


def round_constants():
    """FIPS 202 section 3.2.5: rc(t) from the LFSR x^8 + x^6 + x^5 + x^4 + 1; RC[r] sets
    bit 2^j - 1 to rc(j + 7r)."""
    state, bits = 1, []
    for _ in range(7 * 24):
        bits.append(state & 1)
        state <<= 1
        if state & 0x100:
            state ^= 0x171
    return [sum(bits[7 * r + j] << ((1 << j) - 1) for j in range(7)) for r in range(24)]


RC = round_constants()


def permute(lanes):
    prog = G.program()
    for r in range(24):
        lanes = G.emulate_round(lanes, RC[r] & 0xFFFFFFFF, RC[r] >> 32, prog)
    return lanes


def test_round_constants_are_the_fips_table():
    assert RC[0] == 1 and RC[1] == 0x8082 and RC[23] == 0x8000000080008008
    assert RC[6] == RC[20] == 0x8000000080008081  # the table's one repeated constant


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_program_is_keccak_f(seed):
    rng = random.Random(seed)
    lanes = [rng.getrandbits(64) for _ in range(25)]
    assert permute(lanes) == O.keccak_f1600(lanes)


def test_program_on_special_states():
    for lanes in ([0] * 25, [(1 << 64) - 1] * 25, [1 << i for i in range(25)]):
        assert permute(list(lanes)) == O.keccak_f1600(list(lanes))


def test_layout_and_op_counts():
    prog = G.program()
    G.check_layout(prog)  # raises on a three-source op of one parity or an overlapping role
    ops = [op for op, *_ in prog]
    assert len(prog) == 180
    assert ops.count("bitop3") == 120 and ops.count("alignbit") == 58 and ops.count("xor") == 2
    for op, d, srcs, imm in prog:
        regs = [d] + [r for r in srcs if isinstance(r, int)]
        assert all(8 <= r <= 107 for r in regs)


def test_a_planted_layout_mistake_is_caught():
    """A layout with one column placed on a single parity fails the check."""
    orig = G.s1
    try:
        G.s1 = lambda x, y, half: G.BASE1 + 2 * (x + 5 * y) + half  # low words all even
        with pytest.raises(AssertionError):
            G.check_layout(G.program())
    finally:
        G.s1 = orig


def test_committed_header_is_current():
    with open(os.path.join(CSRC, "keccak_round_asm.hpp")) as f:
        assert f.read() == G.header()
