"""Writes keccak_round_asm.hpp: one Keccak-f[1600] round (theta, rho, pi, chi, iota) on a
one-lane state, as gfx950 inline assembly over pinned VGPRs.

Why pinned registers: on gfx950 a three-source VALU op (v_bitop3_b32) issues at double
rate only when its three VGPR sources are not all of one parity (all even or all odd
register numbers); otherwise it takes the full four cycles of a single-rate op.  The
compiler's register allocator does not know this, and about 40% of the three-source ops
of a compiler-allocated round land on one parity.  Here every register is placed by hand:

  S1 (state A, 50 VGPRs): lane (x, y) -> pair 2i, i = x + 5y; low word parity (x + y) % 2,
      high word the other register of the pair.  Along a column the parities alternate,
      so theta's column sums xor3(A0, A1, A2) and xor3(C, A3, A4) always mix parities.
  S2 (rho/pi output B, 50 VGPRs): lane (x, y) -> pair 2j, j = x + 5y; low word parity
      x % 2, high word the other.  In each row the five low words read E O E O E and the
      high words O E O E O, so chi's window (x, x+1, x+2) is never of one parity.
      B(0, 0) = A(0, 0) (rho offset 0) stays in S1, whose parities match.
  C, R (theta's column parities and their rotations by one): 10 even and 10 odd
      registers of S2, dead before rho writes S2, so the fused xor3(A, C, R) always
      reads one even and one odd temporary.

Per round: 20 xor3 (C), 10 alignbit (R), 50 xor3 (A ^= C[x-1] ^ rot1(C[x+1])), 48 alignbit
(rho into S2), 50 bitop3 0xD2 (chi back into S1, natural order), 2 xor (iota): 180 VALU ops,
118 of them three-source, none of one parity.  The alignbits sit in two runs per round
(interleaving theta's xor3 with rho row by row or lane by lane measured no faster, and
needs 20 more registers for C and R).

v_bitop3_b32 truth-table index is S0 * 4 + S1 * 2 + S2 (0x96 = xor3, 0xD2 = S0 ^ (~S1 & S2)).
v_alignbit_b32 d, a, b, s = low 32 bits of ((a << 32 | b) >> s).

Round semantics follow FIPS 202 (the reference's SHA3 oracle, Sources/AppleZKProver/
SHA3Oracle.swift:181-280).  `emulate_round` interprets the emitted program on the host;
the tests compare it with the oracle's Keccak-f.

Run: python3 gen_keccak_round.py  (rewrites keccak_round_asm.hpp beside this file).
"""
import os

# FIPS 202 rho offsets r[x][y]
RHO = [[0, 36, 3, 41, 18], [1, 44, 10, 45, 2], [62, 6, 43, 15, 61], [28, 55, 25, 21, 56], [27, 20, 39, 8, 14]]

BASE1 = 8    # S1 = v8 .. v57
BASE2 = 58   # S2 = v58 .. v107


def s1(x, y, half):
    """VGPR of state word (x, y), half 0 = low, 1 = high."""
    i = x + 5 * y
    p = (x + y) % 2
    return BASE1 + 2 * i + (p if half == 0 else 1 - p)


def s2(x, y, half):
    if x == 0 and y == 0:
        return s1(0, 0, half)
    j = x + 5 * y
    q = x % 2
    return BASE2 + 2 * j + (q if half == 0 else 1 - q)


def c_reg(x, half):
    """Even register of S2 pair 1 + 2x + half."""
    return BASE2 + 2 * (1 + 2 * x + half)


def r_reg(x, half):
    return c_reg(x, half) + 1


def program():
    """The round as (op, dst, srcs, imm) tuples; srcs are VGPR numbers or 'rclo'/'rchi'."""
    p = []
    for half in (0, 1):  # first-level column sums
        for x in range(5):
            p.append(("bitop3", c_reg(x, half), (s1(x, 0, half), s1(x, 1, half), s1(x, 2, half)), 0x96))
    for half in (0, 1):
        for x in range(5):
            p.append(("bitop3", c_reg(x, half), (c_reg(x, half), s1(x, 3, half), s1(x, 4, half)), 0x96))
    for x in range(5):  # R = rotl64(C, 1)
        p.append(("alignbit", r_reg(x, 0), (c_reg(x, 0), c_reg(x, 1)), 31))
        p.append(("alignbit", r_reg(x, 1), (c_reg(x, 1), c_reg(x, 0)), 31))
    for y in range(5):  # A ^= C[x - 1] ^ R[x + 1]
        for x in range(5):
            for half in (0, 1):
                p.append(("bitop3", s1(x, y, half), (s1(x, y, half), c_reg((x + 4) % 5, half),
                                                      r_reg((x + 1) % 5, half)), 0x96))
    for y in range(5):  # rho and pi: B(y, 2x + 3y) = rotl64(A(x, y), r[x][y]); C and R are dead
        for x in range(5):
            n = RHO[x][y]
            if n == 0:
                continue
            assert n != 32
            tx, ty = y, (2 * x + 3 * y) % 5
            lo, hi = s1(x, y, 0), s1(x, y, 1)
            if n < 32:
                p.append(("alignbit", s2(tx, ty, 0), (lo, hi), 32 - n))
                p.append(("alignbit", s2(tx, ty, 1), (hi, lo), 32 - n))
            else:
                p.append(("alignbit", s2(tx, ty, 0), (hi, lo), 64 - n))
                p.append(("alignbit", s2(tx, ty, 1), (lo, hi), 64 - n))
    for y in range(5):  # chi into S1; row 0 writes (0, 0) last, where B(0, 0) lives
        for x in ((1, 2, 3, 4, 0) if y == 0 else range(5)):
            for half in (0, 1):
                p.append(("bitop3", s1(x, y, half), (s2(x, y, half), s2((x + 1) % 5, y, half),
                                                      s2((x + 2) % 5, y, half)), 0xD2))
    p.append(("xor", s1(0, 0, 0), ("rclo", s1(0, 0, 0)), None))
    p.append(("xor", s1(0, 0, 1), ("rchi", s1(0, 0, 1)), None))
    return p


def check_layout(prog):
    """Every register in range and distinct per role; no three-source op of one parity."""
    regs1 = {s1(x, y, h) for x in range(5) for y in range(5) for h in (0, 1)}
    regs2 = {s2(x, y, h) for x in range(5) for y in range(5) for h in (0, 1) if (x, y) != (0, 0)}
    assert len(regs1) == 50 and len(regs2) == 48 and not regs1 & regs2
    tmp = {c_reg(x, h) for x in range(5) for h in (0, 1)} | {r_reg(x, h) for x in range(5) for h in (0, 1)}
    assert len(tmp) == 20 and tmp <= regs2 | {BASE2, BASE2 + 1}
    for op, d, srcs, imm in prog:
        if op == "bitop3":
            assert len({r % 2 for r in srcs}) == 2, (op, d, srcs)


def emulate_round(state, rclo, rchi, prog=None):
    """Run the program on 25 64-bit lanes (index x + 5y); returns the new lanes."""
    prog = prog or program()
    m = 0xFFFFFFFF
    reg = {}
    for i, v in enumerate(state):
        x, y = i % 5, i // 5
        reg[s1(x, y, 0)] = v & m
        reg[s1(x, y, 1)] = v >> 32
    for op, d, srcs, imm in prog:
        if op == "bitop3":
            a, b, c = (reg[s] for s in srcs)
            out = 0
            for bit in range(32):
                idx = ((a >> bit) & 1) * 4 + ((b >> bit) & 1) * 2 + ((c >> bit) & 1)
                out |= ((imm >> idx) & 1) << bit
            reg[d] = out
        elif op == "alignbit":
            a, b = (reg[s] for s in srcs)
            reg[d] = (((a << 32) | b) >> imm) & m
        else:
            s = rclo if srcs[0] == "rclo" else rchi
            reg[d] = s ^ reg[srcs[1]]
    return [reg[s1(i % 5, i // 5, 0)] | (reg[s1(i % 5, i // 5, 1)] << 32) for i in range(25)]


def asm_text(prog):
    lines = []
    for op, d, srcs, imm in prog:
        if op == "bitop3":
            lines.append(f"v_bitop3_b32 v{d}, v{srcs[0]}, v{srcs[1]}, v{srcs[2]} bitop3:0x{imm:02x}")
        elif op == "alignbit":
            lines.append(f"v_alignbit_b32 v{d}, v{srcs[0]}, v{srcs[1]}, {imm}")
        else:
            lines.append(f"v_xor_b32 v{d}, {'%50' if srcs[0] == 'rclo' else '%51'}, v{srcs[1]}")
    return lines


def header():
    prog = program()
    check_layout(prog)
    outs = []
    for i in range(25):
        x, y = i % 5, i // 5
        outs.append(f'"+{{v{s1(x, y, 0)}}}"(a[{i}].lo)')
        outs.append(f'"+{{v{s1(x, y, 1)}}}"(a[{i}].hi)')
    clob = sorted({s2(x, y, h) for x in range(5) for y in range(5) for h in (0, 1) if (x, y) != (0, 0)} |
                  {c_reg(x, h) for x in range(5) for h in (0, 1)} | {r_reg(x, h) for x in range(5) for h in (0, 1)})
    text = asm_text(prog)
    out = ["// Generated by gen_keccak_round.py; do not edit.  One Keccak-f[1600] round on a one-lane",
           "// state held in pinned VGPRs v8..v57 (word (x, y): pair x + 5y, low word at parity (x + y) % 2)",
           "// with v58..v107 as rho/pi output and theta temporaries, laid out so that no three-source",
           "// VALU op reads three VGPRs of one parity (such an op issues at half rate on gfx950).",
           "// 180 VALU ops: 20 + 50 xor3, 10 + 48 alignbit, 50 chi bitop3, 2 iota xor.",
           "#pragma once", "",
           "__device__ __forceinline__ void kround_pinned(u2 (&a)[25], uint32_t rclo, uint32_t rchi)", "{",
           "    asm(\"" + " \\n\\t\"\n        \"".join(text) + "\"",
           "        : " + ",\n          ".join(outs),
           '        : "s"(rclo), "s"(rchi)',
           "        : " + ", ".join(f'"v{r}"' for r in clob) + ', "memory");', "}", ""]
    return "\n".join(out)


if __name__ == "__main__":
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "keccak_round_asm.hpp")
    with open(path, "w") as f:
        f.write(header())
    print("wrote", path)
