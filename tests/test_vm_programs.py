"""CPU: interpreter programs built from the rewritten plans -- width class,
register allocation (Sethi-Ullman order, registers reused), operand
alignment shifts, and the serving workload's query generator."""
import pytest

from ultraprecise_amd import workloads as W

C15 = [("q", 15, 2), ("p", 15, 2), ("d", 15, 2)]


def test_plain_method_programs(U):
    v = U.VmProgram("p*(1-d)", C15)
    # plan (MUL 31 4 c1 (SUB 16 2 k+64:2 c2)): the constant 1.00 is k0
    # the SUB result is forwarded to the MUL without a register write
    assert v.listing == "W4 regs 1 root acc\nSUB - k0 c2\nMUL - c1 acc\n"
    assert (v.precision, v.scale, v.faithful) == (31, 4, True)
    assert U.VmProgram("q", C15).listing == "W2 regs 1 root acc\nMOV - c0\n"


def test_sweep_program_alignment_and_registers(U):
    v = U.VmProgram("(a+b)*c - a/d", [(n, 38, 10) for n in "abcd"])
    # the division result (scale 14) is aligned to the product's scale 20
    assert v.listing == ("W8 regs 2 root acc\nNEG - c0\nDIV r0 acc c3 #14\nADD - c0 c1\nMUL - acc c2\n"
                         "ADD - r0<<6 acc\n")
    v = U.VmProgram("(a+b)*c - a/d", [(n, 150, 50) for n in "abcd"])
    assert v.width == 32 and not v.faithful
    # P = 300 needs more than 32 limbs: a multi-lane class
    v = U.VmProgram("(a+b)*c - a/d", [(n, 300, 100) for n in "abcd"])
    assert v.width == 64 and not v.faithful
    with pytest.raises(U.UpdecError, match="exceed"):
        U.VmProgram("*".join(["a"] * 17), [("a", 300, 100)])   # about 17,000 bits


def test_series_program(U):
    v = U.VmProgram(W.sin_expr(terms=3), [("x", 300, 100, 334)])
    lines = v.listing.splitlines()
    assert lines[0] == "W32 regs 2 root acc"
    assert lines[1:3] == ["MUL - c0 c0", "DOWN r0 acc #100"]
    assert lines[-2:] == ["MUL - c0 acc", "DOWN - acc #100"]
    # 25 terms: y = cast(x*x) is recomputed per use, registers stay at 2
    v = U.VmProgram(W.sin_expr(), [("x", 300, 100, 334)])
    # per Horner step: y (MUL, DOWN), MUL, DOWN, NEG, ADD; then x * s (MUL, DOWN)
    assert v.nregs == 2 and v.nins == 24 * 6 + 2


def test_registers_bounded_by_tree_need(U):
    cols = [(n, 15, 2) for n in "abcd"]
    # a balanced tree of depth 3 needs 3 registers (Strahler number + 0)
    v = U.VmProgram("((a+b)*(c+d))-((a-c)*(b-d))", cols)
    assert v.nregs == 3
    # a left-deep chain needs one
    assert U.VmProgram("((a*b)*c)*d", cols).nregs == 1
    # the rewrite orders operands by scale: -d (scale 2) is evaluated
    # before the product (scale 4), so the product needs a second register
    assert U.VmProgram("((a+b)*c)-d", cols).listing == (
        "W4 regs 2 root acc\nNEG r0 c3\nADD - c0 c1\nMUL - acc c2\nADD - r0<<2 acc\n")


def test_serving_queries_are_seeded_and_fit(U):
    def ok(text, cols):
        try:
            U.VmProgram(text, cols)
            return True
        except U.UpdecError as e:
            assert "exceed" in str(e) or "JIT path only" in str(e)
            return False

    qs = [W.random_query(4, i, ok) for i in range(300)]
    assert qs == [W.random_query(4, i, ok) for i in range(300)]
    assert {q["P"] for q in qs} == set(W.CONFIG5["precisions"])
    for q in qs:
        assert q["S"] == q["P"] // 3
        assert 1000 <= q["rows"] <= 10_000_000
        assert 2 <= len(q["cols"]) <= 4
        depth, d = 0, 0
        for ch in q["text"]:
            d += ch == "("
            d -= ch == ")"
            depth = max(depth, d)
        assert 2 <= depth <= 6
        v = U.VmProgram(q["text"], q["cols"])
        assert v.width in U.VM_CLASS_WIDTHS
