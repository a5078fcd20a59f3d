"""CPU: the high-precision series workload -- cast(expr, P, S) parsing, rewrite
and typing, the bounded generator, and the oracle's series evaluation pinned
by the restatement (row by row) and by mathematics (sin itself)."""
import math
import random
from fractions import Fraction

import numpy as np
import pytest

import pyref
from ultraprecise_amd import workloads as W


def test_pi_digits():
    # the first 60 decimals of pi (a published constant) and consistency of
    # the fixed-point routine across precisions
    want = "3141592653589793238462643383279502884197169399375105820974944"
    assert str(W.pi_scaled(60)) == want[:61]
    assert W.pi_scaled(140) // 10 ** 40 == W.pi_scaled(100)


def test_sin_constants_are_truncated_factorial_reciprocals():
    cs = W.sin_constants(25, 100)
    assert cs[0] == "1." + "0" * 100
    for k, c in enumerate(cs):
        sign, m, s = pyref.parse(c)
        assert (sign, s) == (0, 100)
        assert m == 10 ** 100 // math.factorial(2 * k + 1)


def test_cast_parse_and_types(U):
    cols = [("a", 38, 10)]
    # scale down: the cast is the plan root and keeps its declared type
    assert U.Expr("cast(a, 20, 3)", cols).plan == "F 20 3 9 (CAST 20 3 c0:38:10)"
    # scale up inside arithmetic: the cast is an operand of the rewritten sum,
    # sorted by scale like any operand (CmpByScale, AriesEngineShell.cpp:344-358)
    p = U.Expr("cast(a, 40, 12) + a", cols).plan
    assert p == "F 41 12 18 (ADD 41 12 c0:38:10 (CAST 40 12 c0:38:10))"
    # a cast argument is rewritten on its own: 1 + a + 2 folds to 3 + a
    k = format(3 * 10 ** 10, "x")
    assert U.Expr("cast(1 + a + 2, 38, 5)", cols).plan == f"F 38 5 16 (CAST 38 5 (ADD 39 10 k+{k}:10 c0:38:10))"
    for bad in ["cast(a, 3)", "cast(a, 2, 5)", "cast(a 3, 2)"]:
        with pytest.raises(U.UpdecError):
            U.Expr(bad, cols)


def test_cast_envelope(U):
    cols = [("a", 38, 10)]
    assert U.Expr("cast(a, 60, 48)", cols).faithful            # up by 38: one table row
    assert not U.Expr("cast(a, 70, 49)", cols).faithful        # up by 39: past the table
    assert U.Expr("cast(a, 38, 0)", cols).faithful             # down: chunked, any shift
    assert not U.Expr("cast(a, 300, 100)", cols).faithful      # result past the byte table


def test_repeated_subtree_is_computed_once(U):
    e = U.Expr(W.sin_expr(terms=4), [("x", 300, 100, 334)])
    src = e.source
    # x*x appears in every Horner step of the text but once in the code
    assert e.plan.count("(MUL 600 200 c0:300:100 c0:300:100)") == 3
    assert src.count("dmul<21, 11, 11>(n0, n0)") == 1


def test_gen_below_bounds(U):
    bound = W.pi_scaled(100) + 1
    rows = 4000
    host = U.gen_compact_host(W.sin_desc(U, 3), rows)
    nb = U.compact_bytes(300)
    vals = [pyref.compact_decode(bytes(host[i * nb:(i + 1) * nb]), 100) for i in range(rows)]
    mags = [v[1] for v in vals]
    assert max(mags) < bound
    # uniform on [0, bound): mean within 5 standard errors, both signs present
    mean = Fraction(sum(mags), rows) / bound
    assert abs(float(mean) - 0.5) < 5 * (1 / math.sqrt(12 * rows))
    negs = sum(v[0] for v in vals)
    assert abs(negs - rows / 2) < 5 * math.sqrt(rows / 4)
    # counter-based: any row range reproduces the same bytes
    part = U.gen_compact_host(W.sin_desc(U, 3), 100, row0=1234)
    assert (part == host[1234 * nb:1334 * nb]).all()


def horner(x, cs, faithful=False):
    """the series in the restatement: s = c_k - cast(y * s), y = cast(x*x)"""
    y = pyref.cast(pyref.mul(x, x, faithful), 100)
    s = cs[-1]
    for k in range(len(cs) - 2, -1, -1):
        t = pyref.cast(pyref.mul(y, s, faithful), 100)
        s = pyref.add(pyref.neg(t), cs[k], faithful)
    return pyref.cast(pyref.mul(x, s, faithful), 100)


def sin_fixed(m, S, guard=30):
    """sin(m / 10^S) * 10^(S+guard), rounded toward zero, by the Taylor
    series in exact integers with guard digits (a textbook evaluation,
    independent of the Horner form)"""
    one = 10 ** (S + guard)
    x = m * 10 ** guard
    term, total, k = x, 0, 1
    while term:
        total += term
        term = -term * x // one * x // one // ((k + 1) * (k + 2))
        k += 2
    return total


def test_series_oracle_matches_restatement_and_sin(U, O):
    rows = 300
    host = U.gen_compact_host(W.sin_desc(U, 3), rows)
    nb = U.compact_bytes(300)
    bits = max(int.from_bytes(bytes(host[i * nb:(i + 1) * nb]), "little") & ((1 << (8 * nb - 1)) - 1)
               for i in range(rows)).bit_length()
    e = U.Expr(W.sin_expr(), [("x", 300, 100, bits)])
    assert not e.faithful and (e.precision, e.scale) == (300, 100)
    out = O.plan_eval(O.Plan(e.plan), [host], [300], [100], rows, e.out_bytes, mode=O.EXTENDED)
    cs = [pyref.parse(c) for c in W.sin_constants(25, 100)]
    worst = 0
    for i in range(rows):
        x = pyref.compact_decode(bytes(host[i * nb:(i + 1) * nb]), 100)
        r = horner(x, cs)
        got = bytes(out[i * e.out_bytes:(i + 1) * e.out_bytes])
        assert got == pyref.compact_encode(r, e.out_bytes), i
        # mathematics: within the series remainder pi^51/51! plus the
        # truncation error of 25 Horner steps
        sm = sin_fixed(x[1], 100)
        want = Fraction(-sm if x[0] else sm, 10 ** 130)
        have = Fraction(-r[1] if r[0] else r[1], 10 ** 100)
        worst = max(worst, abs(have - want))
    assert worst < Fraction(1, 10 ** 40)


def test_series_negative_truncates_to_minus_zero(O):
    # cast keeps the sign of a value that truncates to zero (decimal.cu:471-533)
    d = O.dec("-0.000000001")
    O.lib.od_cast(O.ctypes.byref(d), 10, 3, 0)
    assert d.sign == 1 and O.to_str(d).lstrip("-") in ("0.000", "0")
    assert pyref.cast((1, 1, 9), 3) == (1, 0, 3)


def test_series_kernels_compile(U):
    cols = [("x", 300, 100, 334)]
    e = U.Expr(W.sin_expr(), cols)
    assert "__noinline__" in e.source
    U.check_compile(e.source)
    a = U.Agg([W.sin_expr()], cols)
    U.check_compile(a.source(1))


@pytest.mark.parametrize("seed", [1, 2])
def test_cast_restatement_random(O, seed):
    """oracle cast against the restatement on random values and shifts"""
    rng = random.Random(seed)
    for _ in range(400):
        S = rng.randrange(0, 120)
        T = rng.randrange(0, 160)
        m = rng.getrandbits(rng.randrange(1, 900))
        sign = rng.getrandbits(1)
        d = O.OdDec()
        d.set_magnitude(m)
        d.sign, d.frac, d.prec = sign, S, 300
        O.lib.od_cast(O.ctypes.byref(d), 300, T, O.EXTENDED)
        assert (d.sign, d.magnitude, d.frac) == pyref.cast((sign, m, S), T)
