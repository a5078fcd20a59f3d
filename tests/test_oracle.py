"""CPU: pin the C oracle.

1. the reference's own known answers (tests/golden/known_answers.json,
   transcribed from test/testcases/datatypes/test_decimal.cu);
2. exact rational arithmetic (fractions.Fraction) on brute-force tiny
   operands: what the mathematics fixes for add/sub/mul/div;
3. an independent Python-integer restatement (tests/pyref.py) on random
   operands of every size, both modes, including mod 2^1024 wrap;
4. closed forms of the compact layout.
"""
import json
import os
import random
from fractions import Fraction

import pytest

import pyref

HERE = os.path.dirname(os.path.abspath(__file__))
KA = json.load(open(os.path.join(HERE, "golden", "known_answers.json")))


def val(d):
    """(sign, magnitude, scale) of an oracle value"""
    return d.sign, d.magnitude, d.frac


def mk(O, v, mode=0):
    d = O.OdDec()
    d.sign, d.frac, d.prec = v[0], v[2], 30
    d.set_magnitude(v[1])
    return d


@pytest.mark.parametrize("case", KA["binary"], ids=[c["case"] for c in KA["binary"]])
def test_reference_known_answers(O, case):
    a = O.dec_cast(case["a"][2], case["a"][0], case["a"][1])
    b = O.dec_cast(case["b"][2], case["b"][0], case["b"][1])
    r = O.binop(case["op"], a, b)
    got = O.to_str(r)
    want = case["expect"]
    want_scale = len(want.split(".")[1]) if "." in want else 0
    if case["op"] != "div" and want_scale == r.frac:
        assert got == want, case["case"]
    elif case["op"] != "div":
        # printed with fewer fractional digits by an older scale rule: the
        # printed digits are the exact result truncated to that scale
        assert want_scale < r.frac
        g = Fraction(got)
        t = Fraction(int(g * 10 ** want_scale), 10 ** want_scale)
        assert t == Fraction(want), case["case"]
    else:
        # printed with Sa + 6 digits by the older rule; the current rule keeps Sa + 4
        assert r.frac == case["a"][1] + 4
        w = Fraction(want)
        assert Fraction(got) == Fraction(round(w * 10 ** r.frac), 10 ** r.frac), case["case"]


def test_reference_expression_all_2(O):
    c = KA["expressions"][0]
    d2, d3, d1 = O.dec(c["d2"]), O.dec(c["d3"]), O.dec(c["d1"])
    r = O.binop("div", O.binop("mul", d2, d3), d1)
    assert r.frac == 6
    w = Fraction(c["expect"])
    assert Fraction(O.to_str(r)) == Fraction(round(w * 10 ** 6), 10 ** 6)


@pytest.mark.parametrize("case", KA["truncate"], ids=[str(t["line"]) for t in KA["truncate"]])
def test_reference_truncate(O, case):
    """Decimal(P, S, text).truncate(p) (test_decimal.cu:672-745), every p"""
    import ctypes
    P, S, text = case["a"]
    d = O.dec_cast(text, P, S)
    O.lib.od_truncate(ctypes.byref(d), case["p"], 0)
    assert O.to_str(d) == case["expect"]


@pytest.mark.parametrize("case", [t for t in KA["truncate"] if t["p"] > 0])
def test_reference_cast_down_up(O, case):
    """for p > 0 truncate and cast to scale p agree"""
    import ctypes
    P, S, text = case["a"]
    d = O.dec_cast(text, P, S)
    O.lib.od_cast(ctypes.byref(d), P, case["p"], 0)
    assert O.to_str(d) == case["expect"]


def test_truncate_against_closed_form(O):
    import ctypes
    rng = random.Random(41)
    for _ in range(3000):
        P = rng.randrange(1, 80)
        S = rng.randrange(0, P + 1)
        v = (rng.getrandbits(1), rng.randrange(0, 10 ** P), S)
        p = rng.randrange(-P - 3, 40)
        d = mk(O, v)
        d.prec = P
        O.lib.od_truncate(ctypes.byref(d), p, 1)
        assert (d.sign, d.magnitude, d.frac, d.prec) == pyref.truncate(v, P, p), (v, P, p)


@pytest.mark.parametrize("c", KA["compact_round_trip"])
def test_reference_compact_round_trip(O, c):
    d = O.dec_cast(c["text"], c["P"], c["S"])
    nb = O.lib.od_compact_bytes(c["P"])
    buf = O.to_compact(d, nb)
    e = O.from_compact(buf, c["P"], c["S"])
    assert O.to_str(e) == O.to_str(d)
    assert buf == pyref.compact_encode(val(d), nb)


def test_reference_int_ctors(O):
    import ctypes
    for c in KA["int_ctors"]:
        d = O.OdDec()
        O.lib.od_from_int64(ctypes.byref(d), c["value"])
        assert (d.sign, d.magnitude) == (1 if c["value"] < 0 else 0, abs(c["value"]))
        assert O.to_str(d) == str(c["value"])


def test_compact_bytes_table_holds_every_value(O):
    """B(P) never truncates a DECIMAL(P,*) magnitude plus the sign bit, and
    past the table it is the closed form"""
    for P in range(1, 300):
        need = ((10 ** P - 1).bit_length() + 1 + 7) // 8
        assert O.lib.od_compact_bytes(P) >= need
        assert O.lib.od_compact_bytes(P) <= need + 1
    for P in (300, 301, 450, 602):
        assert O.lib.od_compact_bytes(P) == ((10 ** P - 1).bit_length() + 1 + 7) // 8
    # the widths of the workloads' precisions (decimal.cu:92 table)
    assert [O.lib.od_compact_bytes(P) for P in (15, 18, 38, 76, 150)] == [7, 8, 16, 32, 63]


def frac_of(v):
    s, m, x = v[0], v[1], v[2]
    return Fraction((-1) ** s * m, 10 ** x)


def test_brute_force_tiny_against_rationals(O):
    """every pair of small operands: add/sub/mul exact, div round-half-up of
    the exact quotient, and the sign rules for zero"""
    vals = []
    for s in (0, 1, 2):
        for m in range(0, 13):
            for sign in (0, 1):
                vals.append((sign, m, s))
    for a in vals:
        for b in vals:
            da, db = mk(O, a), mk(O, b)
            r = val(O.binop("add", da, db))
            assert frac_of(r) == frac_of(a) + frac_of(b) and r[2] == max(a[2], b[2])
            if r[1] == 0:
                assert r[0] == (1 if (a[0] == b[0] == 1) else 0)
            else:
                assert r[0] == (1 if frac_of(r) < 0 else 0)
            r = val(O.binop("sub", da, db))
            assert frac_of(r) == frac_of(a) - frac_of(b)
            r = val(O.binop("mul", da, db))
            assert frac_of(r) == frac_of(a) * frac_of(b) and r[2] == a[2] + b[2]
            assert r[0] == a[0] ^ b[0]
            q = O.binop("div", da, db)
            S = a[2] + 4
            assert q.frac == S
            if a[1] == 0:
                assert (q.sign, q.magnitude, q.error) == (0, 0, 0)
            elif b[1] == 0:
                assert (q.sign, q.magnitude, q.error) == (a[0] ^ b[0], a[1], 2)
            else:
                exact = Fraction(a[1], 10 ** a[2]) / Fraction(b[1], 10 ** b[2]) * 10 ** S
                assert q.magnitude == int(exact + Fraction(1, 2)), (a, b)
                assert q.sign == a[0] ^ b[0]


def rand_mag(rng, bits):
    return rng.getrandbits(bits) if bits > 0 else 0


@pytest.mark.parametrize("mode", [0, 1])
def test_random_against_python_ints(O, mode):
    rng = random.Random(1234 + mode)
    faithful = mode == 0
    for i in range(3000):
        ba = rng.choice([0, 8, 31, 32, 33, 63, 64, 65, 100, 200, 500, 1000, 1023, 1024])
        bb = rng.choice([0, 8, 31, 32, 33, 63, 64, 65, 100, 200, 500, 1000, 1023, 1024])
        a = (rng.getrandbits(1), rand_mag(rng, ba), rng.randrange(0, 12))
        b = (rng.getrandbits(1), rand_mag(rng, bb), rng.randrange(0, 12))
        if not faithful:
            a = (a[0], a[1] >> 400, a[2])
            b = (b[0], b[1] >> 400, b[2])
        da, db = mk(O, a), mk(O, b)
        for op in ("add", "sub", "mul"):
            got = val(O.binop(op, da, db, mode))
            want = getattr(pyref, op)(a, b, faithful)
            assert got == want, (op, a, b)
        got = O.binop("div", da, db, mode)
        if faithful and a[1] and b[1]:
            # beyond one chunk the reference's DivOrMod loop is outside its
            # defined behaviour; only single-chunk dividends are compared
            dtz = (a[1].bit_length() - 1) // 32
            if (32 - dtz) * 9 - 10 < b[2] + 4 or (a[1] * 10 ** (b[2] + 4)).bit_length() < b[1].bit_length() - 64:
                continue
        want = pyref.div(a, b, faithful)
        assert (got.sign, got.magnitude, got.frac, got.error) == want, ("div", a, b)


def test_division_paths_hit(O):
    """each DivOrMod path with its quirk, against the restatement"""
    cases = [
        # path (i): uint64 numerator wraps (bitlen estimate 64, true product 65 bits)
        ((0, 2_200_000_000_000_000, 0), (0, 3, 0)),
        # path (i): no wrap
        ((0, 12345, 2), (1, 7, 0)),
        # path (ii): single-word divisor, remainder >= 2^31 makes 2R wrap in uint32
        ((0, (1 << 70) + 5, 0), (0, 3_000_000_001, 0)),
        # path (iii): bitlen(X) == bitlen(B) - 1 gives 1
        # (X = 2^86 * 10^4 has 100 bits, B has 101, and 2X < B)
        ((0, 1 << 86, 0), (0, (1 << 101) - (1 << 97), 0)),
        # path (iii): ordinary
        ((1, (1 << 100) + 12345, 3), (0, (1 << 70) + 999, 1)),
    ]
    for a, b in cases:
        got = O.binop("div", mk(O, a), mk(O, b), 0)
        assert (got.sign, got.magnitude, got.frac, got.error) == pyref.div(a, b, True), (a, b)
    # the wrapping cases do differ from exact rounding
    a, b = cases[0]
    assert pyref.div(a, b, True)[1] != pyref.div(a, b, False)[1]
    a, b = cases[3]
    assert pyref.div(a, b, True)[1] == 1 and pyref.div(a, b, False)[1] == 0


def test_faithful_wraps_mod_2_1024(O):
    big = (0, (1 << 1024) - 1, 0)
    one = (0, 1, 0)
    r = val(O.binop("add", mk(O, big), mk(O, one), 0))
    assert r == (0, 0, 0)
    r = val(O.binop("mul", mk(O, (0, 1 << 1000, 0)), mk(O, (1, 1 << 30, 0)), 0))
    assert r == (1, (1 << 1030) % (1 << 1024), 0)


def test_plan_eval_matches_restatement(O):
    """a rewritten plan evaluated by the oracle equals a direct evaluation
    with the restatement, row by row"""
    import numpy as np
    rng = random.Random(9)
    P, S = 18, 4
    nb = O.lib.od_compact_bytes(P)
    rows = 500
    cols = []
    vals = []
    for c in range(4):
        vs = [(rng.getrandbits(1), rng.randrange(10 ** 17, 10 ** 18), S) for _ in range(rows)]
        vals.append(vs)
        cols.append(np.frombuffer(b"".join(pyref.compact_encode(v, nb) for v in vs), dtype=np.uint8))
    plan = O.Plan("(ADD 38 8 (MUL 37 8 (ADD 19 4 c0:18:4 c1:18:4) c2:18:4) (DIV 9 8 (NEG c0:18:4) c3:18:4))")
    out_b = O.lib.od_compact_bytes(38)
    got = O.plan_eval(plan, cols, [P] * 4, [S] * 4, rows, out_b)
    for i in range(rows):
        a, b, c, d = (vals[k][i] for k in range(4))
        q = pyref.div(pyref.neg(a), d)
        r = pyref.add(pyref.mul(pyref.add(a, b), c), q[:3])
        assert bytes(got[i * out_b:(i + 1) * out_b]) == pyref.compact_encode(r, out_b)


def test_wide_oracle_sums_exact_products(U, O):
    """the 640-limb build (liboracle_wide.so) against Python integers on
    products far past 192 limbs, and against the default build on narrow
    plans"""
    import random
    rng = random.Random(5)
    P, S = 300, 0
    B = O.lib.od_compact_bytes(P)
    rows = 6
    vals = [rng.randrange(10 ** 299, 10 ** 300) * rng.choice((1, -1)) for _ in range(rows)]
    col = bytearray()
    for v in vals:
        m = abs(v).to_bytes(B, "little")
        col += m[:-1] + bytes([m[-1] | (0x80 if v < 0 else 0)])
    import numpy as np
    arr = np.frombuffer(bytes(col), dtype=np.uint8)
    e = U.Expr("a*a*a*a*a*a*a*a*a*a*a*a*a*a*a", [("a", P, S)])    # about 15,000 bits per row
    assert not e.faithful
    img, n = O.plan_sum_wide(e.plan, [arr], [P], [S], rows)
    total = sum(v ** 15 for v in vals)
    assert n == rows
    assert img[1] == (1 if total < 0 else 0)
    assert int.from_bytes(img[5:133], "little") == abs(total) % (1 << 1024)
    e2 = U.Expr("a*a+a", [("a", P, S)])
    assert O.plan_sum_wide(e2.plan, [arr], [P], [S], rows)[0] == O.plan_sum(O.Plan(e2.plan), [arr], [P], [S], rows,
                                                                            mode=O.EXTENDED)[0]
