"""CPU: pin the oracle's division and MOD (closed forms plus path predicates,
oracle/updec_oracle.c od_div / od_mod, and tests/pyref.py) against a literal
step-by-step restatement of the reference's procedures (tests/refproc.py:
DivOrMod, DivByInt64, DivByInt, DivInt, HalfIntDecimal, ModCalc).

The two are written differently on purpose: the oracle picks a path from bit
lengths and applies its quirk in one expression, the restatement runs the
reference's chunk loops, uint64/uint32 arithmetic and bisection as written.
A misread predicate (a path boundary, a wrap, the bisection's collapse range)
makes them disagree on one of the crafted or random cases below.
"""
import random

import pytest

import pyref
import refproc as R


def od(O, v):
    d = O.OdDec()
    d.sign, d.frac, d.prec = v[0], v[2], 30
    d.set_magnitude(v[1])
    return d


def literal_div(a, b):
    r, path, chunks = R.div_or_mod(R.Dec(a[1], a[0], a[2]), R.Dec(b[1], b[0], b[2]))
    return (r.sign, r.mag, r.frac, r.error), path, chunks


def literal_mod(a, b):
    r, path = R.mod_calc(R.Dec(a[1], a[0], a[2]), R.Dec(b[1], b[0], b[2]))
    return (r.sign, r.mag, r.frac), path


def check_div(O, a, b):
    """compare oracle, closed-form restatement and literal procedure; returns
    the reference's path, or None when the reference leaves the case
    undefined or takes more than one chunk (outside the faithful contract)"""
    try:
        want, path, chunks = literal_div(a, b)
    except R.Undefined:
        return None
    if chunks > 1:
        return None
    got = O.binop("div", od(O, a), od(O, b), 0)
    assert (got.sign, got.magnitude, got.frac, got.error) == want, ("div", a, b, path)
    assert pyref.div(a, b, True) == want, ("pyref div", a, b, path)
    return path


def check_mod(O, a, b):
    try:
        want, path = literal_mod(a, b)
    except R.Undefined:
        return None
    got = O.binop("mod", od(O, a), od(O, b), 0)
    assert (got.sign, got.magnitude, got.frac) == want, ("mod", a, b, path)
    assert pyref.mod(a, b, True) == want, ("pyref mod", a, b, path)
    return path


def est_bits(s):
    return (s // 3) * 10 + {0: 0, 1: 4, 2: 7}[s % 3]


def rand_bits(rng, n):
    """a value of exactly n bits (n >= 1)"""
    return (1 << (n - 1)) | rng.getrandbits(n - 1) if n > 1 else 1


def test_literal_restatement_basics():
    """the restatement's own building blocks against plain integers"""
    rng = random.Random(5)
    for _ in range(200):
        m, n = rng.getrandbits(rng.randrange(1, 900)), rng.randrange(0, 39)
        assert R.shl10(R.Dec(m), n).mag == (m * 10 ** n) % R.MOD
        x, y = rng.getrandbits(rng.randrange(1, 1000)), rng.getrandbits(rng.randrange(1, 1000))
        assert R.half_int(R.Dec(x), R.Dec(y)).mag == ((x + y) % R.MOD) // 2
        assert R.lt(R.Dec(x), R.Dec(y)) == (x < y)
        ds, dt = R.Dec(max(x, y)), R.Dec(min(x, y) or 1)
        if (ds.mag.bit_length() >= dt.mag.bit_length() + 1):
            q, res = R.div_int(ds, dt, dt.mag.bit_length())
            assert (q.mag, res.mag, res.sign) == (ds.mag // dt.mag, ds.mag % dt.mag, 0)
    with pytest.raises(R.Undefined):
        R.shl10(R.Dec(1), 39)


def test_div_every_path_random(O):
    rng = random.Random(77)
    seen = {"i": 0, "ii": 0, "iii": 0}
    for i in range(900):
        kind = i % 3
        fb = rng.randrange(0, 31)
        s = fb + 4
        if kind == 0:        # uint64 path, wraps included
            la = rng.randrange(1, max(2, 66 - est_bits(s)) + 1)
            lb = rng.randrange(1, 65)
        elif kind == 1:      # one-word divisor
            la = rng.randrange(1, 900)
            lb = rng.randrange(1, 33)
        else:                # bisection: quotient of at most ~190 bits
            lb = rng.randrange(33, 900)
            lx = lb + rng.randrange(-70, 190)
            la = max(1, min(1000, lx - est_bits(s)))
        a = (rng.getrandbits(1), rand_bits(rng, la), rng.randrange(0, 20))
        b = (rng.getrandbits(1), rand_bits(rng, lb), fb)
        p = check_div(O, a, b)
        if p:
            seen[p] += 1
    assert min(seen.values()) >= 150, seen


def test_div_bisection_gap_sweep(O):
    """DivInt's collapse: a numerator 1..64 bits shorter than the divisor
    gives quotient 1 with no round-up; 65 bits or more stores outside qmax"""
    rng = random.Random(11)
    for gap in range(-3, 70):
        for _ in range(3):
            fb = rng.randrange(0, 25)
            s = fb + 4
            a = (rng.getrandbits(1), rand_bits(rng, rng.randrange(70, 400)), rng.randrange(0, 10))
            lx = (a[1] * 10 ** s).bit_length()
            if lx + gap > 1020:
                continue
            b = (rng.getrandbits(1), rand_bits(rng, lx + gap), fb)
            if gap >= 65:
                with pytest.raises(R.Undefined):
                    literal_div(a, b)
                continue
            assert check_div(O, a, b) == "iii"
            if gap >= 1:
                assert O.binop("div", od(O, a), od(O, b), 0).magnitude == 1


def test_div_uint_wraps(O):
    """path (i)'s uint64 numerator and doubled remainder and path (ii)'s
    uint32 doubled remainder, crafted so that each wrap changes the result"""
    rng = random.Random(3)
    wrapped_i = wrapped_ii = 0
    for _ in range(400):
        s = rng.randrange(4, 20)
        a = (0, rand_bits(rng, 65 - est_bits(s)), 0)
        b = (0, rand_bits(rng, rng.randrange(1, 65)), s - 4)
        if check_div(O, a, b) == "i" and pyref.div(a, b, True) != pyref.div(a, b, False):
            wrapped_i += 1
        b = (0, rand_bits(rng, 32), rng.randrange(0, 20))
        a = (0, rand_bits(rng, rng.randrange(70, 600)), 0)
        if check_div(O, a, b) == "ii" and pyref.div(a, b, True) != pyref.div(a, b, False):
            wrapped_ii += 1
    assert wrapped_i >= 20 and wrapped_ii >= 20, (wrapped_i, wrapped_ii)


def test_div_collapse_known_answer(O):
    """1.00 (9,2) / 12345678901.0000000000 (38,10): X = 10^16 is 13 bits
    shorter than the divisor, so the reference's DivInt returns 1 (scale 6:
    0.000001) where exact rounding gives 0.000000"""
    a = O.dec_cast("1.00", 9, 2)
    b = O.dec_cast("12345678901.0000000000", 38, 10)
    assert O.to_str(O.binop("div", a, b, 0)) == "0.000001"
    assert O.to_str(O.binop("div", a, b, 1)) == "0.000000"
    want, path, _ = literal_div((0, 100, 2), (0, 123456789010000000000, 10))
    assert path == "iii" and want == (0, 1, 6, 0)


def test_mod_every_path_random(O):
    rng = random.Random(91)
    seen = {"i": 0, "ii": 0, "iii": 0}
    for i in range(900):
        kind = i % 3
        if kind == 0:
            la, lb = rng.randrange(1, 65), rng.randrange(1, 65)
        elif kind == 1:
            la, lb = rng.randrange(65, 1000), rng.randrange(1, 33)
        else:
            lb = rng.randrange(33, 1000)
            la = max(1, min(1023, lb + rng.randrange(-70, 160)))
        a = (rng.getrandbits(1), rand_bits(rng, la), rng.randrange(0, 30))
        b = (rng.getrandbits(1), rand_bits(rng, lb), rng.randrange(0, 30))
        p = check_mod(O, a, b)
        if p:
            seen[p] += 1
    assert min(seen.values()) >= 150, seen


def test_mod_bisection_gap_sweep(O):
    rng = random.Random(13)
    for gap in range(-3, 70):
        for _ in range(3):
            la = rng.randrange(70, 900)
            if la + gap > 1020:
                continue
            a = (rng.getrandbits(1), rand_bits(rng, la), 0)
            b = (rng.getrandbits(1), rand_bits(rng, la + gap), 0)
            if gap >= 65:
                with pytest.raises(R.Undefined):
                    literal_mod(a, b)
                continue
            assert check_mod(O, a, b) == "iii"
            if gap >= 0:
                assert O.binop("mod", od(O, a), od(O, b), 0).magnitude == abs(a[1] - b[1])
