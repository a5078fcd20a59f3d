"""CPU: MOD (ModCalc, decimal.cu:2621-2716) -- the oracle's literal
restatement (uint64 path, one-limb long division, DivInt's bisection with
32-limb products) against the closed form in tests/pyref.py, on operand pairs
built to hit every path and bit-length relation; the RSA workload's
arithmetic (modular products of positive integers) against Python's own;
and the planner's MOD typing (P = P of the divisor, S = 0) and rewrite."""
import ctypes
import random

import pytest

import pyref


def odv(O, sign, m, frac=0):
    d = O.OdDec()
    d.sign, d.frac, d.prec = sign, frac, 30
    d.set_magnitude(m)
    return d


def oracle_mod(O, a, b, mode):
    r = O.binop("mod", odv(O, *a), odv(O, *b), mode)
    return r.sign, r.magnitude, r.frac


def pairs(rng):
    out = []
    for _ in range(3000):
        lb = rng.choice([1, 5, 31, 32, 33, 63, 64, 65, 96, 128, 200, 400, 900])
        rel = rng.choice([-40, -3, -2, -1, 0, 0, 0, 1, 1, 2, 5, 60, 300])
        la = max(1, min(1000, lb + rel))
        A = rng.getrandbits(la) | (1 << (la - 1))
        B = rng.getrandbits(lb) | (1 << (lb - 1))
        if rng.random() < 0.03:
            A = 0
        if rng.random() < 0.03:
            B = 0
        out.append(((rng.randint(0, 1), A, rng.randint(0, 6)), (rng.randint(0, 1), B, rng.randint(0, 6))))
    return out


@pytest.mark.parametrize("mode", [0, 1])
def test_mod_literal_matches_closed_form(O, mode):
    rng = random.Random(17 + mode)
    hit = set()
    for a, b in pairs(rng):
        got = oracle_mod(O, a, b, mode)
        want = pyref.mod(a, b, faithful=(mode == 0))
        assert got == want, (a, b)
        la, lb = a[1].bit_length(), b[1].bit_length()
        if a[1] and b[1]:
            hit.add("i" if la <= 64 and lb <= 64 else "ii" if b[1] < 1 << 32 else
                    "iii>" if la > lb else "iii=" if la >= lb - 1 else "iii<")
    assert hit == {"i", "ii", "iii>", "iii=", "iii<"}


def test_mod_faithful_quirk_values(O):
    # same bit length, a < b: the bisection keeps quotient 1, |a - b| is returned
    a, b = (0, 99999999999999999999, 0), (0, 100000000000000000000, 0)
    assert oracle_mod(O, a, b, 0) == (0, 1, 0)
    assert oracle_mod(O, a, b, 1) == (0, 99999999999999999999, 0)
    # the sign is the xor of the signs, scales are dropped
    assert oracle_mod(O, (1, 17, 0), (0, 5, 0), 0) == (1, 2, 0)
    assert oracle_mod(O, (1, 17, 0), (1, 5, 0), 0) == (0, 2, 0)
    assert oracle_mod(O, (0, 25, 1), (0, 2, 0), 0) == (0, 1, 0)
    assert oracle_mod(O, (1, 0, 3), (1, 5, 0), 0) == (0, 0, 0)


def test_rsa_arithmetic(O):
    """col1 * col1 % N * col1 % N for the four RSA key lengths: exact modular
    arithmetic for positive operands (every bisection there has la > lb)"""
    Ns = [514679589732349351, 331972734823750661745273085756461571,
          273420351015258185841522860354335753665897876941082396543621156533322691]
    rng = random.Random(5)
    for N, P in zip(Ns, (17, 35, 71)):
        for _ in range(200):
            x = rng.randrange(10 ** (P - 1), 10 ** P)
            t = O.binop("mul", odv(O, 0, x), odv(O, 0, x), 0)
            t = O.binop("mod", t, odv(O, 0, N), 0)
            t = O.binop("mul", t, odv(O, 0, x), 0)
            t = O.binop("mod", t, odv(O, 0, N), 0)
            assert t.magnitude == x * x % N * x % N


def test_mod_typing(U):
    e = U.Expr("col1 * col1 % 514679589732349351 * col1 % 514679589732349351", [("col1", 17, 0)])
    # the int64 literal becomes Decimal(int64_t), precision 19; MOD types as (P of the divisor, 0)
    assert (e.precision, e.scale, e.faithful) == (19, 0, True)
    assert e.plan.split()[:3] == ["F", "19", "0"]
    e = U.Expr("a % b", [("a", 38, 10), ("b", 20, 5)])
    assert (e.precision, e.scale) == (20, 0)
    assert e.plan.endswith("(MOD 20 0 c0:38:10 c1:20:5)")
    # literals past int64 are decimal literals of their digits
    e = U.Expr("a % 331972734823750661745273085756461571", [("a", 70, 0)])
    assert (e.precision, e.scale) == (36, 0)
    with pytest.raises(U.UpdecError, match="JIT path"):
        U.VmProgram("a % b", [("a", 38, 10), ("b", 20, 5)])


@pytest.fixture(scope="module")
def dmod_host(tmp_path_factory):
    import os
    import subprocess
    here = os.path.dirname(os.path.abspath(__file__))
    csrc = os.path.join(os.path.dirname(here), "ultraprecise_amd", "csrc")
    out = str(tmp_path_factory.mktemp("dmod") / "libdmod_host.so")
    subprocess.check_call(["g++", "-O1", "-std=c++17", "-fPIC", "-shared", "-I" + csrc,
                           os.path.join(here, "quad_host", "dmod_host.cpp"), "-o", out])
    return ctypes.CDLL(out)


@pytest.mark.parametrize("narrow", [0, 1])
@pytest.mark.parametrize("faith", [0, 1])
def test_device_dmod_matches_closed_form(dmod_host, narrow, faith):
    """dec_device.h's dmod, built for the host, against pyref.mod"""
    import numpy as np
    rng = random.Random(23 + 2 * narrow + faith)
    ps = []
    for a, b in pairs(rng):
        if narrow:  # a within 4 limbs, b within 8
            a = (a[0], a[1] & ((1 << 128) - 1), 0)
            b = (b[0], b[1] & ((1 << 256) - 1), 0)
        ps.append((a, b))
    n = len(ps)
    A = np.zeros((n, 32), dtype=np.uint32)
    B = np.zeros((n, 32), dtype=np.uint32)
    sg = np.zeros((n, 2), dtype=np.uint32)
    for i, (a, b) in enumerate(ps):
        for k in range(32):
            A[i, k] = (a[1] >> (32 * k)) & 0xFFFFFFFF
            B[i, k] = (b[1] >> (32 * k)) & 0xFFFFFFFF
        sg[i] = (a[0], b[0])
    out = np.zeros((n, 33), dtype=np.uint32)
    dmod_host.dmod_run(narrow, faith, A.ctypes.data_as(ctypes.c_void_p), B.ctypes.data_as(ctypes.c_void_p),
                       sg.ctypes.data_as(ctypes.c_void_p), n, out.ctypes.data_as(ctypes.c_void_p))
    for i, (a, b) in enumerate(ps):
        m = sum(int(out[i, k]) << (32 * k) for k in range(32))
        want = pyref.mod(a, b, faithful=bool(faith))
        assert (int(out[i, 32]), m) == (want[0], want[1]), (a, b)
