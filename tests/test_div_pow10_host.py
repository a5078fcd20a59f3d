"""CPU: truncating division by a compile-time 10^N (dshr10, the scale-down
cast and the truncation of products) built for the host from dec_device.h:
the Barrett form and the one-word pass form against exact integer division,
on random values of every bit length, the values next to multiples of 5^N
and 10^N, and the extremes, for every instantiated (limbs, N)."""
import ctypes
import os
import random
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "ultraprecise_amd", "csrc")


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("div5") / "libdiv5_host.so")
    subprocess.check_call(["g++", "-O1", "-std=c++17", "-fPIC", "-shared", "-I" + CSRC,
                           os.path.join(HERE, "quad_host", "div5_host.cpp"), "-o", out])
    return ctypes.CDLL(out)


def cases(lib):
    out = []
    for c in range(lib.div5_ncases()):
        wi, n, b = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        lib.div5_case(c, ctypes.byref(wi), ctypes.byref(n), ctypes.byref(b))
        out.append((c, wi.value, n.value, bool(b.value)))
    return out


def values(wi, n, rng):
    top = 1 << (32 * wi)
    d5, d10 = 5 ** n, 10 ** n
    xs = [0, 1, top - 1, top // 2, d10 - 1 if d10 < top else 7, (d10 if d10 < top else 5)]
    for m in (d5 << n, d10):
        for k in (1, 2, 3, rng.randrange(1, max(2, top // m)) if m < top else 1):
            for e in (-2, -1, 0, 1):
                x = k * m + e
                if 0 <= x < top:
                    xs.append(x)
    for _ in range(400):
        bits = rng.randrange(1, 32 * wi + 1)
        xs.append(rng.getrandbits(bits))
    # x with x >> N = k 5^N 2^S - 1 (2^S <= 5^N < 2^(S+1)) near the top: all
    # of floor(x / 10^N)'s fractional part and of the shifted-out bits, where
    # a Barrett estimate falls short
    S = d5.bit_length() - 1
    kmax = (top >> n) // (d5 << S)
    for k in range(max(1, kmax - 20), kmax + 1):
        xs.append((((k * (d5 << S)) - 1) << n) | ((1 << n) - 1))
    # values just below the top
    for _ in range(50):
        xs.append(top - 1 - rng.getrandbits(rng.randrange(1, min(64, 32 * wi))))
    return xs


def run(lib, c, wi, xs, passes):
    X = np.zeros((len(xs), wi), dtype=np.uint32)
    for i, x in enumerate(xs):
        for k in range(wi):
            X[i, k] = (x >> (32 * k)) & 0xFFFFFFFF
    out = np.zeros_like(X)
    lib.div5_run(c, X.ctypes.data_as(ctypes.c_void_p), len(xs), out.ctypes.data_as(ctypes.c_void_p), int(passes))
    return [sum(int(out[i, k]) << (32 * k) for k in range(wi)) for i in range(len(xs))]


def test_division_by_10n_is_exact(lib):
    rng = random.Random(5)
    seen_barrett = seen_passes = 0
    for c, wi, n, barrett in cases(lib):
        xs = values(wi, n, rng)
        want = [x // 10 ** n for x in xs]
        got = run(lib, c, wi, xs, passes=False)
        bad = [(x, g, w) for x, g, w in zip(xs, got, want) if g != w]
        assert not bad, (wi, n, barrett, bad[:3])
        assert run(lib, c, wi, xs, passes=True) == want, (wi, n, "passes")
        seen_barrett += barrett
        seen_passes += not barrett
    # both forms are chosen somewhere (the series' truncation by 10^100 is Barrett)
    assert seen_barrett >= 10 and seen_passes >= 3
    assert any(wi == 21 and n == 100 and b for _, wi, n, b in cases(lib))


def test_reciprocal_word(tmp_path):
    """recip_word (a double quotient corrected by one integer step) equals
    floor((2^64 - 1) / d) - 2^32 for every normalized d in a stride-13 sweep
    and for every d of the first and last 2^20"""
    out = str(tmp_path / "librecip.so")
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-I" + CSRC,
                           os.path.join(HERE, "quad_host", "div5_host.cpp"), "-o", out])
    lib = ctypes.CDLL(out)
    lib.recip_check.restype = ctypes.c_long
    assert lib.recip_check(0, 13) == 0
    assert lib.recip_check(5, 13) == 0
    assert lib.recip_check(0, 1 << 11) == 0
    # the ends of the range: d = 2^31 .. 2^31 + 2^20 and 2^32 - 2^20 .. 2^32 - 1
    assert lib.recip_check_range(0, 1 << 20) == 0
    assert lib.recip_check_range((1 << 31) - (1 << 20), 1 << 20) == 0
