"""CPU: the interpreter's per-row code (vm_device.h) built for the host and
checked row by row against the oracle -- the configuration expressions,
casts, and a seeded stream of random serving queries of every width class.
The GPU tests check the same programs on the device."""
import ctypes
import os
import subprocess

import pytest

from ultraprecise_amd import workloads as W

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "ultraprecise_amd", "csrc")


@pytest.fixture(scope="session")
def vmh(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("vmh") / "libvm_host.so")
    subprocess.check_call(["g++", "-O1", "-std=c++17", "-fPIC", "-shared", "-I" + CSRC,
                           os.path.join(HERE, "vm_host", "vm_host.cpp"), os.path.join(CSRC, "planner.cpp"), "-o", out])
    lib = ctypes.CDLL(out)

    def run(text, cols, data, rows, mode=0):
        n = len(cols)
        names = (ctypes.c_char_p * n)(*[c[0].encode() for c in cols])
        P = (ctypes.c_int * n)(*[c[1] for c in cols])
        S = (ctypes.c_int * n)(*[c[2] for c in cols])
        B = (ctypes.c_int * n)(*[c[3] if len(c) > 3 else 0 for c in cols])
        ptrs = (ctypes.c_void_p * n)(*[d.ctypes.data for d in data])
        img = ctypes.create_string_buffer(133)
        err = ctypes.create_string_buffer(256)
        rc = lib.vmh_sum(text.encode(), names, P, S, B, n, mode, ptrs, ctypes.c_int64(rows), img, err, 256)
        if rc:
            raise RuntimeError(err.value.decode())
        return img.raw
    return run


def oracle_sum(U, O, text, cols, data, rows):
    e = U.Expr(text, cols)
    mode = O.FAITHFUL if e.faithful else O.EXTENDED
    want, n = O.plan_sum(O.Plan(e.plan), data, [c[1] for c in cols], [c[2] for c in cols], rows, mode=mode)
    assert n == rows
    return want


C15 = [("q", 15, 2), ("p", 15, 2), ("d", 15, 2), ("t", 15, 2)]
CASES = [
    ("p*(1-d)", C15), ("q*p", C15), ("p*(1-d)*(1+t)", C15), ("q", C15),
    ("(a+b)*c - a/d", [(n, 18, 4) for n in "abcd"]),
    ("(a+b)*c - a/d", [(n, 38, 10) for n in "abcd"]),
    ("(a+b)*c - a/d", [(n, 76, 20) for n in "abcd"]),
    ("(a+b)*c - a/d", [(n, 150, 50) for n in "abcd"]),
    ("a/b + cast(c, 40, 3) - cast(a*b, 60, 30)", [(n, 38, 10) for n in "abc"]),
    ("1.5 - a*2 + 0.25/b", [(n, 18, 4) for n in "ab"]),
    ("cast(a, 10, 0) - cast(b, 60, 45)", [(n, 38, 20) for n in "ab"]),
    # truncate(p): to a scale, to the integer part with -p digits zeroed,
    # and to zero when the type has no more than -p integer digits
    ("truncate(a, 2) + truncate(b, -3) - truncate(a*b, 0)", [(n, 20, 6) for n in "ab"]),
    ("truncate(a, -14) + truncate(-b, -20) * truncate(a, 9)", [(n, 20, 6) for n in "ab"]),
]


@pytest.mark.parametrize("k", range(len(CASES)))
def test_interpreter_rows_match_oracle(U, O, vmh, k):
    text, cols = CASES[k]
    rows = 3000
    data = [U.gen_compact_host(U.gen_desc(U.GEN_DIGITS, P, 31 + k, 500 + i, digits=max(1, P - (i % 3)),
                                          neg_permille=500), rows) for i, (_, P, S) in enumerate(cols)]
    assert vmh(text, cols, data, rows) == oracle_sum(U, O, text, cols, data, rows)


def test_interpreter_series(U, O, vmh):
    rows = 200
    host = U.gen_compact_host(W.sin_desc(U, 3), rows)
    cols = [("x", 300, 100, 334)]
    assert vmh(W.sin_expr(), cols, [host], rows) == oracle_sum(U, O, W.sin_expr(), cols, [host], rows)


def test_interpreter_random_serving_queries(U, O, vmh):
    def ok(text, cols):
        try:
            U.VmProgram(text, cols)
            return True
        except U.UpdecError as e:
            if "exceed" in str(e) or "JIT path only" in str(e):
                return False
            raise

    cfg = dict(W.CONFIG5, rows=(1e3, 2000))
    widths = set()
    for i in range(150):
        q = W.random_query(21, i, ok, config=cfg)
        rows = q["rows"]
        data = [U.gen_compact_host(U.gen_desc(U.GEN_DIGITS, q["P"], 21, 5000 + 16 * q["P"] + cid, digits=q["P"],
                                              neg_permille=500), rows) for _, cid in q["keys"]]
        assert vmh(q["text"], q["cols"], data, rows) == oracle_sum(U, O, q["text"], q["cols"], data, rows), q
        widths.add(U.VmProgram(q["text"], q["cols"]).width)
    # every one-value-per-lane class from 4 limbs and the first multi-lane
    # class (its programs evaluated here one value per thread)
    assert {4, 8, 16, 32, 64} <= widths
