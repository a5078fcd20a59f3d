"""CPU: the column-iterator kernels (AriesColumnDataIterator inputs) that the
C ABI generates only at first use, generated here by the planner built for
the host and compiled for gfx950 with hipRTC; the row mapping they contain."""
import ctypes
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "ultraprecise_amd", "csrc")


@pytest.fixture(scope="module")
def gen(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("gen") / "libgen_host.so")
    subprocess.check_call(["g++", "-O1", "-std=c++17", "-fPIC", "-shared", "-I" + CSRC,
                           os.path.join(HERE, "vm_host", "gen_host.cpp"), os.path.join(CSRC, "planner.cpp"), "-o", out])
    lib = ctypes.CDLL(out)
    lib.gen_blocks_source.restype = ctypes.c_char_p

    def source(text, cols, agg):
        n = len(cols)
        names = (ctypes.c_char_p * n)(*[c[0].encode() for c in cols])
        P = (ctypes.c_int * n)(*[c[1] for c in cols])
        S = (ctypes.c_int * n)(*[c[2] for c in cols])
        return lib.gen_blocks_source(text.encode(), names, P, S, n, agg).decode()
    return source


CASES = [("p*(1-d)", [("p", 15, 2), ("d", 15, 2)]),
         ("(a+b)*c - a/d", [(n, 38, 10) for n in "abcd"]),
         ("col1 * col1 % 514679589732349351 * col1 % 514679589732349351", [("col1", 17, 0)]),
         ("(a+b)*c - a/d", [(n, 150, 50) for n in "abcd"])]


@pytest.mark.parametrize("text,cols", CASES)
@pytest.mark.parametrize("agg", [0, 1])
def test_block_kernels_compile(U, gen, text, cols, agg):
    src = gen(text, cols, agg)
    assert not src.startswith("error"), src
    assert "IterArgs a" in src and "col_load<" in src
    # every column load goes through its iterator and reports NULL
    assert "const ColIter& t = a.it[c];" in src and ", row, nc" in src
    assert "load_quad" not in src and "stage_wave" not in src
    U.check_compile(src)
