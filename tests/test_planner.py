"""CPU: the rewrite (AlignOptimize), result typing, classification, code
generation and the C ABI surface -- no device needed."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

C15 = [("q", 15, 2), ("p", 15, 2), ("d", 15, 2)]


def plan(U, text, cols, mode=0):
    return U.Expr(text, cols, mode=mode).plan


# Known answers derived by hand from the reference's rewrite rules
# (AriesEngineShell.cpp:90-685) and type rules (decimal.cu:1353-2000).
def test_plain_method_plans(U):
    # 1 - d: the integer folds into Decimal(1, scale 2) = 1.00 (:420-436), then
    # SUB node (16,2) (ListToBinTree :104-116); MUL (31,4) (:117-123); B(31) = 14
    assert plan(U, "p*(1-d)", C15) == "F 31 4 14 (MUL 31 4 c1:15:2 (SUB 16 2 k+64:2 c2:15:2))"
    assert plan(U, "q*p", C15) == "F 30 4 13 (MUL 30 4 c0:15:2 c1:15:2)"


def test_constant_folding_and_alignment_examples(U):
    # query.md "4*(col1+col2)*0.25": the constants fold to 1.00 and are dropped
    assert plan(U, "4*(q+p)*0.25", C15) == "F 16 2 7 (ADD 16 2 c0:15:2 c1:15:2)"
    # "1+col1+2-3": the integers cancel and the column replaces the node (:402-406)
    assert plan(U, "1+q+2-3", C15) == "F 15 2 7 c0:15:2"
    # "1+col1+2+11" with col1 DECIMAL(17,10): 14 folds into Decimal(14, 10)
    # placed first; ADD type max(0-10, 17-10)+1+10 = 18
    k = format(14 * 10 ** 10, "x")
    assert plan(U, "1+c+2+11", [("c", 17, 10)]) == f"F 18 10 8 (ADD 18 10 k+{k}:10 c0:17:10)"
    # alignment scheduling (query.md "col1+col2+col1"): operands sorted by
    # scale, so the two DECIMAL(2,1) operands are added before the (17,11) one
    assert plan(U, "a+b+a", [("a", 2, 1), ("b", 17, 11)]) == \
        "F 18 11 8 (ADD 18 11 (ADD 3 1 c0:2:1 c0:2:1) c1:17:11)"
    # decimal constants fold with the integer into one constant: 1.5 + 2 = 3.5
    # (scale 1), rescaled to 3.50 and inserted before the first operand of
    # larger scale (:437-470)
    assert plan(U, "1.5+q+2", C15) == "F 16 2 7 (ADD 16 2 k+15e:2 c0:15:2)"
    # a multiplication by constants that fold to 1.00 keeps only the column
    assert plan(U, "q*0.5*2", C15) == "F 15 2 7 c0:15:2"


def test_sweep_plans_and_modes(U):
    cols = [(n, 18, 4) for n in "abcd"]
    assert plan(U, "(a+b)*c - a/d", cols) == \
        "F 38 8 16 (ADD 38 8 (MUL 37 8 (ADD 19 4 c0:18:4 c1:18:4) c2:18:4) (DIV 9 8 (NEG c0:18:4) c3:18:4))"
    # at (38,10) the division (scale 14) sorts before the product (scale 20)
    cols = [(n, 38, 10) for n in "abcd"]
    assert plan(U, "(a+b)*c - a/d", cols).startswith("F 78 20 33 (ADD 78 20 (DIV 15 14 (NEG c0:38:10)")
    for P, S, faithful in [(76, 20, True), (150, 50, False), (300, 100, False)]:
        e = U.Expr("(a+b)*c - a/d", [(n, P, S) for n in "abcd"])
        assert e.faithful is faithful
    # faithful mode caps the root precision at 288 (AriesEngineShell.cpp:638-640)
    e = U.Expr("a*b", [("a", 200, 0), ("b", 200, 0)], mode=U.MODE_AUTO)
    assert e.precision == (288 if e.faithful else 400)
    with pytest.raises(U.UpdecError):
        U.Expr("(a+b)*c - a/d", [(n, 150, 50) for n in "abcd"], mode=U.MODE_FAITHFUL)


def test_parse_errors(U):
    with pytest.raises(U.UpdecError, match="unknown column"):
        U.Expr("x+1", C15)
    with pytest.raises(U.UpdecError):
        U.Expr("q+", C15)
    with pytest.raises(U.UpdecError, match="cast"):
        U.Expr("cast(q, 123456789012345678901234, 2)", C15)
    # MOD compiles on the JIT path (it has no interpreter instruction)
    assert U.Expr("q%p", C15).plan.endswith("(MOD 15 0 c0:15:2 c1:15:2)")


def test_rewritten_plan_parses_in_oracle(U, O):
    for text in ["p*(1-d)", "q*p", "(q+p)*(d-q)/p", "-q*p+d", "q-(p-d)", "q*p % d + 1"]:
        pl = O.Plan(U.Expr(text, C15).plan)
        assert pl.P > 0


@pytest.mark.parametrize("P,S", [(18, 4), (38, 10), (76, 20), (150, 50), (300, 100)])
def test_sweep_kernel_compiles(U, P, S):
    e = U.Expr("(a+b)*c - a/d", [(n, P, S) for n in "abcd"])
    U.check_compile(e.source)


def test_agg_kernels_compile(U):
    from ultraprecise_amd import workloads as W
    cols = [(c[0], c[1], c[2], b) for c, b in zip(W.CONFIG3["cols"], [13, 24, 4, 4])]
    a = U.Agg(W.CONFIG3["exprs"], cols, filter=W.CONFIG3["filter"], nkeys=2)
    src = a.source(6)
    assert "LDS 64-bit atomic slots" in src
    U.check_compile(src)
    # without statistics the accumulators are wider and use lane-private slots
    cols = [(c[0], c[1], c[2]) for c in W.CONFIG3["cols"]]
    src = U.Agg(W.CONFIG3["exprs"][:2], cols, nkeys=1).source(3)
    U.check_compile(src)
    src = U.Agg(W.CONFIG1["exprs"], C15).source(1)
    assert "registers" in src
    U.check_compile(src)


def test_fused_avg_kernels(U):
    """AVG expressions: the one-pass final kernel divides by the faithful or
    the extended rule as the mode says; masks past the expressions fail"""
    from ultraprecise_amd import workloads as W
    cols = [(c[0], c[1], c[2], b) for c, b in zip(W.CONFIG3["cols"], [13, 24, 4, 4])]
    for mode, faith in ((U.MODE_AUTO, 1), (U.MODE_FAITHFUL, 1), (U.MODE_EXTENDED, 0)):
        src = U.Agg(W.CONFIG3["exprs"], cols, filter=W.CONFIG3["filter"], nkeys=2, mode=mode, avg=(0, 1, 4)).source(0)
        assert f"div_dec_int64<{faith}, AW2>" in src
        assert "u8* __restrict__ avgs" in src
        U.check_compile(src)
    src = U.Agg(W.CONFIG3["exprs"], cols, filter=W.CONFIG3["filter"], nkeys=2).source(0)
    assert "div_dec_int64" not in src
    with pytest.raises(ValueError):
        U.Agg(["a"], [("a", 15, 2)], avg=[1])
    # the C entry point checks the mask itself
    arr, keep = U._col_array([("a", 15, 2)])
    ex = (ctypes.c_char_p * 1)(b"a")
    d = U.AggDesc(arr, 1, ex, 1, 0, 0, 0, 0, U.MODE_AUTO, 2)
    h = ctypes.c_void_p()
    with pytest.raises(U.UpdecError, match="avg_mask"):
        U._check(U.lib().updec_agg_compile(ctypes.byref(d), ctypes.byref(h)))


def test_c_abi_exports_every_declared_symbol(U):
    hdr = open(os.path.join(ROOT, "include", "updec.h")).read()
    declared = set(re.findall(r"\b(updec_[a-z0-9_]+)\s*\(", hdr))
    lib = U.lib()
    missing = [n for n in sorted(declared) if not hasattr(lib, n)]
    assert not missing, missing
    assert set(U.EXPORTED) <= declared


def test_generator_is_counter_based(U):
    from ultraprecise_amd import workloads as W
    d = W.gen_desc(U, 0, 1, W.CONFIG1["cols"][1])
    full = U.gen_compact_host(d, 1000)
    part = U.gen_compact_host(d, 100, row0=400)
    assert (full[400 * 7:500 * 7] == part).all()
    # price column within its declared range
    import numpy as np
    m = full.reshape(-1, 7).astype(np.uint64)
    vals = sum(m[:, i] << np.uint64(8 * i) for i in range(7))
    assert vals.min() >= 90000 and vals.max() <= 10495000
    dg = W.digits_desc(U, 1, 0, 38)
    x = U.gen_compact_host(dg, 50)
    for i in range(50):
        b = bytearray(x[i * 16:(i + 1) * 16])
        b[-1] &= 0x7F
        assert len(str(int.from_bytes(bytes(b), "little"))) == 38


def test_descriptor_layouts_match_header(U, tmp_path):
    """the ctypes mirrors of the C ABI's descriptor structs have the
    header's size and field offsets (built with the host C compiler)"""
    import shutil
    import subprocess
    if not shutil.which("gcc"):
        pytest.skip("no host C compiler")
    mirrors = {"updec_column_desc": U.ColumnDesc, "updec_column_blocks": U.ColumnBlocks,
               "updec_agg_desc": U.AggDesc, "updec_gen_desc": U.GenDesc, "updec_column_iter": U.ColumnIterDesc}
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "updec.h"', "int main(void) {"]
    for cname, py in mirrors.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'  printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    subprocess.check_call(["gcc", "-I", inc, str(src), "-o", str(exe)])
    got = {}
    for line in subprocess.check_output([str(exe)]).decode().splitlines():
        cname, key, val = line.split()
        got[(cname, key)] = int(val)
    for cname, py in mirrors.items():
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)



@pytest.mark.parametrize("text,why", [
    ("a*1", "rewrites to a constant or integer"),       # RemoveZeroNode makes x*1 integer 0 in the reference
    ("1*b", "rewrites to a constant or integer"),
    ("a*0", "rewrites to a constant or integer"),
    ("cast(7, 20, 3)", "integer operand"),               # an integer literal left as an operand
    ("a + cast(7, 20, 3)", "integer operand"),
    ("(1-2)%a", "integer-typed calculation"),            # RemoveZeroNode leaves the integer SUB without operands
])
def test_documented_rejections(U, text, why):
    with pytest.raises(U.UpdecError, match=why):
        U.Expr(text, [("a", 18, 4), ("b", 15, 2)])


def test_constants_are_kernel_data(U):
    """generated sources carry no constant values: expressions that differ
    only in literals of the same limb widths give one source text (one
    hipRTC compile); the values travel in the plan's constant block"""
    cols = [("p", 15, 2), ("d", 15, 2), ("t", 15, 2)]
    a = U.Expr("p*(1-d)*(1.25+t)", cols)
    b = U.Expr("p*(3-d)*(7.50+t)", cols)
    assert a.source == b.source
    assert a.plan != b.plan
    assert "0x7d" not in a.source and "kld(a.kc, " in a.source
    ga = U.Agg(["p*(1-d)", "d+0.05"], cols, nkeys=1).source(0)
    gb = U.Agg(["p*(2-d)", "d+0.07"], cols, nkeys=1).source(0)
    assert ga == gb
    # a literal too wide for the first one's limbs changes the source
    c = U.Expr("p*(12345678901-d)*(1.25+t)", cols)
    assert c.source != a.source
    U.check_compile(a.source)


def test_iterator_kernels_compile(U):
    """the column-iterator variants (NULL flags, index kinds, nullable
    output) of the calc kernel and of the grouped, filtered aggregation"""
    from ultraprecise_amd import workloads as W
    cols = [("a", 38, 10), ("b", 38, 10), ("c", 15, 2)]
    for text in ("a/b + c", "(a+b)*c - a/c", "a % b"):
        e = U.Expr(text, cols)
        U.check_compile(e.iter_source())
    q1 = [(c[0], c[1], c[2]) for c in W.CONFIG3["cols"]]
    a = U.Agg(W.CONFIG3["exprs"], q1, filter=W.CONFIG3["filter"], nkeys=2)
    src = a.source(6, iterators=True)
    assert "col_load<7>(a, 1, row, nc1)" in src and "p2_nn" in src
    U.check_compile(src)


def test_multi_lane_calc_kernel_compiles(U):
    """an extended-mode plan of 64+ limbs has the multi-lane calc form at
    W/4 and W/8 lanes per row; both sources compile for gfx950, and lane
    counts other than 1, W/4 and W/8 are refused"""
    e = U.Expr("(a+b)*c - a/d", [(n, 300, 100) for n in "abcd"])
    assert e.multi_lane_width == 64
    for lanes in (16, 8):
        src = e.ml_source(lanes)
        assert f"#define TPI_ {lanes}" in src and "ml::div<L_, TPI_>(A, B, 104u, g)" in src
        U.check_compile(src)
    with pytest.raises(U.UpdecError):
        e.set_lanes(32)
    e.set_lanes(8)
    e.set_lanes(1)
    narrow = U.Expr("(a+b)*c - a/d", [(n, 38, 10) for n in "abcd"])
    assert narrow.multi_lane_width == 0
    with pytest.raises(U.UpdecError):
        narrow.set_lanes(8)
