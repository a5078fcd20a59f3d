"""CPU: the planner's rewrite and typing against an independent restatement.

tests/ref_rewrite.py restates the reference's AlignOptimize family
(AriesEngineShell.cpp:90-685), the frontend result types
(CommonBiaodashi.cpp:3338-3377), the literal bridging (AriesExprBridge.cpp)
and the host Decimal constant folding (decimal.cu) from the reference's
source, and plans from the expression text.  The product plans the same
text through the C ABI (updec_expr_compile / updec_expr_plan); the plan
texts must be identical: evaluation order, every node's (P, S), folded
constants with their scales, and the materialized width B(P) of the root.

Both modes are compared.  Faithful mode is skipped where the product
refuses it (the expression leaves the reference's defined envelope), and
where the restatement meets a constant rescaled by more than 38 digits
(operator<< reads past __POW10_ARRAY: no defined result; the product uses
the exact power).  Cases both reject (integer operands left after the
rewrite, integer-typed calculations) must be rejected by both.
"""
import random

import pytest

import ref_rewrite as R
from ultraprecise_amd import workloads as W


def product_plan(U, text, cols, mode):
    try:
        e = U.Expr(text, cols, mode=mode)
    except U.UpdecError as ex:
        return None, str(ex)
    head, body = e.plan.split(" ", 4)[:4], e.plan.split(" ", 4)[4]
    assert (int(head[1]), int(head[2]), int(head[3])) == (e.precision, e.scale, e.out_bytes)
    return (body, e.precision, e.scale, e.out_bytes), None


def compare(U, O, text, cols, mode):
    got, err = product_plan(U, text, cols, mode)
    try:
        body, P, S = R.plan(text, cols, mode)
        want = (body, P, S, O.lib.od_compact_bytes(P))
        why = None
    except R.Undefined as ex:
        want, why = None, str(ex)
    if got is None and mode == R.FAITHFUL and "envelope" in err:
        return "envelope"
    if want is None and "POW10" in why:
        return "departure"
    if got is None and want is None:
        return "both reject"
    assert got == want, (mode, text, cols, got, err, want, why)
    return "same"


Q1_COLS = [("qty", 15, 2), ("price", 15, 2), ("disc", 15, 2), ("tax", 15, 2)]
SMALL = [("col1", 12, 3), ("col2", 20, 5)]

FIXED = [(t, [("p", 15, 2), ("q", 15, 2), ("d", 15, 2)]) for t in ("p*(1-d)", "q*p")]
FIXED += [(W.CONFIG2["expr"], [(n, P, S) for n in "abcd"]) for P, S in W.CONFIG2["precisions"]]
FIXED += [(t, Q1_COLS) for t in W.CONFIG3["exprs"]]
# append_expr_info/expr_query/2_optimizations_in_expressions/query.md shapes
FIXED += [(t, SMALL) for t in ("1+col1+2+11", "1+col1+2-3", "4*(col1+col2)*0.25", "col1+col2+col1",
                               "col1-col2*3+2.5", "-col1-col2", "col1/3", "col1%7", "2/col1", "col1*0.5*2",
                               "col1+0.001-0.001", "0.1986039 % col1", "col1 - 0.00 + col2", "-(col1 - 2.50)")]
FIXED += [("col1 * col1 % 514679589732349351 * col1 % 514679589732349351", [("col1", 17, 0)])]


@pytest.mark.parametrize("text,cols", FIXED)
def test_fixed_expressions(U, O, text, cols):
    for mode in (R.FAITHFUL, R.EXTENDED):
        compare(U, O, text, cols, mode)
    # every fixed expression plans in extended mode
    assert compare(U, O, text, cols, R.EXTENDED) == "same"


def test_series_plan(U, O):
    cols = [("x", 300, 100)]
    assert compare(U, O, W.sin_expr(), cols, R.EXTENDED) == "same"


def test_literal_leaf_type_is_string_to_decimal():
    """a decimal literal's own type counts its leading zero (StringToDecimal),
    while the frontend typing of its parent uses its real precision"""
    body, P, S = R.plan("c % 0.1986039", [("c", 12, 3)], R.FAITHFUL)
    assert body == "(MOD 8 0 c0:12:3 k+1e4df7:7)" and (P, S) == (8, 0)


def _lit(rng):
    k = rng.randrange(10)
    if k == 0:
        return str(rng.choice([0, 1, 2, 10]))
    if k == 1:
        return str(rng.randrange(1, 1000))
    if k == 2:
        return str(rng.randrange(1, 1 << 40))
    if k == 3:
        return rng.choice(["1.00", "1.0", "0.0", "0.00", "1", "2.5", "0.25", "10.000"])
    if k == 4:
        s, i = rng.randrange(1, 12), rng.randrange(0, 6)
        return (str(rng.randrange(0, 10 ** i)) if i else "0") + "." + "".join(rng.choice("0123456789") for _ in range(s))
    return None


def _tree(rng, names, depth):
    if depth == 0 or rng.random() < 0.25:
        lit = _lit(rng) if rng.random() < 0.3 else None
        t = lit if lit is not None else rng.choice(names)
        return "-" + t if rng.random() < 0.15 else t
    op = rng.choice("++--**//%")
    t = f"{_tree(rng, names, depth - 1)} {op} {_tree(rng, names, depth - 1)}"
    if rng.random() < 0.6:
        t = f"({t})"
    if rng.random() < 0.1:
        t = f"-{t}" if t.startswith("(") else f"-({t})"
    if rng.random() < 0.05:
        t = f"cast({t}, {rng.randrange(20, 60)}, {rng.randrange(0, 15)})"
    return t


@pytest.mark.parametrize("seed", range(4))
def test_random_trees(U, O, seed):
    """1,200 seeded random trees (4 x 300) with literals of every kind, unary
    minus, casts and all five operators, over 1..4 columns of random types"""
    rng = random.Random(1000 + seed)
    seen = {}
    for _ in range(300):
        maxp = rng.choice([20, 40, 80, 300])
        cols = []
        for k in range(rng.randrange(1, 5)):
            P = rng.randrange(1, maxp + 1)
            cols.append((f"c{k}x", P, rng.randrange(0, P + 1)))
        text = _tree(rng, [c[0] for c in cols], rng.randrange(1, 6))
        for mode in (R.FAITHFUL, R.EXTENDED):
            r = compare(U, O, text, cols, mode)
            seen[r] = seen.get(r, 0) + 1
    assert seen["same"] >= 450, seen


def test_serving_trees(U, O):
    """the serving workload's trees (depth 2..6 over 2..4 columns, every
    precision of the stream)"""
    for i in range(200):
        q = W.random_query(W.CONFIG5["seed"], i)
        assert compare(U, O, q["text"], q["cols"], R.EXTENDED) == "same"
        compare(U, O, q["text"], q["cols"], R.FAITHFUL)
