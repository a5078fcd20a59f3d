"""GPU parity on seeded random expressions: mixed column precisions and
scales, integer and decimal literals, unary minus, + - * / %, cast, faithful
and extended plans, over columns that hold zeros, small and full-width
values.  Every output byte of the generated kernel must equal the oracle's.
Each case is one expression, so a failure names the text and the seed."""
import random

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROWS = 3000
# (P, S) of the columns a, b, c; b holds small values (zeros included)
SPECS = [(9, 2), (15, 2), (18, 4), (30, 10), (38, 5), (44, 12), (60, 20), (76, 30), (120, 40), (150, 50)]
LITERALS = ["0", "1", "7", "100", "0.25", "3.5", "0.001", "12345678901", "99999999999999999999", "2.71828"]
NCASES = 64


def random_expr(rng, depth, top=False):
    if depth == 0 or (not top and rng.random() < 0.2):
        r = rng.random()
        if r < 0.75:
            return rng.choice("abc")
        return rng.choice(LITERALS)
    r = rng.random()
    if r < 0.08:
        return f"(-{random_expr(rng, depth - 1)})"
    if r < 0.14:
        P = rng.choice([20, 30, 40, 60])
        return f"cast({random_expr(rng, depth - 1)}, {P}, {rng.randint(0, min(P, 12))})"
    op = rng.choice(["+", "-", "*", "*", "/", "%"])
    return f"({random_expr(rng, depth - 1)}{op}{random_expr(rng, depth - 1)})"


# the planner's documented rejections (tests/test_planner.py checks them on
# the CPU): a tree that rewrites to a constant or an integer (x*1, x*0), an
# integer operand left in a decimal operation (a cast of a literal), an
# integer-typed calculation below a decimal one, exact values past the
# widest limb class, or a grouped kernel past the LDS budget
DOCUMENTED = ("constant", "integer operand", "integer-typed", "wide", "limb", "digits", "LDS")


def case(U, k, grouped=False):
    """case k: the first draw of its seeded stream that the planner accepts
    (a rejected draw, which must be a documented rejection, is redrawn from
    the same stream, so every case runs)"""
    rng = random.Random(1009 * k + 17)
    while True:
        specs = [rng.choice(SPECS) for _ in range(3)]
        text = random_expr(rng, rng.randint(2, 4), top=True)
        while not any(ch in text for ch in "abc"):
            text = random_expr(rng, rng.randint(2, 4), top=True)
        cols = [(n, P, S) for n, (P, S) in zip("abc", specs)]
        try:
            U.Expr(text, cols)
            if grouped:
                U.Agg([text], cols, nkeys=1, avg=[0])
            return specs, text
        except U.UpdecError as err:
            assert any(w in str(err) for w in DOCUMENTED), str(err)


def columns(U, specs, seed):
    dev, host = [], []
    for i, (P, S) in enumerate(specs):
        if i == 1:
            d = U.gen_desc(U.GEN_RANGE, P, seed, i, lo=-4, hi=4, step=1)
        else:
            d = U.gen_desc(U.GEN_DIGITS, P, seed, i, digits=P if i == 0 else max(1, P // 3), neg_permille=400)
        t = U.gen_compact_device(d, ROWS)
        dev.append(t)
        host.append(t.numpy(count=ROWS * U.compact_bytes(P)))
    return dev, host


@pytest.mark.parametrize("k", range(NCASES))
def test_random_expression_bit_exact(U, O, gpu, k):
    specs, text = case(U, k)
    cols = [(n, P, S) for n, (P, S) in zip("abc", specs)]
    print(f"case {k}: {text} over {specs}", flush=True)
    e = U.Expr(text, cols)
    dev, host = columns(U, specs, 300 + k)
    out = U.DeviceBuffer(ROWS * e.out_bytes + 4).zero()
    e.calc(dev, ROWS, out)
    full = out.numpy()
    got = full[:ROWS * e.out_bytes]
    plan = O.Plan(e.plan)
    want = O.plan_eval(plan, host, [P for P, S in specs], [S for P, S in specs], ROWS, e.out_bytes,
                       mode=O.FAITHFUL if e.faithful else O.EXTENDED)
    bad = np.nonzero((got != want).reshape(ROWS, e.out_bytes).any(axis=1))[0]
    assert bad.size == 0, f"{text} ({'faithful' if e.faithful else 'extended'}): {bad.size} rows differ, first {bad[:5]}"
    assert not full[ROWS * e.out_bytes:].any()


@pytest.mark.parametrize("k", range(0, NCASES, 3))
def test_random_expression_grouped_sum_avg(U, O, gpu, k):
    """SUM and AVG of the same random expressions, grouped by a 1-byte key"""
    specs, text = case(U, k, grouped=True)
    cols = [(n, P, S) for n, (P, S) in zip("abc", specs)]
    print(f"case {k}: {text} over {specs}", flush=True)
    e = U.Expr(text, cols)
    agg = U.Agg([text], cols, nkeys=1, avg=[0])
    dev, host = columns(U, specs, 500 + k)
    key = U.gen_byte_device(500 + k, 9, bytes([4, 8, 15, 16, 23]), ROWS)
    hkey = key.numpy(count=ROWS).astype(np.int32)
    maxg = 8
    images = U.DeviceBuffer(maxg * 2 * 133).zero()
    counts = U.DeviceBuffer(maxg * 8).zero()
    ng, keys = agg.run(dev, ROWS, images, counts, key0=key, max_groups=maxg)
    hkeys = np.unique(hkey)
    assert ng == len(hkeys) and keys == [int(x) for x in hkeys]
    lut = np.full(256, -1, dtype=np.int32)
    lut[hkeys] = np.arange(ng, dtype=np.int32)
    mode = O.FAITHFUL if e.faithful else O.EXTENDED
    want, wc = O.plan_group_sum(O.Plan(e.plan), host, [P for P, S in specs], [S for P, S in specs], ROWS,
                                lut[hkey].astype(np.int32), ng, mode=mode)
    allimg = images.numpy()
    got = allimg[:maxg * 133].reshape(maxg, 133)
    avg = allimg[maxg * 133:].reshape(maxg, 133)
    gotc = counts.numpy("int64")[:ng]
    assert np.array_equal(gotc, wc)
    for g in range(ng):
        assert bytes(got[g]) == bytes(want[g]), f"SUM group {g}"
        # the aggregation's mode is AUTO: its AVG divides by the faithful rule
        assert bytes(avg[g]) == O.avg_from_sum(bytes(want[g]), int(wc[g])), f"AVG group {g}"
