"""Seeded synthetic workloads shared by tests and bench.py.

Each column is drawn by the library's counter-based generator keyed by
(seed, column id), so host and device produce identical bytes.
"""
GEN_RANGE, GEN_DIGITS, GEN_BELOW = 0, 1, 2

# DECIMAL(15,2) columns of the plain-method workload (1,000,000 rows, seed 0)
# and of the grouped aggregation (60,000,000 rows, seed 2): magnitudes in
# hundredths.
CONFIG1 = {
    "rows": 1_000_000, "seed": 0,
    "cols": [
        # name, P, S, kind, lo, hi, step
        ("q", 15, 2, GEN_RANGE, 1, 50, 100),          # integer 1..50
        ("p", 15, 2, GEN_RANGE, 90000, 10495000, 1),  # 900.00 .. 104950.00
        ("d", 15, 2, GEN_RANGE, 0, 10, 1),            # 0.00 .. 0.10
    ],
    "exprs": ["p*(1-d)", "q*p"],
}

CONFIG3 = {
    "rows": 60_000_000, "seed": 2,
    "cols": [
        ("qty", 15, 2, GEN_RANGE, 1, 50, 100),
        ("price", 15, 2, GEN_RANGE, 90000, 10495000, 1),
        ("disc", 15, 2, GEN_RANGE, 0, 10, 1),
        ("tax", 15, 2, GEN_RANGE, 0, 8, 1),
    ],
    "flag1": b"ANR",
    "flag2": b"FO",
    "ship": (0, 2525),
    "filter": ("<=", 2435),
    # SUM(qty), SUM(price), SUM(price*(1-disc)), SUM(price*(1-disc)*(1+tax)), SUM(disc) (for AVG(disc))
    "exprs": ["qty", "price", "price*(1-disc)", "price*(1-disc)*(1+tax)", "disc"],
}

# precision sweep: 10,000,000 rows, seed 1, (a+b)*c - a/d
CONFIG2 = {
    "rows": 10_000_000, "seed": 1,
    "precisions": [(18, 4), (38, 10), (76, 20), (150, 50), (300, 100)],
    "expr": "(a+b)*c - a/d",
}

# division-heavy high-precision series: 1,000,000 rows, seed 3, x DECIMAL(300,100)
# uniform in [-pi, pi]; sin(x) by a 25-term alternating Taylor series in Horner
# form with the factorial reciprocals as scale-100 constants and every product
# truncated back to scale 100, then SUM over all rows
CONFIG4 = {"rows": 1_000_000, "seed": 3, "P": 300, "S": 100, "terms": 25}


def pi_scaled(S):
    """floor(pi * 10^S), from Machin's formula pi = 16 atan(1/5) - 4 atan(1/239)
    in fixed point with 20 guard digits."""
    one = 10 ** (S + 20)

    def atan_inv(n):
        total, term, k, sign = 0, one // n, 1, 1
        while term:
            total += sign * (term // k)
            term //= n * n
            k += 2
            sign = -sign
        return total

    return (16 * atan_inv(5) - 4 * atan_inv(239)) // 10 ** 20


def sin_constants(terms, S):
    """1/(2k+1)! truncated to S fractional digits, as decimal literals"""
    out, f = [], 1
    for k in range(terms):
        if k:
            f *= (2 * k) * (2 * k + 1)
        m = 10 ** S // f
        out.append(f"{m // 10 ** S}.{m % 10 ** S:0{S}d}")
    return out


def sin_expr(col="x", terms=25, P=300, S=100):
    """Horner form of the series: s = c_k - cast(y * s), y = cast(x*x), result cast(x * s)"""
    c = sin_constants(terms, S)
    y = f"cast({col}*{col}, {P}, {S})"
    s = c[terms - 1]
    for k in range(terms - 2, -1, -1):
        s = f"{c[k]} - cast({y}*({s}), {P}, {S})"
    return f"cast({col}*({s}), {P}, {S})"


def sin_desc(U, seed, col_id=0, P=300, S=100):
    return U.gen_desc(GEN_BELOW, P, seed, col_id, bound=pi_scaled(S) + 1, neg_permille=500)


# batched serving: 10,000 queries, seed 4, Poisson arrivals at 200/400/800
# queries/s; rows log-uniform in [1e3, 1e7]; P from the list with S = P/3; a
# random + - * / tree of depth 2..6 over 2..4 fresh uniform-digit columns,
# then SUM.  Every query is served as drawn: trees whose exact
# intermediates need more than 32 limbs run on the interpreter's multi-lane
# classes (up to 512 limbs, which holds all 10,000 queries of seed 4); an
# `accept` callback may still redraw (a test restricting the widths).
CONFIG5 = {"queries": 10_000, "seed": 4, "rates": (200, 400, 800), "rows": (1e3, 1e7),
           "precisions": (15, 18, 38, 76, 150, 300), "depth": (2, 6), "ncols": (2, 4)}


def random_tree(rng, names, depth):
    """fully parenthesized random tree of exactly `depth` operator levels"""
    if depth == 0:
        return rng.choice(names)
    op = rng.choice("+-*/")
    deep = gen_side = rng.random() < 0.5
    left = random_tree(rng, names, depth - 1 if deep else rng.randrange(0, depth))
    right = random_tree(rng, names, rng.randrange(0, depth) if deep else depth - 1)
    return f"({left}{op}{right})"


def random_query(seed, i, accept=None, config=None):
    """Query i of the serving stream: {"text", "cols", "rows", "P", "S",
    "col_ids", "tries"}.  `accept(text, cols)` -> bool (optional) decides
    whether a drawn tree is kept; without it the first tree is the query."""
    import math
    import random
    C = config or CONFIG5
    rng = random.Random(seed * 1_000_003 + i)
    lo, hi = C["rows"]
    rows = int(round(10 ** rng.uniform(math.log10(lo), math.log10(hi))))
    P = rng.choice(C["precisions"])
    S = P // 3
    ncols = rng.randint(*C["ncols"])
    names = [chr(ord("a") + k) for k in range(ncols)]
    cols = [(n, P, S) for n in names]
    tries = 0
    while True:
        tries += 1
        depth = rng.randint(*C["depth"])
        text = random_tree(rng, names, depth)
        if accept is None or accept(text, cols):
            break
    # resident columns used by the serving loop: distinct ids out of
    # SERVE_COLUMNS_PER_P per precision
    keys = [(P, c) for c in rng.sample(range(SERVE_COLUMNS_PER_P), ncols)]
    return {"text": text, "cols": cols, "rows": rows, "P": P, "S": S, "keys": keys,
            "col_ids": [1000 + 16 * i + k for k in range(ncols)], "tries": tries}


SERVE_COLUMNS_PER_P = 6


def query_column_desc(U, seed, q, k):
    """generator of the k-th column of a serving query: P uniform digits,
    sign 50/50 (never zero, so divisions are defined)"""
    return U.gen_desc(GEN_DIGITS, q["P"], seed, q["col_ids"][k], digits=q["P"], neg_permille=500)


# column ids of the generator streams
COL_FLAG1, COL_FLAG2, COL_SHIP = 100, 101, 102


def gen_desc(U, seed, col_id, spec):
    name, P, S, kind, lo, hi, step = spec
    return U.gen_desc(kind, P, seed, col_id, lo=lo, hi=hi, step=step)


def digits_desc(U, seed, col_id, P, neg_permille=500):
    return U.gen_desc(GEN_DIGITS, P, seed, col_id, digits=P, neg_permille=neg_permille)
