"""Seeded synthetic workloads shared by tests and bench.py.

Each column is drawn by the library's counter-based generator keyed by
(seed, column id), so host and device produce identical bytes.
"""
GEN_RANGE, GEN_DIGITS = 0, 1

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

# column ids of the generator streams
COL_FLAG1, COL_FLAG2, COL_SHIP = 100, 101, 102


def gen_desc(U, seed, col_id, spec):
    name, P, S, kind, lo, hi, step = spec
    return U.gen_desc(kind, P, seed, col_id, lo=lo, hi=hi, step=step)


def digits_desc(U, seed, col_id, P, neg_permille=500):
    return U.gen_desc(GEN_DIGITS, P, seed, col_id, digits=P, neg_permille=neg_permille)
