"""GPU parity of truncate(x, p) (Decimal::truncate, decimal.cu:541-570)
through the generated calc kernel, bit-exact against the oracle: p > 0
(truncation and padding to scale p), p = 0, p < 0 (integer part with -p
digits zeroed), and p at or past the type's integer digits (zero, the sign
kept), over column and expression arguments, faithful and extended."""
import numpy as np
import pytest

from test_gpu_sweep import run_calc

pytestmark = pytest.mark.gpu

ROWS = 20_000
CASES = [
    ("truncate(a, 2)", [("a", 20, 6)]),
    ("truncate(a, 9)", [("a", 20, 6)]),
    ("truncate(a, 0)", [("a", 20, 6)]),
    ("truncate(a, -3)", [("a", 20, 6)]),
    ("truncate(a, -14)", [("a", 20, 6)]),          # intg 14: zero, sign kept
    ("truncate(-a, -20)", [("a", 20, 6)]),
    ("truncate(a*b, -5) + b", [("a", 20, 6), ("b", 18, 4)]),
    ("truncate(a/b, 3) - truncate(a, 40)", [("a", 20, 6), ("b", 18, 4)]),
    ("truncate(a, -30) + truncate(a, 120)", [("a", 150, 50)]),   # extended: a 70-digit shift
]


@pytest.mark.parametrize("k", range(len(CASES)))
def test_truncate_rows_bit_exact(U, O, gpu, k):
    import random
    text, cols = CASES[k]
    print(f"truncate case {k}: {text}", flush=True)
    rng = random.Random(77 + k)
    host = []
    for (_, P, S) in cols:
        nb = U.compact_bytes(P)
        buf = bytearray()
        for i in range(ROWS):
            d = rng.choice([P, P - 1, P // 2, 3, 1])
            m = rng.randrange(10 ** (d - 1), 10 ** d) if i % 11 else 0
            b = bytearray(m.to_bytes(nb, "little"))
            if rng.random() < 0.5:
                b[-1] |= 0x80
            buf += b
        host.append(np.frombuffer(bytes(buf), dtype=np.uint8).copy())
    e, got, want, full = run_calc(U, O, text, cols, host, ROWS)
    bad = np.nonzero((got != want).reshape(ROWS, -1).any(axis=1))[0]
    assert bad.size == 0, f"{text}: {bad.size} rows differ, first {bad[:5]}"
    assert not full[ROWS * e.out_bytes:].any()
