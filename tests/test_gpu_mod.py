"""GPU parity of MOD (ModCalc) through the C ABI, bit-exact against the
oracle's literal restatement: the RSA workload's expressions on the four key
lengths, operand pairs built to hit every path of the reference's %, the
fused SUM of a MOD expression, and extended mode."""
import random

import numpy as np
import pytest

from test_gpu_sweep import encode, run_calc
from ultraprecise_amd import workloads as W

pytestmark = pytest.mark.gpu

RSA = [
    (17, "514679589732349351"),
    (35, "331972734823750661745273085756461571"),
    (71, "273420351015258185841522860354335753665897876941082396543621156533322691"),
    (143, "1686878444203618142189652405784181065581460450325437359510561733117064525110066314054980898011446"
          "29119144418134762919565200448008667563331468801"),
]


@pytest.mark.parametrize("P,N", RSA, ids=["LEN_4", "LEN_8", "LEN_16", "LEN_32"])
def test_rsa_rows_bit_exact(U, O, gpu, P, N):
    rows = 20_000
    col = U.gen_compact_device(W.digits_desc(U, 8, 0, P, neg_permille=0), rows)
    host = col.numpy(count=rows * U.compact_bytes(P))
    text = f"col1 * col1 % {N} * col1 % {N}"
    cols = [("col1", P, 0)]
    e, got, want, full = run_calc(U, O, text, cols, [host], rows)
    assert e.faithful
    bad = np.nonzero((got != want).reshape(rows, -1).any(axis=1))[0]
    assert bad.size == 0, f"{bad.size} rows differ, first {bad[0]}"
    assert not full[rows * e.out_bytes:].any()
    if P < 143:
        # exact modular arithmetic: the result column holds x^3 mod N
        xs = [int.from_bytes(bytes(host[i * U.compact_bytes(P):(i + 1) * U.compact_bytes(P)]), "little") for i in range(50)]
        ob = e.out_bytes
        for i, x in enumerate(xs):
            assert int.from_bytes(bytes(got[i * ob:(i + 1) * ob]), "little") == x * x % int(N) * x % int(N)


def _mod_pairs(rng, rows, Pa, Pb):
    """signed integers for a and b: bit lengths equal, one apart, far apart,
    divisors below 2^32 and 2^64, zeros"""
    a, b = [], []
    amax = (10 ** Pa - 1).bit_length() - 1
    bmax = (10 ** Pb - 1).bit_length() - 1
    for i in range(rows):
        k = i % 8
        lb = rng.randint(1, bmax)
        if k == 0:
            la = lb
        elif k == 1:
            la = max(1, lb - 1)
        elif k == 2:
            la = max(1, lb - rng.randint(2, 40))
        elif k == 3:
            la = min(amax, lb + rng.randint(1, 80))
        elif k == 4:
            lb = min(lb, 32)
            la = rng.randint(1, amax)
        elif k == 5:
            lb = min(lb, 64)
            la = rng.randint(1, min(amax, 64))
        else:
            la = rng.randint(1, amax)
        la = min(la, amax)
        x = rng.getrandbits(la) | (1 << (la - 1))
        y = rng.getrandbits(lb) | (1 << (lb - 1))
        if i % 97 == 0:
            x = 0
        if i % 89 == 0:
            y = 0
        a.append(x if rng.random() < 0.5 else -x)
        b.append(y if rng.random() < 0.5 else -y)
    return a, b


@pytest.mark.parametrize("Pa,Sa,Pb,Sb", [(38, 0, 38, 0), (60, 4, 40, 2), (18, 2, 9, 0), (120, 10, 80, 30), (18, 2, 60, 0)])
@pytest.mark.parametrize("mode", ["auto", "extended"])
def test_mod_paths_bit_exact(U, O, gpu, Pa, Sa, Pb, Sb, mode):
    rng = random.Random(Pa * 100 + Pb)
    rows = 40_000
    a, b = _mod_pairs(rng, rows, Pa, Pb)
    cols = [("a", Pa, Sa), ("b", Pb, Sb)]
    m = U.MODE_AUTO if mode == "auto" else U.MODE_EXTENDED
    e, got, want, _ = run_calc(U, O, "a % b", cols, [encode(a, Pa), encode(b, Pb)], rows, mode=m)
    bad = np.nonzero((got != want).reshape(rows, -1).any(axis=1))[0]
    assert bad.size == 0, f"{bad.size} rows differ; first a={a[bad[0]]} b={b[bad[0]]}"


def test_mod_inside_expression_and_sum(U, O, gpu):
    rng = random.Random(9)
    rows = 30_000
    a, b = _mod_pairs(rng, rows, 30, 20)
    c = [rng.randrange(-10 ** 10, 10 ** 10) for _ in range(rows)]
    cols = [("a", 30, 3), ("b", 20, 1), ("c", 12, 2)]
    host = [encode(a, 30), encode(b, 20), encode(c, 12)]
    text = "(a % b) * c + a % 7"
    e, got, want, _ = run_calc(U, O, text, cols, host, rows)
    assert np.array_equal(got, want)
    dev = [U.DeviceBuffer.from_numpy(h) for h in host]
    agg = U.Agg([text], cols)
    images = U.DeviceBuffer(133).zero()
    counts = U.DeviceBuffer(8).zero()
    agg.run(dev, rows, images, counts, max_groups=1)
    wsum, n = O.plan_sum(O.Plan(e.plan), host, [30, 20, 12], [3, 1, 2], rows,
                         mode=O.FAITHFUL if e.faithful else O.EXTENDED)
    assert n == rows
    assert bytes(images.numpy()) == wsum
