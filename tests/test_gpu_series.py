"""GPU parity for cast(expr, P, S) and the high-precision series workload:
sin(x) by a 25-term Horner series over x DECIMAL(300,100) in [-pi, pi],
per row and as SUM over 1,000,000 rows, bit-exact against the oracle."""
import numpy as np
import pytest

from ultraprecise_amd import workloads as W

pytestmark = pytest.mark.gpu


def maxbits(host, nb, rows):
    m = host.reshape(rows, nb).copy()
    m[:, -1] &= 0x7F
    nz = np.nonzero(m.any(axis=0))[0]
    if len(nz) == 0:
        return 1
    top = nz[-1]
    return int(8 * top + int(m[:, top].max()).bit_length())


def test_gen_below_device_equals_host(U, gpu):
    rows = 50_000
    d = W.sin_desc(U, 3)
    dev = U.gen_compact_device(d, rows)
    host = U.gen_compact_host(d, rows)
    assert np.array_equal(dev.numpy(count=host.size), host)


@pytest.mark.parametrize("P,S,T,digits", [
    (38, 10, 0, 38), (38, 10, 3, 20), (38, 10, 12, 38), (38, 10, 48, 30),
    (150, 50, 0, 150), (150, 50, 37, 80), (300, 100, 0, 300), (300, 100, 100, 250),
    (300, 100, 160, 200), (76, 20, 19, 5), (38, 10, 0, 8),
])
def test_cast_bit_exact(U, O, gpu, P, S, T, digits):
    """scale down (truncation by 10^n in chunks) and up (multiply by 10^n),
    including values that truncate to -0"""
    rows = 100_000
    d = U.gen_desc(U.GEN_DIGITS, P, 17, 0, digits=digits, neg_permille=500)
    host = U.gen_compact_host(d, rows)
    col = U.DeviceBuffer.from_numpy(host)
    nb = U.compact_bytes(P)
    cols = [("a", P, S, maxbits(host, nb, rows))]
    e = U.Expr(f"cast(a, {max(P, T + 1)}, {T})", cols)
    out = U.DeviceBuffer(rows * e.out_bytes + 4).zero()
    e.calc([col], rows, out)
    got = out.numpy(count=rows * e.out_bytes)
    mode = O.FAITHFUL if e.faithful else O.EXTENDED
    want = O.plan_eval(O.Plan(e.plan), [host], [P], [S], rows, e.out_bytes, mode=mode)
    assert np.array_equal(got, want)


def series_column(U, rows, seed):
    d = W.sin_desc(U, seed)
    dev = U.gen_compact_device(d, rows)
    bits = U.column_maxbits(dev, 300, rows)
    return dev, bits


def test_series_rows_bit_exact(U, O, gpu):
    rows = 20_000
    C = W.CONFIG4
    dev, bits = series_column(U, rows, C["seed"])
    host = dev.numpy(count=rows * U.compact_bytes(300))
    e = U.Expr(W.sin_expr(), [("x", 300, 100, bits)])
    out = U.DeviceBuffer(rows * e.out_bytes + 4).zero()
    e.calc([dev], rows, out)
    got = out.numpy(count=rows * e.out_bytes)
    want = O.plan_eval(O.Plan(e.plan), [host], [300], [100], rows, e.out_bytes, mode=O.EXTENDED)
    assert np.array_equal(got, want)


def test_series_sum_1m_bit_exact(U, O, gpu):
    C = W.CONFIG4
    rows = C["rows"]
    dev, bits = series_column(U, rows, C["seed"])
    host = dev.numpy(count=rows * U.compact_bytes(300))
    cols = [("x", 300, 100, bits)]
    agg = U.Agg([W.sin_expr()], cols)
    images = U.DeviceBuffer(133).zero()
    counts = U.DeviceBuffer(8).zero()
    ng, _ = agg.run([dev], rows, images, counts, max_groups=1)
    assert ng == 1
    e = U.Expr(W.sin_expr(), cols)
    want, n = O.plan_sum(O.Plan(e.plan), [host], [300], [100], rows, mode=O.EXTENDED, threads=16)
    assert n == rows
    assert counts.numpy("int64", count=1)[0] == rows
    assert bytes(images.numpy()) == want
