"""GPU: the workloads at their full sizes against the oracle's results
committed in tests/golden/fullsize.json (written by
tests/golden/make_fullsize.py in the build container, with plans from the
independent rewrite restatement).  Every output is compared whole: the
grouped aggregation's images and counts over all 60,000,000 rows, the SHA-256
of every precision's 10,000,000-row output column, the series' SUM image over
all 1,000,000 rows."""
import ctypes
import hashlib
import json
import os

import numpy as np
import pytest

from ultraprecise_amd import workloads as W

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "fullsize.json")))


def test_plain_1M(U, gpu):
    G = GOLD["plain"]
    C = W.CONFIG1
    rows = C["rows"]
    assert rows == G["rows"]
    dev = [U.gen_compact_device(W.gen_desc(U, C["seed"], i, c), rows) for i, c in enumerate(C["cols"])]
    cols = [(c[0], c[1], c[2]) for c in C["cols"]]
    agg = U.Agg(C["exprs"], cols)
    images = U.DeviceBuffer(len(C["exprs"]) * 133).zero()
    counts = U.DeviceBuffer(8).zero()
    agg.run(dev, rows, images, counts)
    got = images.numpy().reshape(-1, 133)
    for k, text in enumerate(C["exprs"]):
        assert bytes(got[k]).hex() == G["sums"][text]["image"], text
    assert int(counts.numpy("int64")[0]) == rows


@pytest.mark.parametrize("stats", ["column_bits", "no_statistics"])
def test_q1_60M(U, gpu, stats):
    """the headline aggregation, with the load-time column statistics the
    bench passes and with none (general limb widths)"""
    G = GOLD["q1"]
    C = W.CONFIG3
    rows, seed = C["rows"], C["seed"]
    assert rows == G["rows"] and C["exprs"] == G["exprs"]
    dev = [U.gen_compact_device(W.gen_desc(U, seed, i, c), rows) for i, c in enumerate(C["cols"])]
    f1 = U.gen_byte_device(seed, W.COL_FLAG1, C["flag1"], rows)
    f2 = U.gen_byte_device(seed, W.COL_FLAG2, C["flag2"], rows)
    ship = U.gen_int32_device(seed, W.COL_SHIP, C["ship"][0], C["ship"][1], rows)
    if stats == "column_bits":
        bits = [U.column_maxbits(t, c[1], rows) for t, c in zip(dev, C["cols"])]
    else:
        bits = [0] * len(dev)
    cols = [(c[0], c[1], c[2], b) for c, b in zip(C["cols"], bits)]
    agg = U.Agg(C["exprs"], cols, filter=C["filter"], nkeys=2, avg=tuple(G["avg_of"]))
    ne, maxg = len(C["exprs"]), 16
    images = U.DeviceBuffer(maxg * (ne + 3) * 133).zero()
    counts = U.DeviceBuffer(maxg * 8).zero()
    ng, keys = agg.run(dev, rows, images, counts, filter_col=ship, key0=f1, key1=f2, max_groups=maxg)
    assert ng == len(G["groups"])
    got = images.numpy().reshape(-1, 133)
    gc = counts.numpy("int64")
    for g, want in enumerate(G["groups"]):
        assert [int(keys[g]) & 255, int(keys[g]) >> 8] == want["key"], g
        assert int(gc[g]) == want["count"], g
        for k in range(ne):
            assert bytes(got[g * ne + k]).hex() == want["sums"][k], (g, C["exprs"][k])
        for j in range(3):
            assert bytes(got[maxg * ne + g * 3 + j]).hex() == want["avgs"][j], (g, j)


@pytest.mark.parametrize("P", [18, 38, 76, 150, 300])
def test_sweep_10M(U, gpu, P):
    G = GOLD["sweep"]
    C = W.CONFIG2
    rows = C["rows"]
    want = G["precisions"][str(P)]
    S = want["scale"]
    cols = [(n, P, S) for n in "abcd"]
    e = U.Expr(C["expr"], cols)
    assert e.faithful == (want["mode"] == "faithful")
    assert [e.precision, e.scale] == want["out_type"] and e.out_bytes == want["out_bytes"]
    dev = [U.gen_compact_device(W.digits_desc(U, C["seed"], i, P), rows) for i in range(4)]
    out = U.DeviceBuffer(rows * e.out_bytes + 16)
    e.calc(dev, rows, out)
    h = hashlib.sha256()
    chunk = 1_000_000 * e.out_bytes
    total = rows * e.out_bytes
    for off in range(0, total, chunk):
        n = min(chunk, total - off)
        h.update(_dtoh(U, out, off, n))
    assert h.hexdigest() == want["sha256"], P


def _dtoh(U, buf, off, n):
    a = np.empty(n, dtype=np.uint8)
    U._check(U.lib().updec_memcpy_dtoh(a.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(buf.ptr + off), n, None))
    return a.tobytes()


def test_series_1M(U, gpu):
    G = GOLD["series"]
    C = W.CONFIG4
    rows = C["rows"]
    dev = U.gen_compact_device(W.sin_desc(U, C["seed"]), rows)
    bits = U.column_maxbits(dev, 300, rows)
    agg = U.Agg([W.sin_expr()], [("x", 300, 100, bits)])
    images = U.DeviceBuffer(133).zero()
    counts = U.DeviceBuffer(8).zero()
    agg.run([dev], rows, images, counts, max_groups=1)
    assert bytes(images.numpy()).hex() == G["image"]
    assert int(counts.numpy("int64")[0]) == rows
