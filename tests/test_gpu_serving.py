"""GPU: the serving loop -- batched (graph) and plain (one launch per query)
give identical SUM images for every query of a seeded stream, equal to the
oracle's, with the stream as drawn (multi-lane classes included) on fresh
per-query columns; admission waits for device memory when the budget is
small (a resident column pool, and fresh columns under a byte cap)."""
import numpy as np
import pytest

from ultraprecise_amd import workloads as W
from ultraprecise_amd import serving

pytestmark = pytest.mark.gpu

ROWS = 40_000
CFG = dict(W.CONFIG5, rows=(1e3, ROWS))


def accept(U):
    def ok(text, cols):
        try:
            U.VmProgram(text, cols)
            return True
        except U.UpdecError as e:
            if "exceed" in str(e) or "JIT path only" in str(e):
                return False
            raise
    return ok


def stream(U, n, seed=4):
    return [W.random_query(seed, i, accept(U), config=CFG) for i in range(n)]


def test_serving_batched_equals_plain_equals_oracle(U, O, gpu):
    n = 80
    qs = stream(U, n)
    store = serving.ColumnStore(U, 8 << 30, seed=4, rows=ROWS)
    arrivals = serving.poisson_arrivals(n, 2000.0, seed=4)
    out = {}
    for batched in (True, False):
        imgs = U.DeviceBuffer(133 * n).zero()
        cnts = U.DeviceBuffer(8 * n).zero()
        srv = serving.Server(U, store, batched=batched, max_batch=32)
        st = srv.run(qs, arrivals, imgs, cnts)
        assert st["queries"] == n
        out[batched] = (imgs.numpy().reshape(n, 133), cnts.numpy("int64"))
        if batched:
            assert st["mean_batch"] > 1
    assert np.array_equal(out[True][0], out[False][0])
    assert np.array_equal(out[True][1], [q["rows"] for q in qs])
    # the oracle on every query (resident column prefixes)
    for i, q in enumerate(qs):
        rows = q["rows"]
        host = [store.cols[k].numpy(count=rows * U.compact_bytes(q["P"])) for k in q["keys"]]
        e = U.Expr(q["text"], q["cols"])
        P, S = [c[1] for c in q["cols"]], [c[2] for c in q["cols"]]
        if e.faithful:
            want, _ = O.plan_sum(O.Plan(e.plan), host, P, S, rows, mode=O.FAITHFUL, threads=16)
        else:   # 640-limb values: the widest programs pass the default 192
            want, _ = O.plan_sum_wide(e.plan, host, P, S, rows, mode=O.EXTENDED, threads=16)
        assert bytes(out[True][0][i]) == want, (i, q["text"], q["P"])


def test_serving_admission_by_device_memory(U, gpu):
    """a budget that holds only a few columns: queries wait for running ones
    to release theirs, evicted columns are regenerated, results unchanged"""
    n = 30
    qs = stream(U, n, seed=9)
    big = serving.ColumnStore(U, 8 << 30, seed=9, rows=ROWS)
    need = max(sum(big.nbytes(k[0]) for k in q["keys"]) for q in qs)
    small = serving.ColumnStore(U, need, seed=9, rows=ROWS)
    arrivals = np.zeros(n)
    res = []
    for store in (big, small):
        imgs = U.DeviceBuffer(133 * n).zero()
        cnts = U.DeviceBuffer(8 * n).zero()
        serving.Server(U, store, batched=True, max_batch=16).run(qs, arrivals, imgs, cnts)
        res.append(imgs.numpy())
        assert store.used <= store.budget
    assert small.generated > big.generated
    assert np.array_equal(res[0], res[1])


def test_serving_fresh_columns_as_drawn(U, O, gpu):
    """every query as drawn (no redraw: programs up to 512 limbs), each on
    its own columns generated at admission and freed at completion; a cap
    that holds about two queries' columns makes admission wait"""
    n = 40
    qs = [W.random_query(4, i, config=CFG) for i in range(n)]
    widths = {U.VmProgram(q["text"], q["cols"]).width for q in qs}
    assert max(widths) >= 64
    arrivals = serving.poisson_arrivals(n, 2000.0, seed=5)
    out = {}
    for name, cap in (("free", None), ("capped", None)):
        store = serving.FreshColumns(U, 4, reserve_bytes=1 << 30)
        if name == "capped":
            store.cap = 2 * max(store.nbytes(q) for q in qs)
        imgs = U.DeviceBuffer(133 * n).zero()
        cnts = U.DeviceBuffer(8 * n).zero()
        serving.Server(U, store, batched=True, max_batch=16).run(qs, arrivals, imgs, cnts)
        assert store.used == 0 and not store.live
        if name == "capped":
            assert store.refused > 0 and store.peak <= store.cap
        out[name] = (imgs.numpy().reshape(n, 133), cnts.numpy("int64"))
    assert np.array_equal(out["free"][0], out["capped"][0])
    assert np.array_equal(out["free"][1], [q["rows"] for q in qs])
    for i, q in enumerate(qs):
        rows = q["rows"]
        host = [U.gen_compact_host(W.query_column_desc(U, 4, q, k), rows) for k in range(len(q["cols"]))]
        e = U.Expr(q["text"], q["cols"])
        P, S = [c[1] for c in q["cols"]], [c[2] for c in q["cols"]]
        if e.faithful:
            want, _ = O.plan_sum(O.Plan(e.plan), host, P, S, rows, mode=O.FAITHFUL, threads=16)
        else:
            want, _ = O.plan_sum_wide(e.plan, host, P, S, rows, mode=O.EXTENDED, threads=16)
        assert bytes(out["free"][0][i]) == want, (i, q["text"], q["P"])
