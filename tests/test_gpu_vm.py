"""GPU parity for the precompiled interpreter and its batches: the same
SUM images as the JIT aggregation kernels and as the oracle, alone and in
mixed batches, eagerly and from a replayed HIP graph."""
import numpy as np
import pytest

from ultraprecise_amd import workloads as W

pytestmark = pytest.mark.gpu


def columns(U, specs, rows, seed):
    dev = []
    for i, (name, P, S) in enumerate(specs):
        d = U.gen_desc(U.GEN_DIGITS, P, seed, 500 + i, digits=max(1, P - (i % 3)), neg_permille=500)
        dev.append(U.gen_compact_device(d, rows))
    return dev


def host_of(U, dev, specs, rows):
    return [t.numpy(count=rows * U.compact_bytes(c[1])) for t, c in zip(dev, specs)]


def jit_sum(U, text, cols, dev, rows):
    agg = U.Agg([text], cols)
    img = U.DeviceBuffer(133).zero()
    cnt = U.DeviceBuffer(8).zero()
    agg.run(dev, rows, img, cnt, max_groups=1)
    return bytes(img.numpy()), int(cnt.numpy("int64")[0])


def vm_sum(U, prog, dev, rows, graph=False, batch=None):
    b = batch or U.Batch(max_queries=4, max_items=4096)
    img = U.DeviceBuffer(133).zero()
    cnt = U.DeviceBuffer(8).zero()
    b.add(prog, dev, rows, img, cnt)
    b.launch(graph=graph)
    U.device_sync()
    return bytes(img.numpy()), int(cnt.numpy("int64")[0])


C15 = [("q", 15, 2), ("p", 15, 2), ("d", 15, 2), ("t", 15, 2)]
CASES = [
    ("p*(1-d)", C15), ("q*p", C15), ("p*(1-d)*(1+t)", C15), ("q", C15),
    ("(a+b)*c - a/d", [(n, 18, 4) for n in "abcd"]),
    ("(a+b)*c - a/d", [(n, 38, 10) for n in "abcd"]),
    ("(a+b)*c - a/d", [(n, 76, 20) for n in "abcd"]),
    ("(a+b)*c - a/d", [(n, 150, 50) for n in "abcd"]),
    ("a/b + cast(c, 40, 3) - cast(a*b, 60, 30)", [(n, 38, 10) for n in "abc"]),
    ("1.5 - a*2 + 0.25/b", [(n, 18, 4) for n in "ab"]),
    ("truncate(a, 2) + truncate(b, -3) - truncate(a*b, 0)", [(n, 20, 6) for n in "ab"]),
    ("truncate(a, -14) + truncate(-b, -20) * truncate(a, 9)", [(n, 20, 6) for n in "ab"]),
]


@pytest.mark.parametrize("k", range(len(CASES)))
def test_vm_matches_jit_and_oracle(U, O, gpu, k):
    text, cols = CASES[k]
    rows = 100_003
    dev = columns(U, cols, rows, 31 + k)
    prog = U.VmProgram(text, cols)
    got = vm_sum(U, prog, dev, rows)
    assert got == jit_sum(U, text, cols, dev, rows)
    assert got[1] == rows
    e = U.Expr(text, cols)
    mode = O.FAITHFUL if e.faithful else O.EXTENDED
    want, n = O.plan_sum(O.Plan(e.plan), host_of(U, dev, cols, rows), [c[1] for c in cols], [c[2] for c in cols],
                         rows, mode=mode, threads=16)
    assert got[0] == want


def test_vm_series(U, O, gpu):
    """the 25-term series in the widest register class"""
    rows = 20_000
    dev = U.gen_compact_device(W.sin_desc(U, 3), rows)
    bits = U.column_maxbits(dev, 300, rows)
    cols = [("x", 300, 100, bits)]
    prog = U.VmProgram(W.sin_expr(), cols)
    assert prog.width == 32
    got = vm_sum(U, prog, [dev], rows)
    assert got == jit_sum(U, W.sin_expr(), cols, [dev], rows)


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


def test_vm_batch_of_random_queries(U, O, gpu):
    """a mixed batch (every width class, uneven rows) gives each query the
    result it gets alone, and the oracle's"""
    n = 48
    qs = [W.random_query(7, i, accept(U)) for i in range(n)]
    batch = U.Batch(max_queries=64, max_items=8192)
    batch.tune(0, 4096, 4)   # small items: many items per query
    imgs = U.DeviceBuffer(n * 133).zero()
    cnts = U.DeviceBuffer(n * 8).zero()
    data = []
    for i, q in enumerate(qs):
        rows = min(q["rows"], 30_000) if i % 7 else 0
        dev = [U.gen_compact_device(W.query_column_desc(U, 7, q, k), rows) for k in range(len(q["cols"]))]
        prog = U.VmProgram(q["text"], q["cols"])
        batch.add(prog, dev, rows, imgs, cnts, image_offset=133 * i, count_offset=8 * i)
        data.append((q, rows, dev, prog))
    classes = {d[3].width for d in data}
    assert len(classes) >= 4, classes
    batch.launch(graph=True)
    U.device_sync()
    got = imgs.numpy().reshape(n, 133)
    gotc = cnts.numpy("int64")
    for i, (q, rows, dev, prog) in enumerate(data):
        alone = vm_sum(U, prog, dev, rows)
        assert bytes(got[i]) == alone[0], (i, q["text"])
        assert gotc[i] == rows
        e = U.Expr(q["text"], q["cols"])
        host = host_of(U, dev, q["cols"], rows) if rows else [np.zeros(1, np.uint8)] * len(dev)
        P, S = [c[1] for c in q["cols"]], [c[2] for c in q["cols"]]
        if e.faithful:
            want, _ = O.plan_sum(O.Plan(e.plan), host, P, S, rows, mode=O.FAITHFUL, threads=16)
        else:   # 640-limb values: the widest programs pass the default 192
            want, _ = O.plan_sum_wide(e.plan, host, P, S, rows, mode=O.EXTENDED, threads=16)
        assert bytes(got[i]) == want, (i, q["text"], q["P"])


def test_vm_graph_replays(U, gpu):
    """the captured graph replayed several times, with eager launches and
    synchronizations in between, gives the eager result every time"""
    cols = [(n, 38, 10) for n in "abc"]
    texts = ["(a+b)*c", "a/b - c", "a*b*c", "(a-b)/(c+a)"]
    rows = [50_000, 7, 120_001, 3_000]
    dev = [columns(U, cols, r, 90 + i) for i, r in enumerate(rows)]
    progs = [U.VmProgram(t, cols) for t in texts]
    eager = [vm_sum(U, p, d, r) for p, d, r in zip(progs, dev, rows)]
    batch = U.Batch(max_queries=8, max_items=1024)
    imgs = U.DeviceBuffer(len(texts) * 133)
    cnts = U.DeviceBuffer(len(texts) * 8)
    other = U.Batch(max_queries=2, max_items=64)
    for rep in range(4):
        imgs.zero()
        cnts.zero()
        for i in range(len(texts)):
            batch.add(progs[i], dev[i], rows[i], imgs, cnts, image_offset=133 * i, count_offset=8 * i)
        batch.launch(graph=True)
        U.device_sync()
        got = imgs.numpy().reshape(-1, 133)
        gc = cnts.numpy("int64")
        for i in range(len(texts)):
            assert (bytes(got[i]), int(gc[i])) == eager[i], (rep, i)
        # eager work on the device between replays
        vm_sum(U, progs[rep % 4], dev[rep % 4], rows[rep % 4], batch=other)


def test_query_larger_than_batch_items(U, gpu):
    """a query whose rows need more items than the batch holds at the tuned
    rows per item takes larger items instead; the SUM is unchanged"""
    cols = [(n, 38, 10) for n in "abcd"]
    text = "(a+b)*c - a/d"
    rows = 1_000_003
    dev = columns(U, cols, rows, 77)
    prog = U.VmProgram(text, cols)
    small = U.Batch(max_queries=4, max_items=64)
    small.tune(prog.width_class, 1024, 4)   # 977 items at 1024 rows: 64 larger ones instead
    got = vm_sum(U, prog, dev, rows, batch=small)
    assert got == vm_sum(U, prog, dev, rows, batch=U.Batch(max_queries=4, max_items=4096))
    assert got[1] == rows
