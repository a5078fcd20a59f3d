"""GPU parity for the interpreter's multi-lane classes (W = 64, 128, 256 and
512 limbs, one value per group of W / 4 or W / 8 lanes, at most 64):
serving-shaped queries whose exact intermediates need more than 32 limbs,
batched, against the oracle in extended mode; both lane counts give the
same images."""
import numpy as np
import pytest

from ultraprecise_amd import workloads as W

pytestmark = pytest.mark.gpu


def wide_queries(U, want_widths, per_width, seed=91, max_draws=3000):
    found = {w: [] for w in want_widths}
    for i in range(max_draws):
        q = W.random_query(seed, i)
        try:
            v = U.VmProgram(q["text"], q["cols"])
        except U.UpdecError:
            continue
        if v.width in found and len(found[v.width]) < per_width:
            found[v.width].append((q, v))
        if all(len(x) >= per_width for x in found.values()):
            break
    return found


def run_batch(U, items, lanes):
    b = U.Batch(max_queries=64, max_items=1 << 14)
    for c in range(U.VM_FIRST_MULTI_LANE, len(U.VM_CLASS_WIDTHS)):
        w = U.VM_CLASS_WIDTHS[c]
        b.tune(c, 256, 4, lanes=min(64, w // lanes))
    imgs = U.DeviceBuffer(133 * len(items)).zero()
    cnts = U.DeviceBuffer(8 * len(items)).zero()
    views = []
    for k, (v, dev, rows) in enumerate(items):
        iv = U.DeviceView(imgs, 133 * k, 133)
        cv = U.DeviceView(cnts, 8 * k, 8)
        b.add(v, dev, rows, iv, cv)
        views.append((iv, cv))
    b.launch(graph=False)
    U.device_sync()
    return imgs.numpy().reshape(-1, 133), cnts.numpy("int64")


@pytest.mark.parametrize("width", [64, 128, 256, 512])
def test_multi_lane_class_matches_oracle(U, O, gpu, width):
    found = wide_queries(U, [width], 2 if width == 512 else 4, max_draws=8000)
    assert found[width], f"no seeded query needs {width} limbs"
    items, hosts = [], []
    for n, (q, v) in enumerate(found[width]):
        rows = 700 + 131 * n
        dev = [U.gen_compact_device(U.gen_desc(U.GEN_DIGITS, q["P"], 5, 7000 + 10 * n + k, digits=q["P"],
                                               neg_permille=500), rows) for k in range(len(q["cols"]))]
        items.append((v, dev, rows))
        hosts.append([t.numpy(count=rows * U.compact_bytes(q["P"])) for t in dev])
    got4, cnt4 = run_batch(U, items, 4)     # 4 limbs per lane
    got8, cnt8 = run_batch(U, items, 8)     # 8 limbs per lane
    for n, (q, v) in enumerate(found[width]):
        rows = items[n][2]
        e = U.Expr(q["text"], q["cols"])
        assert not e.faithful
        want, _ = O.plan_sum_wide(e.plan, hosts[n], [c[1] for c in q["cols"]], [c[2] for c in q["cols"]], rows,
                                  mode=O.EXTENDED, threads=16)
        assert bytes(got4[n]) == want, (width, q["text"])
        assert bytes(got8[n]) == want, (width, q["text"], "8 limbs per lane")
        assert int(cnt4[n]) == rows and int(cnt8[n]) == rows


def test_multi_lane_sweep_p300(U, O, gpu):
    """the sweep expression at P = 300 (64 limbs): every operation of the
    multi-lane path (alignment, product, long division, rounding)"""
    cols = [(n, 300, 100) for n in "abcd"]
    text = "(a+b)*c - a/d"
    v = U.VmProgram(text, cols)
    assert v.width == 64
    rows = 3000
    dev = [U.gen_compact_device(W.digits_desc(U, 1, i, 300), rows) for i in range(4)]
    got, cnt = run_batch(U, [(v, dev, rows)], 4)
    e = U.Expr(text, cols)
    want, _ = O.plan_sum(O.Plan(e.plan), [t.numpy(count=rows * U.compact_bytes(300)) for t in dev], [300] * 4,
                         [100] * 4, rows, mode=O.EXTENDED, threads=16)
    assert bytes(got[0]) == want
