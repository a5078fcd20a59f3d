"""GPU parity for the multi-lane form of the calc kernel (one row spread
over W / 4 or W / 8 lanes, the reference's XMP form): the same output rows
as the one-row-per-lane kernel and as the oracle, for the sweep expression
at P = 300 and for serving-shaped trees whose exact intermediates need 64
to 512 limbs."""
import numpy as np
import pytest

from ultraprecise_amd import workloads as W

pytestmark = pytest.mark.gpu


def calc(U, e, dev, rows, lanes):
    e.set_lanes(lanes)
    out = U.DeviceBuffer(rows * e.out_bytes + 4).zero()
    e.calc(dev, rows, out)
    full = out.numpy()
    assert not full[rows * e.out_bytes:].any(), "write past the last row"
    return full[: rows * e.out_bytes]


def test_sweep_p300_every_lane_count(U, O, gpu):
    cols = [(n, 300, 100) for n in "abcd"]
    e = U.Expr(W.CONFIG2["expr"], cols)
    assert e.multi_lane_width == 64 and not e.faithful
    rows = 3001
    host = [U.gen_compact_host(W.digits_desc(U, W.CONFIG2["seed"], i, 300), rows) for i in range(4)]
    dev = [U.DeviceBuffer.from_numpy(h) for h in host]
    want = O.plan_eval(O.Plan(e.plan), host, [300] * 4, [100] * 4, rows, e.out_bytes, mode=O.EXTENDED)
    for lanes in (1, 16, 8):
        got = calc(U, e, dev, rows, lanes)
        bad = np.nonzero((got != want).reshape(rows, -1).any(axis=1))[0]
        assert bad.size == 0, f"lanes={lanes}: {bad.size} rows differ, first {bad[:5]}"


def test_serving_trees_multi_lane(U, O, gpu):
    """one tree of the serving stream per width of 64, 128 and 256 limbs,
    with odd output widths (byte stores) and columns of odd widths, at both
    lane counts against the oracle (the one-row-per-lane kernels of these
    widths take 20 s and more to compile; the sweep test covers that form)"""
    seen = {}
    for i in range(4000):
        q = W.random_query(93, i)
        e = U.Expr(q["text"], q["cols"])
        w = e.multi_lane_width
        if w in (64, 128, 256) and w not in seen:
            seen[w] = (q, e)
        if len(seen) == 3:
            break
    assert sorted(seen) == [64, 128, 256]
    for n, (w, (q, e)) in enumerate(sorted(seen.items())):
        rows = 257 + 61 * n
        ids = dict(q, col_ids=[800 + 10 * n + k for k in range(4)])
        host = [U.gen_compact_host(W.query_column_desc(U, 7, ids, k), rows) for k in range(len(q["cols"]))]
        dev = [U.DeviceBuffer.from_numpy(h) for h in host]
        P = [c[1] for c in q["cols"]]
        S = [c[2] for c in q["cols"]]
        want = O.plan_eval_wide(e.plan, host, P, S, rows, e.out_bytes, mode=O.EXTENDED)
        for lanes in (w // 8, w // 4):
            if lanes > 64:
                continue
            got = calc(U, e, dev, rows, lanes)
            assert bytes(got) == bytes(want), (w, lanes, q["text"])
