"""GPU parity: the HIP path through the C ABI against the oracle, bit-exact,
on every row."""
import numpy as np
import pytest

from ultraprecise_amd import workloads as W

pytestmark = pytest.mark.gpu


def _gen_cols(U, spec, rows, seed):
    """device columns + identical host copies"""
    dev, host = [], []
    for i, c in enumerate(spec):
        d = W.gen_desc(U, seed, i, c)
        t = U.gen_compact_device(d, rows)
        h = t.numpy(count=rows * U.compact_bytes(c[1]))
        dev.append(t)
        host.append(h)
    return dev, host


def test_generator_host_equals_device(U, gpu):
    for spec in W.CONFIG1["cols"]:
        d = W.gen_desc(U, 7, 3, spec)
        h = U.gen_compact_host(d, 5000, row0=123)
        g = U.gen_compact_device(d, 5000, row0=123).numpy(count=h.size)
        assert np.array_equal(h, g)
    for P in (18, 38, 76, 150, 300):
        d = W.digits_desc(U, 1, 2, P)
        h = U.gen_compact_host(d, 777, row0=5)
        g = U.gen_compact_device(d, 777, row0=5).numpy(count=h.size)
        assert np.array_equal(h, g)


@pytest.mark.parametrize("rows", [1, 255, 256, 1000, 100_003, 1_000_000])
def test_config1_calc_rows_bit_exact(U, O, gpu, rows):
    spec = W.CONFIG1["cols"]
    cols = [(c[0], c[1], c[2]) for c in spec]
    dev, host = _gen_cols(U, spec, rows, W.CONFIG1["seed"])
    for text in W.CONFIG1["exprs"]:
        e = U.Expr(text, cols)
        out = U.DeviceBuffer(rows * e.out_bytes + 4).zero()
        e.calc(dev, rows, out)
        full = out.numpy()
        got = full[: rows * e.out_bytes]
        plan = O.Plan(e.plan)
        want = O.plan_eval(plan, host, [c[1] for c in cols], [c[2] for c in cols], rows, e.out_bytes)
        assert np.array_equal(got, want), f"{text}: first mismatch row {np.argmax(got != want) // e.out_bytes}"
        assert not full[rows * e.out_bytes:].any(), "kernel wrote past the last row"


@pytest.mark.parametrize("rows", [1, 1000, 1_000_000])
def test_config1_fused_sum_bit_exact(U, O, gpu, rows):
    spec = W.CONFIG1["cols"]
    cols = [(c[0], c[1], c[2]) for c in spec]
    dev, host = _gen_cols(U, spec, rows, W.CONFIG1["seed"])
    agg = U.Agg(W.CONFIG1["exprs"], cols)
    images = U.DeviceBuffer(len(W.CONFIG1["exprs"]) * 133).zero()
    counts = U.DeviceBuffer(8).zero()
    ng, keys = agg.run(dev, rows, images, counts)
    assert ng == 1
    got = images.numpy().reshape(-1, 133)
    assert int(counts.numpy("int64")[0]) == rows
    for k, text in enumerate(W.CONFIG1["exprs"]):
        e = U.Expr(text, cols)
        want, n = O.plan_sum(O.Plan(e.plan), host, [c[1] for c in cols], [c[2] for c in cols], rows)
        assert n == rows
        assert bytes(got[k]) == want, f"SUM({text}): {O.image_value(got[k])} vs {O.image_value(want)}"


# output row widths: 4 (dword stores), 8 (dwordx2), 16 and 32 (dwordx4), 5
# and 17 (LDS tile); each at a 16-byte aligned output and at one 4 bytes past
# it, where only the tile path may run
@pytest.mark.parametrize("P,S,ob", [(8, 2, 4), (17, 4, 8), (37, 9, 16), (75, 20, 32), (9, 2, 5), (38, 9, 17)])
@pytest.mark.parametrize("offset", [0, 4])
def test_calc_output_store_paths(U, O, gpu, P, S, ob, offset):
    rows = 70_001
    cols = [("a", P, S), ("b", P, S)]
    dev = [U.gen_compact_device(W.digits_desc(U, 11, i, P), rows) for i in range(2)]
    host = [t.numpy(count=rows * U.compact_bytes(P)) for t in dev]
    e = U.Expr("a+b", cols)
    assert e.out_bytes == ob
    buf = U.DeviceBuffer(rows * ob + 64).zero()
    e.calc(dev, rows, U.DeviceView(buf, offset, rows * ob))
    full = buf.numpy()
    got = full[offset: offset + rows * ob]
    want = O.plan_eval(O.Plan(e.plan), host, [P, P], [S, S], rows, ob)
    assert np.array_equal(got, want), f"first mismatch row {np.argmax(got != want) // ob}"
    assert not full[:offset].any() and not full[offset + rows * ob:].any(), "kernel wrote outside the output"
