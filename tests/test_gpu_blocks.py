"""GPU parity of block-column inputs (the reference's AriesColumnDataIterator:
data blocks with start-row prefix sums, optional int32 row indices with -1
for NULL) through the C ABI, bit-exact against the oracle evaluated at the
gathered positions."""
import numpy as np
import pytest

from ultraprecise_amd import workloads as W

pytestmark = pytest.mark.gpu


def split_column(U, host, B, rows, rng, nblocks):
    """upload one host column as nblocks separate device blocks of random sizes"""
    cuts = np.sort(rng.choice(np.arange(1, rows), size=nblocks - 1, replace=False)) if nblocks > 1 else []
    edges = [0] + list(cuts) + [rows]
    blocks, counts = [], []
    for lo, hi in zip(edges[:-1], edges[1:]):
        blocks.append(U.DeviceBuffer.from_numpy(host[lo * B:hi * B]))
        counts.append(hi - lo)
    return U.BlockColumn(blocks, counts)


def make_cols(U, specs, rows, seed):
    host = []
    for i, (P, S) in enumerate(specs):
        t = U.gen_compact_device(W.digits_desc(U, seed, i, P), rows)
        host.append(t.numpy(count=rows * U.compact_bytes(P)))
    return host


def gather(host, B, pos):
    out = np.zeros(len(pos) * B, dtype=np.uint8)
    v = out.reshape(len(pos), B)
    ok = pos >= 0
    v[ok] = host.reshape(-1, B)[pos[ok]]
    return out


CASES = [("p*(1-d)", [(15, 2), (15, 2)]),
         ("(a+b)*c - a/d", [(38, 10)] * 4),
         ("a % b + c", [(30, 0), (20, 0), (18, 2)]),
         # extended plan with 126-byte rows: the 64-row block form
         ("(a+b)*c - a/d", [(150, 50)] * 4)]


@pytest.mark.parametrize("text,specs", CASES)
@pytest.mark.parametrize("indexed", [False, True])
def test_calc_blocks_bit_exact(U, O, gpu, text, specs, indexed):
    rng = np.random.default_rng(len(text) + indexed)
    rows = 50_000
    names = "pd" if len(specs) == 2 else "abcd"[:len(specs)]
    cols = [(names[i], P, S) for i, (P, S) in enumerate(specs)]
    host = make_cols(U, specs, rows, seed=4)
    bcols = [split_column(U, h, U.compact_bytes(P), rows, rng, int(rng.integers(1, 9))) for h, (P, S) in zip(host, specs)]
    e = U.Expr(text, cols)
    ob = e.out_bytes
    if indexed:
        n = 70_000
        pos = rng.integers(0, rows, size=n).astype(np.int64)
        pos[rng.random(n) < 0.05] = -1                   # NULL rows
        pos[rng.random(n) < 0.01] = rows + 7             # past the columns
        idx = U.DeviceBuffer.from_numpy(pos.astype(np.int32))
    else:
        n = rows
        pos = np.arange(rows, dtype=np.int64)
        idx = None
    out = U.DeviceBuffer(n * ob + 8).zero()
    valid = U.DeviceBuffer(n).zero()
    e.calc_blocks(bcols, n, out, indices=idx, valid=valid)
    got = out.numpy(count=n * ob).reshape(n, ob)
    vflags = valid.numpy(count=n)
    ok = (pos >= 0) & (pos < rows)
    assert np.array_equal(vflags, ok.astype(np.uint8))
    assert not got[~ok].any(), "NULL rows must be zero"
    gpos = np.where(ok, pos, -1)
    ghost = [gather(h, U.compact_bytes(P), gpos) for h, (P, S) in zip(host, specs)]
    want = O.plan_eval(O.Plan(e.plan), ghost, [P for P, S in specs], [S for P, S in specs], n, ob,
                       mode=O.FAITHFUL if e.faithful else O.EXTENDED).reshape(n, ob)
    bad = np.nonzero((got[ok] != want[ok]).any(axis=1))[0]
    assert bad.size == 0, f"{bad.size} rows differ"
    assert not out.numpy()[n * ob:].any()


@pytest.mark.parametrize("text,specs", CASES[:2])
def test_agg_blocks_bit_exact(U, O, gpu, text, specs):
    rng = np.random.default_rng(7)
    rows = 60_000
    names = "pd" if len(specs) == 2 else "abcd"
    cols = [(names[i], P, S) for i, (P, S) in enumerate(specs)]
    host = make_cols(U, specs, rows, seed=5)
    bcols = [split_column(U, h, U.compact_bytes(P), rows, rng, 5) for h, (P, S) in zip(host, specs)]
    n = 80_000
    pos = rng.integers(0, rows, size=n).astype(np.int64)
    pos[rng.random(n) < 0.1] = -1
    idx = U.DeviceBuffer.from_numpy(pos.astype(np.int32))
    agg = U.Agg([text], cols, avg=[0])
    images = U.DeviceBuffer(2 * 133).zero()
    counts = U.DeviceBuffer(8).zero()
    agg.run_blocks(bcols, n, images, counts, indices=idx)
    keep = pos[pos >= 0]
    ghost = [gather(h, U.compact_bytes(P), keep) for h, (P, S) in zip(host, specs)]
    e = U.Expr(text, cols)
    want, cnt = O.plan_sum(O.Plan(e.plan), ghost, [P for P, S in specs], [S for P, S in specs], len(keep),
                           mode=O.FAITHFUL if e.faithful else O.EXTENDED)
    assert int(counts.numpy("int64")[0]) == len(keep) == cnt
    img = bytes(images.numpy())
    assert img[:133] == want
    # MODE_AUTO divides by the faithful rule, as div_int64 does
    assert img[133:] == O.avg_from_sum(want, cnt, mode=O.FAITHFUL)


def test_blocks_argument_errors(U, gpu):
    rng = np.random.default_rng(1)
    host = make_cols(U, [(15, 2)], 100, seed=1)
    b = split_column(U, host[0], U.compact_bytes(15), 100, rng, 3)
    e = U.Expr("p*2", [("p", 15, 2)])
    out = U.DeviceBuffer(200 * e.out_bytes)
    with pytest.raises(U.UpdecError, match="exceed"):
        e.calc_blocks([b], 200, out)           # no indices: rows must be within the columns
    agg = U.Agg(["p"], [("p", 15, 2)], filter=("<=", 3))
    with pytest.raises(U.UpdecError, match="unfiltered"):
        agg.run_blocks([b], 100, out, out)
