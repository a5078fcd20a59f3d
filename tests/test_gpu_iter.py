"""GPU: expressions and grouped aggregation over column iterators (the
reference's AriesColumnDataIterator): two columns with their own data
blocks and their own index arrays, one with nullable rows and a
nullable_type<int32> index split into index blocks, the other with an int32
index holding -1 and positions past its end; checked against the oracle on
the rows gathered on the host, with the NULL rules applied there."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PA, SA = 38, 10
PB, SB = 15, 2


def _blocks(U, rows_bytes, counts, stride):
    out, off = [], 0
    for n in counts:
        out.append(U.DeviceBuffer.from_numpy(np.ascontiguousarray(rows_bytes[off:off + n * stride])))
        off += n * stride
    return out


def _split(rng, n, k):
    cuts = np.sort(rng.choice(np.arange(1, n), size=k - 1, replace=False))
    return np.diff(np.concatenate([[0], cuts, [n]])).astype(np.int64)


def _setup(U, seed, rows):
    rng = np.random.default_rng(seed)
    Ba, Bb = U.compact_bytes(PA), U.compact_bytes(PB)
    na, nb = rows + 500, rows // 2 + 300          # positions held by each column
    a = U.gen_compact_host(U.gen_desc(U.GEN_DIGITS, PA, seed, 1, digits=30, neg_permille=500), na).reshape(na, Ba)
    b = U.gen_compact_host(U.gen_desc(U.GEN_RANGE, PB, seed, 2, lo=0, hi=40, step=25), nb).reshape(nb, Bb)
    # column a: nullable rows (flag byte, then the compact bytes), 10% NULL
    aflag = (rng.random(na) >= 0.1).astype(np.uint8)
    arows = np.concatenate([aflag[:, None], a], axis=1)
    acounts = _split(rng, na, 7)
    # its index: nullable_type<int32> (5 packed bytes), 5% NULL, in 3 blocks
    apos = rng.integers(0, na, size=rows).astype(np.int32)
    aiflag = (rng.random(rows) >= 0.05).astype(np.uint8)
    aidx = np.concatenate([aiflag[:, None], apos.view(np.uint8).reshape(rows, 4)], axis=1)
    aicounts = _split(rng, rows, 3)
    # column b: plain rows in 5 blocks; int32 index with -1 and positions
    # past the column's end
    bcounts = _split(rng, nb, 5)
    bpos = rng.integers(-1, nb + 20, size=rows).astype(np.int32)
    colA = U.ColumnIter(_blocks(U, arows.reshape(-1), acounts, Ba + 1), list(acounts), nullable=True,
                        index=_blocks(U, aidx.reshape(-1), aicounts, 5), index_counts=list(aicounts),
                        index_kind=U.INDEX_NULLABLE_INT32)
    colB = U.ColumnIter(_blocks(U, b.reshape(-1), bcounts, Bb), list(bcounts),
                        index=[U.DeviceBuffer.from_numpy(bpos)])
    # host gather: values (zero rows for NULL) and validity per launch row
    a_ok = aiflag.astype(bool) & (aflag[apos] != 0)
    b_ok = (bpos >= 0) & (bpos < nb)
    ga = np.where(a_ok[:, None], a[apos], 0).astype(np.uint8)
    gb = np.where(b_ok[:, None], b[np.clip(bpos, 0, nb - 1)], 0).astype(np.uint8)
    b_zero = ~np.any(gb & np.array([0xff] * (Bb - 1) + [0x7f], dtype=np.uint8), axis=1)
    return [colA, colB], ga, gb, a_ok, b_ok, b_zero


COLS = [("a", PA, SA), ("b", PB, SB)]


@pytest.mark.parametrize("text", ["a/b + a*b", "(a - b) * b", "a % b"])
def test_calc_iter_matches_gathered_oracle(U, O, gpu, text):
    rows = 30_000
    cols, ga, gb, a_ok, b_ok, b_zero = _setup(U, 41, rows)
    e = U.Expr(text, COLS)
    OB = e.out_bytes
    mode = O.FAITHFUL if e.faithful else O.EXTENDED
    want = O.plan_eval(O.Plan(e.plan), [ga.reshape(-1), gb.reshape(-1)], [PA, PB], [SA, SB], rows, OB,
                       mode=mode).reshape(rows, OB)
    null = ~(a_ok & b_ok)
    if "/" in text or "%" in text:
        null |= b_zero    # both columns can be NULL: a zero divisor gives NULL
    want[null] = 0
    for out_nullable in (False, True):
        stride = OB + (1 if out_nullable else 0)
        out = U.DeviceBuffer(rows * stride + 16)
        valid = U.DeviceBuffer(rows)
        e.calc_iter(cols, rows, out, out_nullable=out_nullable, valid=valid)
        got = out.numpy(count=rows * stride).reshape(rows, stride)
        v = valid.numpy(count=rows)
        assert np.array_equal(v, (~null).astype(np.uint8)), text
        if out_nullable:
            assert np.array_equal(got[:, 0], (~null).astype(np.uint8))
            got = got[:, 1:]
        bad = np.flatnonzero(np.any(got != want, axis=1))
        assert bad.size == 0, (text, out_nullable, bad[:5])


def test_agg_iter_grouped_filtered(U, O, gpu):
    rows = 40_000
    cols, ga, gb, a_ok, b_ok, b_zero = _setup(U, 43, rows)
    rng = np.random.default_rng(7)
    f = rng.integers(0, 100, size=rows).astype(np.int32)
    k0 = rng.choice(np.frombuffer(b"XYZ", dtype=np.uint8), size=rows)
    k1 = rng.choice(np.frombuffer(b"PQ", dtype=np.uint8), size=rows)
    exprs = ["a", "a*b", "a/b"]
    agg = U.Agg(exprs, COLS, filter=("<", 70), nkeys=2)
    maxg = 8
    images = U.DeviceBuffer(maxg * len(exprs) * 133).zero()
    counts = U.DeviceBuffer(maxg * 8).zero()
    ng, keys = agg.run_iter(cols, rows, images, counts, filter_col=U.DeviceBuffer.from_numpy(f),
                            key0=U.DeviceBuffer.from_numpy(k0), key1=U.DeviceBuffer.from_numpy(k1), max_groups=maxg)
    mask = f < 70
    key = (k0.astype(np.int32) << 8) | k1
    ukeys = np.unique(key[mask])
    assert ng == len(ukeys)
    assert keys == [int((k >> 8) | ((k & 0xff) << 8)) for k in ukeys]
    lut = np.full(65536, -1, dtype=np.int32)
    lut[ukeys] = np.arange(len(ukeys), dtype=np.int32)
    gid = np.where(mask, lut[key], -1).astype(np.int32)
    got = images.numpy().reshape(-1, 133)
    gc = counts.numpy("int64")[:ng]
    assert np.array_equal(gc, np.bincount(gid[gid >= 0], minlength=ng))
    nulls = [~a_ok, ~(a_ok & b_ok), ~(a_ok & b_ok) | b_zero]
    for k, text in enumerate(exprs):
        e = U.Expr(text, COLS)
        g = np.where(nulls[k], -1, gid).astype(np.int32)
        want, _ = O.plan_group_sum(O.Plan(e.plan), [ga.reshape(-1), gb.reshape(-1)], [PA, PB], [SA, SB], rows, g, ng)
        for gg in range(ng):
            assert bytes(got[gg * len(exprs) + k]) == bytes(want[gg]), (text, gg)


def test_agg_iter_refuses_avg_over_nullable(U, gpu):
    rows = 1000
    cols, *_ = _setup(U, 45, rows)
    agg = U.Agg(["a"], COLS, avg=(0,))
    images = U.DeviceBuffer(2 * 133).zero()
    counts = U.DeviceBuffer(8).zero()
    with pytest.raises(U.UpdecError, match="AVG over columns that can be NULL"):
        agg.run_iter(cols, rows, images, counts, max_groups=1)
