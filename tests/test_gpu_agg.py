"""GPU parity for the grouped aggregation (filter, two 1-byte keys, fused
SUMs, COUNT, AVG) and the segmented SUM / AVG entry points."""
import numpy as np
import pytest

from ultraprecise_amd import workloads as W

pytestmark = pytest.mark.gpu


def config3_columns(U, rows, seed):
    C = W.CONFIG3
    dev, host = [], []
    for i, c in enumerate(C["cols"]):
        d = W.gen_desc(U, seed, i, c)
        t = U.gen_compact_device(d, rows)
        dev.append(t)
        host.append(t.numpy(count=rows * U.compact_bytes(c[1])))
    f1 = U.gen_byte_device(seed, W.COL_FLAG1, C["flag1"], rows)
    f2 = U.gen_byte_device(seed, W.COL_FLAG2, C["flag2"], rows)
    ship = U.gen_int32_device(seed, W.COL_SHIP, C["ship"][0], C["ship"][1], rows)
    hf1, hf2, hship = f1.numpy(count=rows), f2.numpy(count=rows), ship.numpy("int32", count=rows)
    return dev, host, (f1, f2, ship), (hf1, hf2, hship)


def host_groups(hf1, hf2, hship, limit):
    mask = hship <= limit
    key = (hf1.astype(np.int32) << 8) | hf2
    keys = np.unique(key[mask])
    gid = np.full(len(key), -1, dtype=np.int32)
    for g, k in enumerate(keys):
        gid[(key == k) & mask] = g
    return keys, gid


@pytest.mark.parametrize("rows", [1000, 262_144, 1_000_000])  # 262,144: whole chunks of 1,024 rows only
def test_config3_grouped_bit_exact(U, O, gpu, rows):
    C = W.CONFIG3
    dev, host, (f1, f2, ship), (hf1, hf2, hship) = config3_columns(U, rows, C["seed"])
    bits = [U.column_maxbits(t, c[1], rows) for t, c in zip(dev, C["cols"])]
    # statistics agree with the host view of the data
    for b, h, c in zip(bits, host, C["cols"]):
        nb = U.compact_bytes(c[1])
        mags = [int.from_bytes(bytes(h[i * nb:(i + 1) * nb]), "little") & ((1 << (8 * nb - 1)) - 1)
                for i in range(0, rows, max(1, rows // 2000))]
        assert max(m.bit_length() for m in mags) <= b
    cols = [(c[0], c[1], c[2], b) for c, b in zip(C["cols"], bits)]
    avg = [0, 1, 4]
    agg = U.Agg(C["exprs"], cols, filter=C["filter"], nkeys=2, avg=avg)
    maxg = 16
    NE = len(C["exprs"])
    images = U.DeviceBuffer(maxg * (NE + len(avg)) * 133).zero()
    counts = U.DeviceBuffer(maxg * 8).zero()
    ng, keys = agg.run(dev, rows, images, counts, filter_col=ship, key0=f1, key1=f2, max_groups=maxg)
    hkeys, gid = host_groups(hf1, hf2, hship, C["filter"][1])
    assert ng == len(hkeys)
    assert keys == [int((k >> 8) | ((k & 0xff) << 8)) for k in hkeys]
    allimg = images.numpy()
    got = allimg[:maxg * NE * 133].reshape(maxg, NE, 133)
    fused = allimg[maxg * NE * 133:].reshape(maxg, len(avg), 133)
    gotc = counts.numpy("int64")[:ng]
    P = [c[1] for c in cols]
    S = [c[2] for c in cols]
    for k, text in enumerate(C["exprs"]):
        e = U.Expr(text, cols)
        want, wc = O.plan_group_sum(O.Plan(e.plan), host, P, S, rows, gid, ng)
        assert np.array_equal(gotc, wc)
        for g in range(ng):
            assert bytes(got[g, k]) == bytes(want[g]), f"SUM({text}) group {g}"
    # AVG(qty), AVG(price), AVG(disc): fused into the aggregation, and
    # through the separate division kernel
    avg2 = U.DeviceBuffer(maxg * 133).zero()
    for j, k in enumerate(avg):
        # contiguous copy of column k's sums
        sums = U.DeviceBuffer.from_numpy(np.ascontiguousarray(got[:ng, k]).reshape(-1))
        U.div_int64(sums, 133, counts, avg2, ng)
        ga2 = avg2.numpy().reshape(maxg, 133)
        for g in range(ng):
            want = O.avg_from_sum(bytes(got[g, k]), int(gotc[g]))
            assert bytes(ga2[g]) == want, f"AVG expr {k} group {g}"
            assert bytes(fused[g, j]) == want, f"fused AVG expr {k} group {g}"


@pytest.mark.parametrize("nf1,nf2,nkeys", [(40, 3, 2), (3, 2, 2), (7, 1, 1), (60, 1, 1)])
def test_grouping_many_keys_and_fallback(U, O, gpu, nf1, nf2, nkeys):
    """more distinct keys than a block's one-pass table holds (16) forces
    the two-pass form; results must not depend on which form ran"""
    rows = 200_000
    seed = 11
    C = W.CONFIG3
    spec = C["cols"][:2]
    dev = [U.gen_compact_device(W.gen_desc(U, seed, i, c), rows) for i, c in enumerate(spec)]
    host = [t.numpy(count=rows * U.compact_bytes(c[1])) for t, c in zip(dev, spec)]
    letters1 = bytes(range(1, 1 + nf1))
    letters2 = bytes(range(60, 60 + nf2))
    f1 = U.gen_byte_device(seed, 1, letters1, rows)
    f2 = U.gen_byte_device(seed, 2, letters2, rows)
    hf1, hf2 = f1.numpy(count=rows), f2.numpy(count=rows)
    cols = [(c[0], c[1], c[2]) for c in spec]
    exprs = ["qty", "price*qty"]
    agg = U.Agg(exprs, cols, nkeys=nkeys, avg=[1])
    maxg = 255
    images = U.DeviceBuffer(maxg * (len(exprs) + 1) * 133).zero()
    counts = U.DeviceBuffer(maxg * 8).zero()
    ng, keys = agg.run(dev, rows, images, counts, key0=f1, key1=f2 if nkeys == 2 else None, max_groups=maxg)
    key = (hf1.astype(np.int32) << 8) | hf2 if nkeys == 2 else hf1.astype(np.int32)
    hkeys = np.unique(key)
    assert ng == len(hkeys)
    lut = np.full(65536, -1, dtype=np.int32)
    lut[hkeys] = np.arange(len(hkeys), dtype=np.int32)
    gid = lut[key].astype(np.int32)
    allimg = images.numpy()
    got = allimg[:maxg * len(exprs) * 133].reshape(maxg, len(exprs), 133)
    fused = allimg[maxg * len(exprs) * 133:].reshape(maxg, 133)
    gotc = counts.numpy("int64")[:ng]
    for k, text in enumerate(exprs):
        e = U.Expr(text, cols)
        want, wc = O.plan_group_sum(O.Plan(e.plan), host, [c[1] for c in cols], [c[2] for c in cols], rows, gid, ng)
        assert np.array_equal(gotc, wc)
        assert np.array_equal(got[:ng, k], want), text
    for g in range(ng):
        assert bytes(fused[g]) == O.avg_from_sum(bytes(got[g, 1]), int(gotc[g])), f"fused AVG group {g}"


def test_repeated_runs_reuse_state(U, O, gpu):
    """one aggregation object run again and again on inputs with different
    key sets, a block-table overflow (two-pass fallback) and a too-small
    max_groups in between: each run's one-pass state is cleared by the
    previous run's final kernel, so every result must match the oracle"""
    spec = W.CONFIG3["cols"][:2]
    cols = [(c[0], c[1], c[2]) for c in spec]
    exprs = ["qty", "price*qty"]
    agg = U.Agg(exprs, cols, nkeys=1, avg=[0])
    maxg = 64
    images = U.DeviceBuffer(maxg * 3 * 133)
    counts = U.DeviceBuffer(maxg * 8)
    cases = [(100_000, bytes([5, 9])), (70_000, bytes([3])), (150_000, bytes(range(1, 41))),
             (90_000, bytes([2, 7, 200])), (30_000, bytes(range(100, 110))), (110_000, bytes([9, 5, 11]))]
    for i, (rows, letters) in enumerate(cases):
        seed = 40 + i
        dev = [U.gen_compact_device(W.gen_desc(U, seed, j, c), rows) for j, c in enumerate(spec)]
        host = [t.numpy(count=rows * U.compact_bytes(c[1])) for t, c in zip(dev, spec)]
        f1 = U.gen_byte_device(seed, 1, letters, rows)
        hf1 = f1.numpy(count=rows).astype(np.int32)
        images.zero()
        counts.zero()
        mg = 8 if i == 4 else maxg  # case 4: 10 groups, no table overflow
        if i == 4:
            with pytest.raises(U.UpdecError, match="max_groups"):
                agg.run(dev, rows, images, counts, key0=f1, max_groups=mg)
            continue
        ng, keys = agg.run(dev, rows, images, counts, key0=f1, max_groups=mg)
        hkeys = np.unique(hf1)
        assert ng == len(hkeys) and keys == [int(k) for k in hkeys], i
        lut = np.full(256, -1, dtype=np.int32)
        lut[hkeys] = np.arange(len(hkeys), dtype=np.int32)
        gid = lut[hf1].astype(np.int32)
        allimg = images.numpy()
        got = allimg[:mg * 2 * 133].reshape(mg, 2, 133)
        fused = allimg[mg * 2 * 133:mg * 3 * 133].reshape(mg, 133)
        gotc = counts.numpy("int64")[:ng]
        for k, text in enumerate(exprs):
            e = U.Expr(text, cols)
            want, wc = O.plan_group_sum(O.Plan(e.plan), host, [c[1] for c in cols], [c[2] for c in cols], rows, gid, ng)
            assert np.array_equal(gotc, wc), i
            assert np.array_equal(got[:ng, k], want), (i, text)
        for g in range(ng):
            assert bytes(fused[g]) == O.avg_from_sum(bytes(got[g, 0]), int(gotc[g])), (i, g)


@pytest.mark.parametrize("mode", ["faithful", "extended"])
def test_div_int64_edges(U, O, gpu, mode):
    """AVG division on crafted sums: zero, negative, path (i) wrap region,
    sums of exactly 8 limbs and wider (the kernel's two division widths),
    zero count"""
    rng = np.random.default_rng(5)
    imgs, cnts = [], []
    for i in range(420):
        d = O.OdDec()
        k = i % 6
        if k == 0:
            m = 0
        elif k == 1:
            m = int(rng.integers(1, 1 << 40))
        elif k == 2:
            m = int(rng.integers(1 << 48, 1 << 51))
        elif k == 3:
            m = int.from_bytes(rng.bytes(40), "little")
        elif k == 4:
            m = int.from_bytes(rng.bytes(32), "little") | (1 << 255)
        else:
            m = int(rng.integers(1 << 60, 1 << 63))
        d.set_magnitude(m)
        d.sign = int(rng.integers(0, 2))
        d.frac = int(rng.integers(0, 8))
        d.prec = 30
        imgs.append(O.image(d))
        cnts.append(0 if i % 97 == 0 else int(rng.integers(1, 1 << 40)))
    n = len(imgs)
    src = U.DeviceBuffer.from_numpy(np.frombuffer(b"".join(imgs), dtype=np.uint8))
    cnt = U.DeviceBuffer.from_numpy(np.array(cnts, dtype=np.int64))
    out = U.DeviceBuffer(n * 133).zero()
    um, om = (U.MODE_FAITHFUL, O.FAITHFUL) if mode == "faithful" else (U.MODE_EXTENDED, O.EXTENDED)
    U.div_int64(src, 133, cnt, out, n, mode=um)
    got = out.numpy().reshape(n, 133)
    for i in range(n):
        assert bytes(got[i]) == O.avg_from_sum(imgs[i], cnts[i], mode=om), i


@pytest.mark.parametrize("mode", ["faithful", "extended"])
@pytest.mark.parametrize("nkeys", [0, 1])
def test_fused_avg_modes(U, O, gpu, mode, nkeys):
    """fused AVG of a wide sum (DECIMAL(38,10) products) in both modes,
    ungrouped and grouped, against the division of the SUM image"""
    rows = 50_000
    seed = 23
    spec = [("a", 38, 10), ("b", 20, 5)]
    dev = [U.gen_compact_device(U.gen_desc(U.GEN_DIGITS, P, seed, i, digits=P - 2, neg_permille=400), rows)
           for i, (n, P, S) in enumerate(spec)]
    f1 = U.gen_byte_device(seed, 7, bytes([3, 9, 12]), rows)
    exprs = ["a", "a*b", "b"]
    um, om = (U.MODE_FAITHFUL, O.FAITHFUL) if mode == "faithful" else (U.MODE_EXTENDED, O.EXTENDED)
    agg = U.Agg(exprs, spec, nkeys=nkeys, mode=um, avg=[0, 1])
    maxg = 4
    images = U.DeviceBuffer(maxg * 5 * 133).zero()
    counts = U.DeviceBuffer(maxg * 8).zero()
    ng, keys = agg.run(dev, rows, images, counts, key0=f1 if nkeys else None, max_groups=maxg)
    assert ng == (3 if nkeys else 1)
    allimg = images.numpy()
    got = allimg[:maxg * 3 * 133].reshape(maxg, 3, 133)
    fused = allimg[maxg * 3 * 133:maxg * 5 * 133].reshape(maxg, 2, 133)
    gotc = counts.numpy("int64")[:ng]
    assert int(gotc.sum()) == rows
    for g in range(ng):
        for j, k in enumerate([0, 1]):
            assert bytes(fused[g, j]) == O.avg_from_sum(bytes(got[g, k]), int(gotc[g]), mode=om), (g, k)


@pytest.mark.parametrize("P,S", [(15, 2), (38, 10), (76, 20), (150, 50), (290, 90)])
@pytest.mark.parametrize("nseg", [40, 20_000])
def test_segsum_compact_bit_exact(U, O, gpu, P, S, nseg):
    """large segments (every tile inside one) and small ones (tiles crossing
    many starts), with empty segments, a segment at the last row and
    positions in a random `associated` order"""
    rng = np.random.default_rng(P + nseg)
    rows = 300_000
    host = U.gen_compact_host(U.gen_desc(U.GEN_DIGITS, P, 9, P, digits=P - 3, neg_permille=500), rows)
    assoc = rng.permutation(rows).astype(np.int32)
    cuts = np.sort(rng.choice(np.arange(1, rows), size=nseg, replace=False))
    groups = np.concatenate([[0], cuts, [rows - 5, rows - 5, rows - 1]]).astype(np.int32)  # includes empty segments
    groups = np.sort(groups)
    col = U.DeviceBuffer.from_numpy(host)
    da = U.DeviceBuffer.from_numpy(assoc)
    dg = U.DeviceBuffer.from_numpy(groups)
    out = U.DeviceBuffer(len(groups) * 133).zero()
    U.segsum_compact(col, P, S, rows, da, dg, len(groups), out)
    got = out.numpy().reshape(-1, 133)
    want = O.segsum_compact(host, P, S, rows, assoc, groups)
    assert np.array_equal(got, want)


def test_segsum_compact_identity_order_and_edges(U, O, gpu):
    """no `associated` (position i is row i); positions before the first
    start belong to no segment; zero rows give +0 images"""
    P, S = 38, 10
    rows = 50_000
    host = U.gen_compact_host(U.gen_desc(U.GEN_DIGITS, P, 11, 2, digits=30, neg_permille=500), rows)
    groups = np.array([100, 100, 4096, 4097, 8191, 30_000], dtype=np.int32)
    col = U.DeviceBuffer.from_numpy(host)
    dg = U.DeviceBuffer.from_numpy(groups)
    out = U.DeviceBuffer(len(groups) * 133).zero()
    U.segsum_compact(col, P, S, rows, None, dg, len(groups), out)
    got = out.numpy().reshape(-1, 133)
    want = O.segsum_compact(host, P, S, rows, None, groups)
    assert np.array_equal(got, want)
    # zero rows: every start is 0, every segment empty
    g0 = np.zeros(4, dtype=np.int32)
    dg0 = U.DeviceBuffer.from_numpy(g0)
    out0 = U.DeviceBuffer(len(g0) * 133).zero()
    U.segsum_compact(col, P, S, 0, None, dg0, len(g0), out0)
    zero = O.segsum_compact(host, P, S, 0, None, g0)
    assert np.array_equal(out0.numpy().reshape(-1, 133), zero)
