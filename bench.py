#!/usr/bin/env python3
"""Throughput of the MI355X decimal expression + SUM path on one GPU.

Headline workload: grouped aggregation of the TPC-H Q1 shape, 60,000,000
rows (seed 2): filter ship <= 2435, group by (flag1, flag2), SUM(qty),
SUM(price), SUM(price*(1-disc)), SUM(price*(1-disc)*(1+tax)), AVG(qty),
AVG(price), AVG(disc), COUNT(*).  One step = one full pass (the fused
filter/key-hash/expression/SUM kernel, the device merge in key order, the
group-count readback, the three AVG divisions) over the device-resident
columns.

Also reported: the plain-method workload (1,000,000 rows, SUM(p*(1-d)) and
SUM(q*p)), the precision sweep ((a+b)*c - a/d over 10,000,000 rows at
P = 18, 38, 76, 150, 300), the high-precision series (sin(x) by a 25-term
Horner series over 1,000,000 DECIMAL(300,100) rows, then SUM), the batched
serving loop, the compact codec (text ingest and output formatting of
10,000,000 DECIMAL(15,2) rows) and the RSA workload (col1*col1%N*col1%N over
10,000,000 rows for four key lengths), each with its kernel time and bytes.

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
METRIC = "Mrows/s decimal expr+SUM (fraction of HBM roofline); batched vs per-query p99"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def headline(U, W, steps, warmup, cpu_rows, threads):
    C = W.CONFIG3
    rows = C["rows"]
    seed = C["seed"]
    dev = [U.gen_compact_device(W.gen_desc(U, seed, i, c), rows) for i, c in enumerate(C["cols"])]
    f1 = U.gen_byte_device(seed, W.COL_FLAG1, C["flag1"], rows)
    f2 = U.gen_byte_device(seed, W.COL_FLAG2, C["flag2"], rows)
    ship = U.gen_int32_device(seed, W.COL_SHIP, C["ship"][0], C["ship"][1], rows)
    U.device_sync()
    # column statistics are computed once at load time, outside the timed region
    bits = [U.column_maxbits(t, c[1], rows) for t, c in zip(dev, C["cols"])]
    cols = [(c[0], c[1], c[2], b) for c, b in zip(C["cols"], bits)]
    t0 = time.time()
    # AVG(qty), AVG(price), AVG(disc) are divided in the aggregation's final
    # kernel; their images follow the SUM images
    agg = U.Agg(C["exprs"], cols, filter=C["filter"], nkeys=2, avg=(0, 1, 4))
    ne = len(C["exprs"])
    maxg = 16
    images = U.DeviceBuffer(maxg * (ne + 3) * 133).zero()
    counts = U.DeviceBuffer(maxg * 8).zero()

    def step():
        ng, keys = agg.run(dev, rows, images, counts, filter_col=ship, key0=f1, key1=f2, max_groups=maxg)
        return ng

    ng = step()
    U.device_sync()
    compile_s = time.time() - t0
    for _ in range(warmup):
        step()
    U.device_sync()
    kms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
        kms.append(agg.last_kernel_ms())
    U.device_sync()
    total = time.perf_counter() - t0
    ms_per_step = 1000.0 * total / steps
    kernel_ms = sum(kms) / len(kms)
    alg_bytes = 34 * rows  # flag1 + flag2 + ship(4) + 4 x DECIMAL(15,2) (7 B)
    achieved = alg_bytes / (kernel_ms / 1000.0) / 1e9
    res = {
        "rows": rows, "groups": ng, "ms_per_step": ms_per_step, "kernel_ms": kernel_ms,
        "alg_bytes": alg_bytes, "achieved_gbps": achieved, "compile_s": compile_s,
        "column_bits": bits,
    }
    # the timed run's group images, keys and counts against the oracle's
    # full-size results committed with the tests (outside the timed region)
    G = fixtures()["q1"]
    got = images.numpy(count=maxg * (ne + 3) * 133).reshape(-1, 133)
    gcnt = counts.numpy("int64", count=maxg)
    same = ng == len(G["groups"])
    for g, want in enumerate(G["groups"][:ng]):
        same = same and int(gcnt[g]) == want["count"]
        same = same and all(bytes(got[g * ne + k]).hex() == want["sums"][k] for k in range(ne))
        same = same and all(bytes(got[maxg * ne + g * 3 + j]).hex() == want["avgs"][j] for j in range(3))
    res["gpu_equals_fixture"] = bool(same)
    # CPU baseline: the oracle's restatement of the reference's host Decimal
    # path over the same rows (all of them by default), built -march=native on
    # this host; its group images and counts are then compared byte for byte
    # with the GPU's from the last timed step
    if cpu_rows > 0:
        from oracle import oracle as O
        import numpy as np
        native = O.use_native()
        n = min(cpu_rows, rows)
        host = [t.numpy(count=n * U.compact_bytes(c[1])) for t, c in zip(dev, C["cols"])]
        hf1, hf2 = f1.numpy(count=n), f2.numpy(count=n)
        hship = ship.numpy("int32", count=n)
        P = [c[1] for c in cols]
        S = [c[2] for c in cols]
        plans = [O.Plan(U.Expr(t, cols).plan) for t in C["exprs"]]
        t0 = time.perf_counter()
        mask = hship <= C["filter"][1]
        key = (hf1.astype(np.int32) << 8) | hf2
        keys = np.unique(key[mask])
        lut = np.full(65536, -1, dtype=np.int32)
        lut[keys] = np.arange(len(keys), dtype=np.int32)
        gid = np.where(mask, lut[key], -1).astype(np.int32)
        sums = []
        for p in plans:
            img, cnt = O.plan_group_sum(p, host, P, S, n, gid, len(keys), threads=threads)
            sums.append(img)
        avgs = {k: [O.avg_from_sum(bytes(sums[k][g]), int(cnt[g])) for g in range(len(keys))] for k in (0, 1, 4)}
        cpu_s = time.perf_counter() - t0
        model, visible, allowed = O.host_cpu()
        res["cpu"] = {"rows": n, "seconds": cpu_s, "mrows_per_s": n / cpu_s / 1e6, "threads": threads,
                      "cpu_model": model, "cpus_visible": visible, "cpus_allowed": allowed,
                      "build": "gcc -O3 -march=native" if native else "gcc -O3 (portable)"}
        if n == rows:
            got = images.numpy(count=maxg * (ne + 3) * 133).reshape(-1, 133)
            gcnt = counts.numpy("int64", count=maxg)
            same = ng == len(keys) and all(int(gcnt[g]) == int(cnt[g]) for g in range(ng))
            for k in range(ne):
                for g in range(ng):
                    same = same and bytes(got[g * ne + k]) == bytes(sums[k][g])
            for j, k in enumerate((0, 1, 4)):
                for g in range(ng):
                    same = same and bytes(got[maxg * ne + g * 3 + j]) == avgs[k][g]
            res["gpu_equals_cpu_all_rows"] = bool(same)
            if not same:
                log("headline: GPU images differ from the CPU oracle over all rows")
    return res


def headline_general(U, W, steps, warmup):
    """the headline query without column statistics: every value at the
    width its declared precision allows (a caller that passes bits = 0);
    results compared with the statistics form's"""
    C = W.CONFIG3
    rows = C["rows"]
    seed = C["seed"]
    dev = [U.gen_compact_device(W.gen_desc(U, seed, i, c), rows) for i, c in enumerate(C["cols"])]
    f1 = U.gen_byte_device(seed, W.COL_FLAG1, C["flag1"], rows)
    f2 = U.gen_byte_device(seed, W.COL_FLAG2, C["flag2"], rows)
    ship = U.gen_int32_device(seed, W.COL_SHIP, C["ship"][0], C["ship"][1], rows)
    bits = [U.column_maxbits(t, c[1], rows) for t, c in zip(dev, C["cols"])]
    ne, maxg = len(C["exprs"]), 16
    out = {}
    for name, cols in (("statistics", [(c[0], c[1], c[2], b) for c, b in zip(C["cols"], bits)]),
                       ("general", [c[:3] for c in C["cols"]])):
        agg = U.Agg(C["exprs"], cols, filter=C["filter"], nkeys=2, avg=(0, 1, 4))
        img = U.DeviceBuffer(maxg * (ne + 3) * 133).zero()
        cnt = U.DeviceBuffer(maxg * 8).zero()

        def step():
            return agg.run(dev, rows, img, cnt, filter_col=ship, key0=f1, key1=f2, max_groups=maxg)[0]

        for _ in range(warmup + 1):
            step()
        U.device_sync()
        kms = []
        t0 = time.perf_counter()
        for _ in range(steps):
            step()
            kms.append(agg.last_kernel_ms())
        U.device_sync()
        ms = 1000.0 * (time.perf_counter() - t0) / steps
        km = sum(kms) / len(kms)
        out[name] = {"ms_per_step": ms, "kernel_ms": km, "mrows_per_s": rows / ms / 1e3,
                     "achieved_gbps": 34 * rows / (km / 1000.0) / 1e9, "images": bytes(img.numpy()) + bytes(cnt.numpy())}
    same = out["general"].pop("images") == out["statistics"].pop("images")
    r = dict(out["general"], column_bits=0, identical_results_to_statistics_form=same)
    return r


def plain_method(U, W, steps, warmup):
    C = W.CONFIG1
    rows = C["rows"]
    dev = [U.gen_compact_device(W.gen_desc(U, C["seed"], i, c), rows) for i, c in enumerate(C["cols"])]
    U.device_sync()
    cols = [(c[0], c[1], c[2]) for c in C["cols"]]
    agg = U.Agg(C["exprs"], cols)
    images = U.DeviceBuffer(len(C["exprs"]) * 133).zero()
    counts = U.DeviceBuffer(8).zero()
    for _ in range(warmup + 1):
        agg.run(dev, rows, images, counts)
    U.device_sync()
    kms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        agg.run(dev, rows, images, counts)
        kms.append(agg.last_kernel_ms())
    U.device_sync()
    ms = 1000.0 * (time.perf_counter() - t0) / steps
    kernel_ms = sum(kms) / len(kms)
    fx = fixtures()["plain"]["sums"]
    got = images.numpy().reshape(-1, 133)
    same = all(bytes(got[k]).hex() == fx[t]["image"] for k, t in enumerate(C["exprs"]))
    return {"rows": rows, "ms_per_step": ms, "kernel_ms": kernel_ms, "mrows_per_s": rows / ms / 1e3,
            "kernel_mrows_per_s": rows / kernel_ms / 1e3, "alg_bytes": 21 * rows,
            "achieved_gbps": 21 * rows / (kernel_ms / 1000.0) / 1e9, "images_equal_oracle": same}


def fixtures():
    """the oracle's full-size results, computed in advance by
    tests/golden/make_fullsize.py (SUM images, sweep output hashes)"""
    with open(os.path.join(ROOT, "tests", "golden", "fullsize.json")) as f:
        return json.load(f)


def precision_sweep(U, W, steps, rows_override=None):
    import hashlib
    C = W.CONFIG2
    rows = rows_override or C["rows"]
    fx = fixtures()["sweep"]["precisions"] if rows == C["rows"] else {}
    out = {}
    for P, S in C["precisions"]:
        cols = [(n, P, S) for n in "abcd"]
        dev = [U.gen_compact_device(W.digits_desc(U, C["seed"], i, P), rows) for i in range(4)]
        e = U.Expr(C["expr"], cols)
        obuf = U.DeviceBuffer(rows * e.out_bytes + 16)
        n = steps if P <= 76 else max(1, steps // 3)
        alg = (4 * U.compact_bytes(P) + e.out_bytes) * rows
        lp = e.limb_products

        def timed(lanes):
            e.set_lanes(lanes)
            t0 = time.time()
            e.calc(dev, rows, obuf)
            U.device_sync()
            first_s = time.time() - t0
            ev0, ev1 = U.Event(), U.Event()
            ev0.record()
            for _ in range(n):
                e.calc(dev, rows, obuf)
            ev1.record()
            ms = ev0.elapsed_ms(ev1) / n
            r = {"kernel_ms": ms, "mrows_per_s": rows / ms / 1e3, "achieved_gbps": alg / (ms / 1000.0) / 1e9,
                 "first_call_s": first_s, "limb_products_per_s": lp * rows / (ms / 1000.0)}
            if str(P) in fx:
                digest = hashlib.sha256(obuf.numpy(count=rows * e.out_bytes).tobytes()).hexdigest()
                r["output_sha256_equals_oracle"] = digest == fx[str(P)]["sha256"]
            return r

        r = timed(1)
        r.update({"rows": rows, "alg_bytes": alg, "out_type": [e.precision, e.scale], "faithful": e.faithful,
                  "limb_products_per_row": lp, "lanes_per_row": 1})
        w = e.multi_lane_width
        if w:
            # the multi-lane form: one row over W/4 or W/8 lanes
            r["multi_lane"] = {}
            for lanes in (w // 4, w // 8):
                if lanes <= 64:
                    r["multi_lane"][str(lanes)] = timed(lanes)
                    log(f"sweep P={P} lanes {lanes}:", r["multi_lane"][str(lanes)])
            e.set_lanes(1)
            opts = {"1": r["kernel_ms"], **{k: v["kernel_ms"] for k, v in r["multi_lane"].items()}}
            best = min(opts, key=opts.get)
            r["fastest_lanes_per_row"] = int(best)
            r["fastest_kernel_ms"] = opts[best]
            r["fastest_mrows_per_s"] = rows / opts[best] / 1e3
        out[f"P{P}"] = r
        del dev, obuf
    return out


def series(U, W, steps, cpu_rows, threads):
    """sin(x) by the 25-term Horner series over x DECIMAL(300,100) in
    [-pi, pi], SUM over 1,000,000 rows (one fused aggregation launch)"""
    C = W.CONFIG4
    rows = C["rows"]
    dev = U.gen_compact_device(W.sin_desc(U, C["seed"]), rows)
    bits = U.column_maxbits(dev, 300, rows)
    cols = [("x", 300, 100, bits)]
    t0 = time.time()
    agg = U.Agg([W.sin_expr()], cols)
    images = U.DeviceBuffer(133).zero()
    counts = U.DeviceBuffer(8).zero()
    agg.run([dev], rows, images, counts, max_groups=1)
    U.device_sync()
    first_s = time.time() - t0
    kms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        agg.run([dev], rows, images, counts, max_groups=1)
        kms.append(agg.last_kernel_ms())
    U.device_sync()
    ms = 1000.0 * (time.perf_counter() - t0) / steps
    kernel_ms = sum(kms) / len(kms)
    lp = U.Expr(W.sin_expr(), cols).limb_products
    res = {"rows": rows, "column_bits": bits, "ms_per_step": ms, "kernel_ms": kernel_ms,
           "mrows_per_s": rows / ms / 1e3, "kernel_mrows_per_s": rows / kernel_ms / 1e3,
           "first_call_s": first_s, "alg_bytes": rows * U.compact_bytes(300), "limb_products_per_row": lp,
           "limb_products_per_s": lp * rows / (kernel_ms / 1000.0),
           "image_equals_oracle": bytes(images.numpy()).hex() == fixtures()["series"]["image"]}
    if cpu_rows > 0:
        from oracle import oracle as O
        n = min(cpu_rows, rows)
        host = dev.numpy(count=n * U.compact_bytes(300))
        plan = O.Plan(U.Expr(W.sin_expr(), cols).plan)
        t0 = time.perf_counter()
        O.plan_sum(plan, [host], [300], [100], n, mode=O.EXTENDED, threads=threads)
        cpu_s = time.perf_counter() - t0
        res["cpu"] = {"rows": n, "seconds": cpu_s, "mrows_per_s": n / cpu_s / 1e6, "threads": threads,
                      "kind": "port"}
    return res


def codec(U, W, steps, cpu_rows):
    """the compact codec at the edges over 10,000,000 rows of the Q1 price
    column, DECIMAL(15,2): format as decimal text and as GetDecimal hex,
    and ingest the packed text column back (one lane per row each)"""
    import numpy as np
    rows = 10_000_000
    P, S = 15, 2
    spec = W.CONFIG3["cols"][1]
    B = U.compact_bytes(P)
    col = U.gen_compact_device(W.gen_desc(U, W.CONFIG3["seed"], 1, spec), rows)
    out = {}
    ev0, ev1 = U.Event(), U.Event()
    slots = None
    for name, form in (("format_text", U.FORMAT_TEXT), ("format_hex", U.FORMAT_HEX)):
        stride = U.format_stride(P, S, form)
        buf = U.DeviceBuffer(rows * stride)
        U.format_compact(col, rows, P, S, form, buf, stride)
        ev0.record()
        for _ in range(steps):
            U.format_compact(col, rows, P, S, form, buf, stride)
        ev1.record()
        ms = ev0.elapsed_ms(ev1) / steps
        alg = rows * (B + stride)
        out[name] = {"rows": rows, "stride": stride, "kernel_ms": ms, "mrows_per_s": rows / ms / 1e3,
                     "alg_bytes": alg, "achieved_gbps": alg / ms / 1e6}
        if form == U.FORMAT_TEXT:
            slots = buf.numpy(count=rows * stride).reshape(rows, stride)
        del buf
    # packed text column (no padding): the text rows of the formatted slots
    lens = (slots != 0).sum(axis=1).astype(np.int64)
    chars = slots[slots != 0]
    offs = np.zeros(rows + 1, dtype=np.int64)
    offs[1:] = np.cumsum(lens)
    dch = U.DeviceBuffer.from_numpy(chars)
    doff = U.DeviceBuffer.from_numpy(offs)
    back = U.DeviceBuffer(rows * B)
    U.parse_compact(dch, doff, rows, P, S, back)
    U.device_sync()
    same = bool(np.array_equal(back.numpy(), col.numpy(count=rows * B)))
    ev0.record()
    for _ in range(steps):
        U.parse_compact(dch, doff, rows, P, S, back)
    ev1.record()
    ms = ev0.elapsed_ms(ev1) / steps
    alg = int(chars.size) + 8 * (rows + 1) + rows * B
    out["parse_text"] = {"rows": rows, "text_bytes": int(chars.size), "kernel_ms": ms, "mrows_per_s": rows / ms / 1e3,
                         "alg_bytes": alg, "achieved_gbps": alg / ms / 1e6, "round_trip_identical": same}
    if cpu_rows > 0:
        from oracle import oracle as O
        n = min(cpu_rows, rows)
        texts = [bytes(chars[offs[i]:offs[i + 1]]) for i in range(n)]
        t0 = time.perf_counter()
        O.parse_compact(texts, P, S)
        cpu_s = time.perf_counter() - t0
        out["parse_text"]["cpu"] = {"rows": n, "seconds": cpu_s, "mrows_per_s": n / cpu_s / 1e6, "threads": 1,
                                    "kind": "port"}
    return out


def segsum_bench(U, W, steps, cpu_rows, threads):
    """the segmented SUM (aggregate_column over `associated`/`groups`): the
    Q1 price column, 60,000,000 rows of DECIMAL(15,2), in the order a sort
    by (flag1, flag2) gives (associated = the stably sorted row numbers,
    groups = the six segment starts), and the same column in row order cut
    into 1,000,000 segments of 60 rows"""
    import numpy as np
    C = W.CONFIG3
    rows = C["rows"]
    P, S = 15, 2
    B = U.compact_bytes(P)
    col = U.gen_compact_device(W.gen_desc(U, C["seed"], 1, C["cols"][1]), rows)
    f1 = U.gen_byte_host(C["seed"], W.COL_FLAG1, C["flag1"], rows)
    f2 = U.gen_byte_host(C["seed"], W.COL_FLAG2, C["flag2"], rows)
    key = (f1.astype(np.int32) << 8) | f2
    assoc = np.argsort(key, kind="stable").astype(np.int32)
    skey = key[assoc]
    starts = np.flatnonzero(np.concatenate([[True], skey[1:] != skey[:-1]])).astype(np.int32)
    del key, skey
    cases = {"by_key_6_segments": (U.DeviceBuffer.from_numpy(assoc), starts),
             "row_order_1M_segments": (None, (np.arange(1_000_000, dtype=np.int64) * (rows // 1_000_000)).astype(np.int32))}
    out = {}
    ev0, ev1 = U.Event(), U.Event()
    for name, (da, groups) in cases.items():
        dg = U.DeviceBuffer.from_numpy(groups)
        img = U.DeviceBuffer(len(groups) * 133)
        U.segsum_compact(col, P, S, rows, da, dg, len(groups), img)
        U.device_sync()
        ev0.record()
        for _ in range(steps):
            U.segsum_compact(col, P, S, rows, da, dg, len(groups), img)
        ev1.record()
        ms = ev0.elapsed_ms(ev1) / steps
        alg = rows * (B + (4 if da is not None else 0)) + len(groups) * (4 + 133)
        r = {"rows": rows, "segments": len(groups), "ms": ms, "mrows_per_s": rows / ms / 1e3, "alg_bytes": alg,
             "achieved_gbps": alg / ms / 1e6}
        if cpu_rows > 0 and name.startswith("by_key"):
            from oracle import oracle as O
            hcol = col.numpy(count=rows * B)
            t0 = time.perf_counter()
            want = O.segsum_compact(hcol, P, S, rows, assoc, groups, threads=threads)
            cpu_s = time.perf_counter() - t0
            r["cpu"] = {"rows": rows, "seconds": cpu_s, "mrows_per_s": rows / cpu_s / 1e6, "threads": threads,
                        "kind": "port"}
            r["gpu_equals_cpu_all_rows"] = bool(np.array_equal(img.numpy().reshape(-1, 133), want))
        out[name] = r
        del dg, img
    return out


RSA_KEYS = [
    (17, "514679589732349351"),
    (35, "331972734823750661745273085756461571"),
    (71, "273420351015258185841522860354335753665897876941082396543621156533322691"),
    (143, "1686878444203618142189652405784181065581460450325437359510561733117064525110066314054980898011446"
          "29119144418134762919565200448008667563331468801"),
]


def rsa(U, W, steps, cpu_rows, threads):
    """the RSA workload: col1 * col1 % N * col1 % N over 10,000,000 rows of
    DECIMAL(P,0) integers with P digits, for the four key lengths"""
    rows = 10_000_000
    out = {}
    for P, N in RSA_KEYS:
        col = U.gen_compact_device(W.digits_desc(U, 8, 0, P, neg_permille=0), rows)
        e = U.Expr(f"col1 * col1 % {N} * col1 % {N}", [("col1", P, 0)])
        obuf = U.DeviceBuffer(rows * e.out_bytes + 16)
        t0 = time.time()
        e.calc([col], rows, obuf)
        U.device_sync()
        first_s = time.time() - t0
        ev0, ev1 = U.Event(), U.Event()
        ev0.record()
        for _ in range(steps):
            e.calc([col], rows, obuf)
        ev1.record()
        ms = ev0.elapsed_ms(ev1) / steps
        alg = rows * (U.compact_bytes(P) + e.out_bytes)
        lp = e.limb_products
        r = {"rows": rows, "kernel_ms": ms, "mrows_per_s": rows / ms / 1e3, "alg_bytes": alg,
             "achieved_gbps": alg / ms / 1e6, "out_type": [e.precision, e.scale], "faithful": e.faithful,
             "first_call_s": first_s, "limb_products_per_row": lp, "limb_products_per_s": lp * rows / (ms / 1000.0)}
        if cpu_rows > 0:
            from oracle import oracle as O
            n = min(cpu_rows, rows)
            host = col.numpy(count=n * U.compact_bytes(P))
            t0 = time.perf_counter()
            O.plan_eval(O.Plan(e.plan), [host], [P], [0], n, e.out_bytes, threads=threads)
            cpu_s = time.perf_counter() - t0
            r["cpu"] = {"rows": n, "seconds": cpu_s, "mrows_per_s": n / cpu_s / 1e6, "threads": threads, "kind": "port"}
        out[f"LEN_{4 * (1 << RSA_KEYS.index((P, N)))}"] = r
        del col, obuf
    return out


def serving_bench(U, W, nq, rates, reserve_gb, jit_queries=4, saturated=True):
    """the batched serving workload: nq seeded queries at the stated
    distribution (every query as drawn: programs of up to 512 limbs run on
    the interpreter's multi-lane classes), each on its own fresh columns
    generated on the device when the query is admitted (admission by free
    device memory, hipMemGetInfo) and freed when it completes; saturated
    (all queued at once) and at Poisson arrival rates, batched (HIP graph)
    against one plain launch per query"""
    import numpy as np
    from ultraprecise_amd import serving
    from ultraprecise_amd import autotune
    C = W.CONFIG5
    t0 = time.time()
    tuning = autotune.tune(U, use_cache=False)
    tune_s = time.time() - t0
    log("serving: autotuned (rows per item, blocks per CU[, lanes]) per width class:", tuning, f"in {tune_s:.1f} s")
    t0 = time.time()
    qs = [W.random_query(C["seed"], i) for i in range(nq)]
    progs = [U.VmProgram(q["text"], q["cols"]) for q in qs]
    setup_s = time.time() - t0
    imgs = {m: U.DeviceBuffer(133 * nq).zero() for m in ("batched", "plain")}
    cnts = {m: U.DeviceBuffer(8 * nq).zero() for m in ("batched", "plain")}
    out = {"queries": nq, "stream": f"the first {nq:,} queries of the seeded {C['queries']:,}-query stream",
           "rows": sum(q["rows"] for q in qs), "setup_s": setup_s, "tune_s": tune_s,
           "tuning": {str(U.VM_CLASS_WIDTHS[k]): v for k, v in tuning.items()},
           "admission": f"fresh columns per query, admitted while they fit in free device memory less "
                        f"{reserve_gb} GiB (hipMemGetInfo)",
           "width_classes": {w: sum(1 for p in progs if p.width == w) for w in U.VM_CLASS_WIDTHS},
           "rows_by_width_class": {w: sum(q["rows"] for q, p in zip(qs, progs) if p.width == w)
                                   for w in U.VM_CLASS_WIDTHS}}
    zero = np.zeros(nq)
    sat = {}
    for mode in (("plain", "batched") if saturated else ()):
        store = serving.FreshColumns(U, C["seed"], reserve_bytes=reserve_gb << 30)
        srv = serving.Server(U, store, batched=(mode == "batched"), tuning=tuning)
        st = srv.run(qs, zero, imgs[mode], cnts[mode])
        sat[mode] = {k: v for k, v in st.items() if k != "latency_s"}
        sat[mode]["peak_column_gb"] = store.peak / 2 ** 30
        log(f"serving saturated {mode}:", sat[mode])
    same = bytes(imgs["batched"].numpy()) == bytes(imgs["plain"].numpy()) if saturated else None
    # the drop-in JIT path, one query at a time (hipRTC compile of the fused
    # SUM kernel, then one launch), on the first queries of the stream
    njit = min(jit_queries, nq) if saturated else 0
    if njit:
        got = imgs["batched"].numpy().reshape(nq, 133)
        jimg = U.DeviceBuffer(133)
        jcnt = U.DeviceBuffer(8)
        store = serving.FreshColumns(U, C["seed"], reserve_bytes=reserve_gb << 30)
        t1 = time.perf_counter()
        jit_same = True
        for i in range(njit):
            q = qs[i]
            cols = store.acquire_query(q, i)
            agg = U.Agg([q["text"]], q["cols"])
            agg.run(cols, q["rows"], jimg, jcnt, max_groups=1)
            jit_same = jit_same and bytes(jimg.numpy()) == bytes(got[i])
            store.release_query(q, i)
        jit_s = time.perf_counter() - t1
        jrows = sum(q["rows"] for q in qs[:njit])
        sat["plain_jit"] = {"queries": njit, "rows": jrows, "seconds": jit_s, "mrows_per_s": jrows / jit_s / 1e6,
                            "identical_results": jit_same}
        log("serving saturated plain_jit (first queries):", sat["plain_jit"])
    out["saturated"] = sat
    if saturated:
        out["saturated_speedup_vs_plain_interpreter"] = sat["batched"]["mrows_per_s"] / sat["plain"]["mrows_per_s"]
    if "plain_jit" in sat:
        out["saturated_speedup_vs_plain_jit"] = sat["batched"]["mrows_per_s"] / sat["plain_jit"]["mrows_per_s"]
    out["identical_results"] = same
    pois = {}
    for rate in rates:
        arr = serving.poisson_arrivals(nq, rate, seed=C["seed"] + rate)
        pois[rate] = {}
        for mode in ("batched", "plain"):
            store = serving.FreshColumns(U, C["seed"], reserve_bytes=reserve_gb << 30)
            srv = serving.Server(U, store, batched=(mode == "batched"), tuning=tuning)
            st = srv.run(qs, arr, imgs[mode], cnts[mode])
            pois[rate][mode] = {k: v for k, v in st.items() if k != "latency_s"}
            log(f"serving {rate}/s {mode}:", pois[rate][mode])
        pois[rate]["identical_results"] = bytes(imgs["batched"].numpy()) == bytes(imgs["plain"].numpy())
    out["poisson"] = pois
    return out


def copy_bandwidth(U, nbytes=1 << 30, reps=5):
    a = U.DeviceBuffer(nbytes)
    b = U.DeviceBuffer(nbytes)
    U.copy_kernel(b, a, nbytes)
    U.device_sync()
    ev0, ev1 = U.Event(), U.Event()
    ev0.record()
    for _ in range(reps):
        U.copy_kernel(b, a, nbytes)
    ev1.record()
    ms = ev0.elapsed_ms(ev1) / reps
    return {"bytes_moved": 2 * nbytes, "ms": ms, "gbps": 2 * nbytes / (ms / 1000.0) / 1e9}


def load_traffic():
    """HBM bytes per launch of the fused kernel from the committed rocprofv3
    PMC summary (FETCH_SIZE + WRITE_SIZE scaled by the copy-kernel
    calibration), when present."""
    path = os.path.join(ROOT, "profiles", "agg_keys_main_q1_60M_pmc.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            try:
                return float(line.split(":", 1)[1].strip())
            except ValueError:
                return None
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-rows", type=int, default=60_000_000,
                    help="rows of the CPU baseline (all 60,000,000 by default; then also compared with the GPU)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-extra", action="store_true", help="headline only")
    ap.add_argument("--extra", default="general,copy,plain,sweep,series,serving,codec,rsa,segsum",
                    help="comma list of extra workloads")
    ap.add_argument("--sweep-rows", type=int, default=0)
    ap.add_argument("--series-cpu-rows", type=int, default=50_000, help="rows of the series CPU sample")
    ap.add_argument("--codec-cpu-rows", type=int, default=1_000_000, help="rows of the text-ingest CPU sample")
    ap.add_argument("--rsa-cpu-rows", type=int, default=200_000, help="rows of the RSA CPU sample (per key length)")
    ap.add_argument("--serve-queries", type=int, default=1_000,
                    help="queries per serving run: the first N of the seeded 10,000-query stream (the whole "
                         "stream takes about 150 s per run on one MI355X)")
    ap.add_argument("--serve-rates", default="200,400,800")
    ap.add_argument("--serve-saturated", type=int, default=1, help="0: only the Poisson-rate runs")
    ap.add_argument("--serve-reserve-gb", type=int, default=8,
                    help="device memory kept free by the serving admission (GiB)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported: the path runs on one "
              "GPU here and has no multi-GPU form", file=sys.stderr)
        sys.exit(2)
    import ultraprecise_amd as U
    from ultraprecise_amd import workloads as W
    U.init(0)
    info = U.device_info()
    # the CPU leg's threads: every CPU this process may run on, up to the CPU
    # share the GPU host gives one GPU (OMP_NUM_THREADS there)
    try:
        allowed = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        allowed = os.cpu_count() or 1
    share = int(os.environ.get("OMP_NUM_THREADS") or allowed)
    threads = args.cpu_threads or max(1, min(allowed, share))
    log("device:", info)
    h = headline(U, W, args.steps, args.warmup, args.cpu_rows, threads)
    log("headline:", h)
    extra = {}
    want = set() if args.no_extra else set(args.extra.split(","))
    if "general" in want:
        extra["headline_without_column_statistics"] = headline_general(U, W, max(5, args.steps // 2), args.warmup)
        log("headline without column statistics:", extra["headline_without_column_statistics"])
    if "copy" in want:
        extra["copy"] = copy_bandwidth(U)
        log("copy:", extra["copy"])
        ms, macs, rate = U.mad_peak()
        extra["mad_peak"] = {"ms": ms, "multiply_adds": macs, "multiply_adds_per_s": rate,
                             "note": "32x32->64 multiply-add chains (v_mad_u64_u32), 8 independent per lane: the "
                                     "measured issue rate the limb_products_per_s figures compare with"}
        log("mad peak:", extra["mad_peak"])
    if "plain" in want:
        extra["plain_1M"] = plain_method(U, W, args.steps, args.warmup)
        log("plain:", extra["plain_1M"])
    if "sweep" in want:
        extra["sweep_10M"] = precision_sweep(U, W, max(3, args.steps // 4), args.sweep_rows or None)
        log("sweep:", extra["sweep_10M"])
    if "series" in want:
        extra["series_1M"] = series(U, W, max(2, args.steps // 5), args.series_cpu_rows, threads)
        log("series:", extra["series_1M"])
    if "rsa" in want:
        extra["rsa_10M"] = rsa(U, W, max(3, args.steps // 4), args.rsa_cpu_rows, threads)
        log("rsa:", extra["rsa_10M"])
    if "segsum" in want:
        extra["segsum_60M"] = segsum_bench(U, W, max(3, args.steps // 4), args.cpu_rows, threads)
        log("segsum:", extra["segsum_60M"])
    if "codec" in want:
        extra["codec_10M"] = codec(U, W, max(3, args.steps // 4), args.codec_cpu_rows)
        log("codec:", extra["codec_10M"])
    if "serving" in want:
        rates = [int(r) for r in args.serve_rates.split(",")]
        extra["serving"] = serving_bench(U, W, args.serve_queries, rates, args.serve_reserve_gb,
                                         saturated=bool(args.serve_saturated))
    value = h["rows"] / h["ms_per_step"] / 1e3  # Mrows/s
    achieved = h["achieved_gbps"]
    cpu = h.get("cpu")
    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Mrows/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(h["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (seeded counter-based generator, device resident)",
        "config": {
            "workload": "grouped aggregation, TPC-H Q1 shape: 60,000,000 rows, seed 2, filter ship<=2435, "
                        "group by (flag1,flag2), 4 SUM + 3 AVG + COUNT over DECIMAL(15,2)",
            "rows": h["rows"], "groups": h["groups"], "column_bits": h["column_bits"],
        },
        "roofline": {
            "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": load_traffic(),
            "kernel": "agg_keys_main", "kernel_ms": round(h["kernel_ms"], 4), "alg_bytes_per_launch": h["alg_bytes"],
        },
        "cpu_baseline": None if cpu is None else {
            "value": round(cpu["mrows_per_s"], 4), "unit": "Mrows/s", "cores": cpu["threads"], "kind": "port",
            "sample": (f"all {h['rows']:,} rows" if cpu["rows"] == h["rows"] else
                       f"first {cpu['rows']:,} of {h['rows']:,} rows") +
                      f", same query (4 SUM + 3 AVG + COUNT), oracle restatement of the reference's host Decimal "
                      f"({cpu['build']}), {cpu['threads']} threads on {cpu['cpu_model']} "
                      f"({cpu['cpus_allowed']} of {cpu['cpus_visible']} CPUs allowed; the host's share for one GPU "
                      f"is the thread count)",
            "cpu_model": cpu["cpu_model"],
        },
        "exact": {"headline_gpu_equals_cpu_all_rows": h.get("gpu_equals_cpu_all_rows"),
                  "headline_gpu_equals_full_size_fixture": h.get("gpu_equals_fixture")},
        "device": info["name"],
        "extra": extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
