"""Write tests/golden/fullsize.json: the oracle's results for the workloads
at their full sizes, so the GPU tests and bench.py can check the HIP path's
full-size outputs byte for byte without running the oracle there.

Inputs come from the library's counter-based generator on the host (the
same bytes the device generator writes), plans from the independent
restatement of the reference's rewrite (tests/ref_rewrite.py, not the
product planner), values from the C oracle (oracle/updec_oracle.c).

  plain     1,000,000 rows, seed 0: SUM(p*(1-d)), SUM(q*p)
  q1        60,000,000 rows, seed 2: per (flag1, flag2) group in ascending
            key order, COUNT(*), the five SUM images and AVG(qty),
            AVG(price), AVG(disc)
  sweep     10,000,000 rows, seed 1, (a+b)*c - a/d at each P: the output
            type and the SHA-256 of the whole materialized output column
  series    1,000,000 rows, seed 3: the SUM image of the sin(x) series

Run from the repository root (about 10 minutes on 8 threads):
    python tests/golden/make_fullsize.py [--threads N] [--only plain,q1,...]
"""
import argparse
import hashlib
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import numpy as np  # noqa: E402

import ref_rewrite as R  # noqa: E402
import ultraprecise_amd as U  # noqa: E402
from oracle import oracle as O  # noqa: E402
from ultraprecise_amd import workloads as W  # noqa: E402

OUT = os.path.join(HERE, "fullsize.json")


def plan(text, cols, mode):
    body, P, S = R.plan(text, cols, R.FAITHFUL if mode == O.FAITHFUL else R.EXTENDED)
    return O.Plan(body), P, S, O.lib.od_compact_bytes(P)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def plain(threads):
    C = W.CONFIG1
    rows = C["rows"]
    spec = C["cols"]
    cols = [(c[0], c[1], c[2]) for c in spec]
    host = [U.gen_compact_host(W.gen_desc(U, C["seed"], i, c), rows) for i, c in enumerate(spec)]
    out = {"rows": rows, "mode": "faithful", "sums": {}}
    for text in C["exprs"]:
        p, P, S, B = plan(text, cols, O.FAITHFUL)
        img, n = O.plan_sum(p, host, [c[1] for c in cols], [c[2] for c in cols], rows, threads=threads)
        assert n == rows
        out["sums"][text] = {"type": [P, S], "image": img.hex()}
    return out


def q1(threads):
    C = W.CONFIG3
    rows = C["rows"]
    seed = C["seed"]
    cols = [(c[0], c[1], c[2]) for c in C["cols"]]
    host = [U.gen_compact_host(W.gen_desc(U, seed, i, c), rows) for i, c in enumerate(C["cols"])]
    f1 = U.gen_byte_host(seed, W.COL_FLAG1, C["flag1"], rows)
    f2 = U.gen_byte_host(seed, W.COL_FLAG2, C["flag2"], rows)
    ship = U.gen_int32_host(seed, W.COL_SHIP, C["ship"][0], C["ship"][1], rows)
    assert C["filter"][0] == "<="
    mask = ship <= C["filter"][1]
    key = (f1.astype(np.int32) << 8) | f2
    keys = np.unique(key[mask])
    lut = np.full(65536, -1, dtype=np.int32)
    lut[keys] = np.arange(len(keys), dtype=np.int32)
    gid = np.where(mask, lut[key], -1).astype(np.int32)
    del f1, f2, ship, key, mask
    P = [c[1] for c in cols]
    S = [c[2] for c in cols]
    sums, cnt = [], None
    for text in C["exprs"]:
        p, _, _, _ = plan(text, cols, O.FAITHFUL)
        img, c = O.plan_group_sum(p, host, P, S, rows, gid, len(keys), threads=threads)
        cnt = c if cnt is None else cnt
        assert np.array_equal(cnt, c)
        sums.append(img)
    groups = []
    for g, k in enumerate(keys):
        groups.append({
            "key": [int(k) >> 8, int(k) & 255],
            "count": int(cnt[g]),
            "sums": [bytes(sums[e][g]).hex() for e in range(len(C["exprs"]))],
            "avgs": [O.avg_from_sum(bytes(sums[e][g]), int(cnt[g])).hex() for e in (0, 1, 4)],
        })
    return {"rows": rows, "mode": "faithful", "exprs": C["exprs"], "avg_of": [0, 1, 4], "groups": groups}


def sweep(threads, chunk=1_000_000):
    C = W.CONFIG2
    rows = C["rows"]
    out = {"rows": rows, "expr": C["expr"], "precisions": {}}
    for P, S in C["precisions"]:
        mode = O.FAITHFUL if P <= 76 else O.EXTENDED
        cols = [(n, P, S) for n in "abcd"]
        p, OP, OS, OB = plan(C["expr"], cols, mode)
        h = hashlib.sha256()
        t0 = time.time()
        for r0 in range(0, rows, chunk):
            n = min(chunk, rows - r0)
            host = [U.gen_compact_host(W.digits_desc(U, C["seed"], i, P), n, row0=r0) for i in range(4)]
            col = O.plan_eval(p, host, [P] * 4, [S] * 4, n, OB, mode=mode, threads=threads)
            h.update(col.tobytes())
        out["precisions"][str(P)] = {"scale": S, "mode": "faithful" if mode == O.FAITHFUL else "extended",
                                     "out_type": [OP, OS], "out_bytes": OB, "sha256": h.hexdigest()}
        log(f"sweep P={P}: {time.time() - t0:.1f} s")
    return out


def series(threads):
    C = W.CONFIG4
    rows = C["rows"]
    host = U.gen_compact_host(W.sin_desc(U, C["seed"]), rows)
    cols = [("x", 300, 100)]
    p, P, S, B = plan(W.sin_expr(), cols, O.EXTENDED)
    img, n = O.plan_sum(p, [host], [300], [100], rows, mode=O.EXTENDED, threads=threads)
    assert n == rows
    return {"rows": rows, "mode": "extended", "type": [P, S], "image": img.hex()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 8)
    ap.add_argument("--only", default="plain,q1,sweep,series")
    args = ap.parse_args()
    data = json.load(open(OUT)) if os.path.exists(OUT) else {}
    for name in args.only.split(","):
        t0 = time.time()
        data[name] = {"plain": plain, "q1": q1, "sweep": sweep, "series": series}[name](args.threads)
        log(f"{name}: {time.time() - t0:.1f} s")
        with open(OUT, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
            f.write("\n")


if __name__ == "__main__":
    main()
