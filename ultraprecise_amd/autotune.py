"""Autotuner for the interpreter's persistent kernels.

Per register-width class (W = 2 .. 32 limbs one value per lane, W = 64 ..
512 limbs one value per group of lanes) it times representative queries of
that width -- seeded serving queries whose programs land in the class --
over every candidate (rows per work item, resident blocks per compute unit,
and for the multi-lane classes lanes per value: W / 4 or W / 8, at most 64)
and keeps the fastest.  Results are cached per device name in a JSON file
so a process pays the search once.
"""
import json
import os

DEFAULT_CANDIDATES = ((1024, 4), (2048, 4), (4096, 2), (4096, 4), (8192, 4), (8192, 8), (16384, 4), (32768, 4))
MULTI_LANE_ITEMS = ((128, 4), (256, 4), (512, 4), (512, 8), (1024, 4), (1024, 8), (2048, 4))


def class_rows(width, rows):
    """rows timed per query of a class: fewer for the costly wide classes"""
    return rows if width <= 32 else max(20_000, rows * 32 // (width * 8))


def candidates_for(U, cls, candidates):
    w = U.VM_CLASS_WIDTHS[cls]
    if cls < U.VM_FIRST_MULTI_LANE:
        return [(rpi, bpc) for rpi, bpc in candidates]
    lanes = sorted({min(64, w // 4), w // 8})
    return [(rpi, bpc, ln) for rpi, bpc in MULTI_LANE_ITEMS for ln in lanes if ln <= 64]


def cache_path(device_name):
    base = os.environ.get("UPDEC_TUNE_DIR", os.path.join(os.path.expanduser("~"), ".cache", "ultraprecise_amd"))
    safe = "".join(ch if ch.isalnum() else "_" for ch in device_name)
    return os.path.join(base, f"vm_tuning_{safe}.json")


def representative_queries(U, W, per_class=2, seed=77, max_draws=12000):
    """seeded serving queries, `per_class` of each width class (as drawn)"""
    found = {c: [] for c in range(len(U.VM_CLASS_WIDTHS))}
    for i in range(max_draws):
        q = W.random_query(seed, i)
        try:
            cls = U.VmProgram(q["text"], q["cols"]).width_class
        except U.UpdecError:
            continue
        if len(found[cls]) < per_class:
            found[cls].append(q)
        if all(len(v) >= per_class for v in found.values()):
            break
    return found


def tune(U, rows=2_000_000, candidates=DEFAULT_CANDIDATES, reps=3, use_cache=True):
    """returns {class: (rows_per_item, blocks_per_cu[, lanes per value])}"""
    from . import workloads as W
    name = U.device_info()["name"]
    path = cache_path(name)
    if use_cache and os.path.exists(path):
        with open(path) as f:
            return {int(k): tuple(v) for k, v in json.load(f).items()}
    reps = max(1, reps)
    result = {}
    for cls, qs in representative_queries(U, W).items():
        if not qs:
            continue
        progs, cols = [], []
        crows = class_rows(U.VM_CLASS_WIDTHS[cls], rows)
        for q in qs:
            progs.append(U.VmProgram(q["text"], q["cols"]))
            cols.append([U.gen_compact_device(U.gen_desc(U.GEN_DIGITS, q["P"], 77, 900 + k, digits=q["P"],
                                                         neg_permille=500), crows) for k in range(len(q["cols"]))])
        img = U.DeviceBuffer(133 * len(qs))
        cnt = U.DeviceBuffer(8 * len(qs))
        best = None
        for cand in candidates_for(U, cls, candidates):
            b = U.Batch(max_queries=len(qs), max_items=1 << 16)
            b.tune(cls, cand[0], cand[1], lanes=cand[2] if len(cand) > 2 else None)

            def launch():
                for k, (p, c) in enumerate(zip(progs, cols)):
                    if b.add(p, c, crows, img, cnt, image_offset=133 * k, count_offset=8 * k) is None:
                        raise RuntimeError("tuning batch overflow")
                b.launch(graph=False)

            launch()
            U.device_sync()
            e0, e1 = U.Event(), U.Event()
            e0.record()
            for _ in range(reps):
                launch()
            e1.record()
            ms = e0.elapsed_ms(e1) / reps
            if best is None or ms < best[0]:
                best = (ms, cand)
        result[cls] = tuple(best[1])
        del cols
    if use_cache:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        with open(path, "w") as f:
            json.dump({str(k): list(v) for k, v in result.items()}, f)
    return result


def tune_expr_lanes(U, expr, columns, rows, out, reps=3, sample_rows=1_000_000):
    """Lanes per row of a JIT calc expression with a multi-lane form: times
    calc over the first `sample_rows` rows at one row per lane and at W/4
    and W/8 lanes per row (at most 64), keeps the fastest on `expr` and
    returns {lanes: ms}.  An expression without the form keeps one row per
    lane."""
    w = expr.multi_lane_width
    if not w:
        return {1: None}
    n = min(rows, sample_rows)
    times = {}
    for lanes in sorted({1, w // 4, w // 8}):
        if lanes > 64:
            continue
        expr.set_lanes(lanes)
        expr.calc(columns, n, out)      # compile and warm up
        U.device_sync()
        ev0, ev1 = U.Event(), U.Event()
        ev0.record()
        for _ in range(reps):
            expr.calc(columns, n, out)
        ev1.record()
        ev1.sync()
        times[lanes] = ev0.elapsed_ms(ev1) / reps
    best = min(times, key=times.get)
    expr.set_lanes(best)
    return times
