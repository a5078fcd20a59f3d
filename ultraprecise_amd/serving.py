"""Serving loop for decimal queries on one GPU.

Queries (an expression over 2..4 compact columns, then SUM) arrive over
time.  Each query is admitted when its columns fit in device memory --
FreshColumns: the query's own columns are generated on the device when the
free device memory (hipMemGetInfo) minus a reserve holds them, and freed
when it completes; ColumnStore: a resident column pool with a byte budget,
least recently used columns not held by a running query evicted -- then
compiled for the precompiled interpreter (host only, no hipRTC), and either

  * batched: every admitted query not yet running goes into the next batch,
    launched as one HIP graph of persistent kernels, with up to two batches
    in flight (the next is staged while the previous runs), or
  * plain: one launch per query, waited for before the next one (the
    reference's pattern: one kernel per expression, synchronous).

Latency is completion time minus arrival time; results are the 133-byte SUM
images and the COUNT of every query.
"""
import collections
import time

import numpy as np


class ColumnStore:
    """Resident compact columns keyed by (precision, column id)."""

    def __init__(self, U, budget_bytes, seed, rows, stream=None):
        self.U = U
        self.budget = int(budget_bytes)
        self.seed = seed
        self.rows = int(rows)
        self.stream = stream
        self.cols = collections.OrderedDict()   # key -> DeviceBuffer (LRU order)
        self.pins = collections.Counter()
        self.used = 0
        self.generated = 0

    def nbytes(self, P):
        return self.rows * self.U.compact_bytes(P) + 16

    def desc(self, key):
        P, cid = key
        return self.U.gen_desc(self.U.GEN_DIGITS, P, self.seed, 5000 + 16 * P + cid, digits=P, neg_permille=500)

    def _make_room(self, need, keep=()):
        for key in list(self.cols):
            if self.used + need <= self.budget:
                break
            if self.pins[key] == 0 and key not in keep:
                buf = self.cols.pop(key)
                self.used -= buf.nbytes
                del buf
        return self.used + need <= self.budget

    def acquire(self, keys):
        """pin the columns of one query, generating missing ones; None when
        the budget cannot hold them now"""
        missing = [k for k in dict.fromkeys(keys) if k not in self.cols]
        need = sum(self.nbytes(k[0]) for k in missing)
        # the query's resident columns are not candidates for eviction
        if need and not self._make_room(need, keep=set(keys)):
            return None
        for k in missing:
            buf = self.U.DeviceBuffer(self.nbytes(k[0]))
            self.U.gen_compact_into(self.desc(k), self.rows, buf, stream=self.stream)
            self.cols[k] = buf
            self.used += buf.nbytes
            self.generated += 1
        out = []
        for k in keys:
            self.cols.move_to_end(k)
            self.pins[k] += 1
            out.append(self.cols[k])
        return out

    def release(self, keys):
        for k in keys:
            self.pins[k] -= 1

    def acquire_query(self, q, i):
        return self.acquire(q["keys"])

    def release_query(self, q, i):
        self.release(q["keys"])


class FreshColumns:
    """Every query's own columns (the generator streams of the query,
    workloads.query_column_desc), generated on the device at admission and
    freed at completion.  A query is admitted while its columns fit in the
    free device memory less `reserve_bytes` (and, with `cap_bytes`, while
    the live columns stay under that cap)."""

    def __init__(self, U, seed, reserve_bytes=4 << 30, cap_bytes=None, stream=None):
        from . import workloads as W
        self.U = U
        self.W = W
        self.seed = seed
        self.reserve = int(reserve_bytes)
        self.cap = cap_bytes
        self.stream = stream
        self.live = {}
        self.used = 0
        self.peak = 0
        self.generated = 0
        self.refused = 0

    def nbytes(self, q):
        return len(q["cols"]) * (q["rows"] * self.U.compact_bytes(q["P"]) + 16)

    def acquire_query(self, q, i):
        need = self.nbytes(q)
        if self.cap is not None and self.used + need > self.cap:
            self.refused += 1
            return None
        free, _ = self.U.mem_info()
        if need > free - self.reserve:
            self.refused += 1
            return None
        bufs = []
        for k in range(len(q["cols"])):
            b = self.U.DeviceBuffer(q["rows"] * self.U.compact_bytes(q["P"]) + 16, ordered=True, stream=self.stream)
            self.U.gen_compact_into(self.W.query_column_desc(self.U, self.seed, q, k), q["rows"], b, stream=self.stream)
            bufs.append(b)
        self.live[i] = bufs
        self.used += need
        self.peak = max(self.peak, self.used)
        self.generated += len(bufs)
        return bufs

    def release_query(self, q, i):
        bufs = self.live.pop(i)
        self.used -= self.nbytes(q)
        del bufs


class Server:
    def __init__(self, U, store, batched=True, graph=True, max_batch=512, max_items=1 << 16, stream=None,
                 tuning=None):
        self.U = U
        self.store = store
        self.batched = batched
        self.graph = graph and batched
        self.max_batch = max_batch if batched else 1
        self.inflight_max = 2 if batched else 1
        self.stream = stream
        self.batches = [U.Batch(max_queries=self.max_batch, max_items=max_items) for _ in range(self.inflight_max)]
        for b in self.batches:
            for cls, t in (tuning or {}).items():
                b.tune(cls, t[0], t[1], lanes=t[2] if len(t) > 2 else None)
        self.programs = {}

    def program(self, q):
        key = (q["text"], tuple(q["cols"]))
        p = self.programs.get(key)
        if p is None:
            p = self.U.VmProgram(q["text"], q["cols"])
            self.programs[key] = p
        return p

    def run(self, queries, arrivals, images, counts):
        """Serve `queries` arriving at `arrivals` (seconds from the start,
        ascending); SUM images go to images[133 i], COUNTs to counts[8 i].
        Returns per-query latency (s), completion times and run statistics."""
        U = self.U
        n = len(queries)
        lat = np.zeros(n)
        fin = np.zeros(n)
        waiting = collections.deque()
        ready = []
        inflight = collections.deque()
        free_batches = list(self.batches)
        nxt = 0
        done = 0
        nbatches = 0
        max_wait = 0
        t0 = time.perf_counter()
        while done < n:
            now = time.perf_counter() - t0
            while nxt < n and arrivals[nxt] <= now:
                waiting.append(nxt)
                nxt += 1
            max_wait = max(max_wait, len(waiting) + len(ready))
            while waiting:
                i = waiting[0]
                q = queries[i]
                cols = self.store.acquire_query(q, i)
                if cols is None:
                    if not inflight and not ready:
                        raise RuntimeError("a query's columns exceed the device-memory budget on their own")
                    break
                waiting.popleft()
                ready.append((i, cols, self.program(q)))
            progressed = False
            if ready and free_batches:
                b = free_batches.pop()
                take = []
                for i, cols, prog in ready[:self.max_batch]:
                    if b.add(prog, cols, queries[i]["rows"], images, counts, image_offset=133 * i,
                             count_offset=8 * i) is None:
                        break
                    take.append((i, cols, prog))
                ready = ready[len(take):]
                b.launch(graph=self.graph, stream=self.stream)
                ev = U.Event()
                ev.record(self.stream)
                inflight.append((ev, take, b))
                nbatches += 1
                progressed = True
            if inflight and (inflight[0][0].done() if self.batched else (inflight[0][0].sync() or True)):
                ev, take, b = inflight.popleft()
                t = time.perf_counter() - t0
                for i, cols, prog in take:
                    fin[i] = t
                    lat[i] = t - arrivals[i]
                    self.store.release_query(queries[i], i)
                done += len(take)
                free_batches.append(b)
                progressed = True
            if not progressed:
                wait = arrivals[nxt] - (time.perf_counter() - t0) if nxt < n and not inflight and not ready else 0.0
                time.sleep(min(max(wait, 0.0), 0.002) if wait > 0 else 0.00005)
        makespan = fin.max() - (arrivals[0] if n else 0.0)
        rows = sum(q["rows"] for q in queries)
        return {
            "queries": n, "rows": rows, "makespan_s": makespan, "mrows_per_s": rows / makespan / 1e6 if makespan else 0.0,
            "p50_ms": 1e3 * float(np.percentile(lat, 50)), "p99_ms": 1e3 * float(np.percentile(lat, 99)),
            "mean_ms": 1e3 * float(lat.mean()), "batches": nbatches, "mean_batch": n / max(1, nbatches),
            "max_queue": max_wait, "latency_s": lat,
        }


def poisson_arrivals(n, rate, seed):
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, size=n))
