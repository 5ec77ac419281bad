"""Served perft: a request loop with continuous batching on one GPU.

Queries (position, depth) arrive over time.  Waiting queries are grouped by
depth; whenever a slot is idle it takes the depth group whose oldest query
has waited longest and admits queries of that depth while their estimated
level buffers fit the slot's HBM region.  Each admitted batch is one graph
launch on the slot's own stream, so batches on different slots overlap on
the device; all slots share one transposition table.  Latency is measured
from a query's arrival to the moment its batch is seen complete.
"""
import collections
import ctypes
import time

import numpy as np

from . import BatchStats, SLIDER_AUTO, _check, _ptr, lib, positions_array, u64

BYTES_PER_BOARD = 110      # board, key, count, perft, state, dup, offsets, transient move/parent/set
BRANCHING_BOUND = 45       # bound used for admission; an overflowing batch is finished in halves


def estimate_bytes(depth, n=1):
    """Level-buffer estimate of n roots at `depth` (largest materialized ply
    is depth-2; the leaf plies stay in LDS)."""
    if depth <= 2:
        return 4096 + n * BYTES_PER_BOARD
    return n * BYTES_PER_BOARD * (BRANCHING_BOUND ** (depth - 2)) + (1 << 20)


class Server:
    """pg_server: `slots` slots, each with its own stream, graphs and region."""

    def __init__(self, dev, slots=4, slot_bytes=None, max_roots=4096, tt=None, slider=SLIDER_AUTO, dedup=True):
        L = lib()
        if not hasattr(L, "_serve_sigs"):
            P = ctypes.POINTER
            for name, res, args in [
                ("pg_server_create", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, u64, ctypes.c_uint32,
                                                    ctypes.c_void_p, P(ctypes.c_void_p)]),
                ("pg_server_destroy", None, [ctypes.c_void_p]),
                ("pg_server_submit", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p,
                                                    ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, ctypes.c_int]),
                ("pg_server_poll", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, P(BatchStats)]),
                ("pg_device_memory", ctypes.c_int, [ctypes.c_void_p, P(u64), P(u64)]),
            ]:
                f = getattr(L, name)
                f.restype = res
                f.argtypes = args
            L._serve_sigs = True
        free, total = u64(), u64()
        _check(L.pg_device_memory(dev._h, ctypes.byref(free), ctypes.byref(total)), "pg_device_memory")
        self.free_hbm = int(free.value)
        if slot_bytes is None:
            slot_bytes = min(24 << 30, int(self.free_hbm * 0.6) // slots)
        self.slot_bytes = slot_bytes
        h = ctypes.c_void_p()
        _check(L.pg_server_create(dev._h, slots, slot_bytes, max_roots, tt._h if tt is not None else None,
                                  ctypes.byref(h)), "pg_server_create")
        self._h = h
        self.slots = slots
        self.max_roots = max_roots
        self.slider = slider
        self.dedup = dedup
        self._busy = [None] * slots

    def close(self):
        if self._h:
            lib().pg_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def submit(self, slot, positions, depth, tag=None):
        arr = positions_array(positions)
        _check(lib().pg_server_submit(self._h, slot, _ptr(arr), len(arr), depth, self.slider,
                                      1 if self.dedup else 0), "pg_server_submit")
        self._busy[slot] = (len(arr), tag)

    def poll(self, slot):
        """None while running, else (counts, stats, tag)."""
        n, tag = self._busy[slot]
        out = np.zeros(n, dtype=np.uint64)
        st = BatchStats()
        rc = lib().pg_server_poll(self._h, slot, _ptr(out), ctypes.byref(st))
        if rc == 0:
            return None
        _check(rc if rc < 0 else 0, "pg_server_poll")
        self._busy[slot] = None
        return out, st.as_dict(), tag

    def idle_slots(self):
        return [i for i, b in enumerate(self._busy) if b is None]

    def busy_slots(self):
        return [i for i, b in enumerate(self._busy) if b is not None]


def poisson_arrivals(n, rate, seed=0):
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, size=n))


def serve(server, positions, depths, arrivals, max_batch=1024, long_depth=6, long_slots=None):
    """Runs the request loop until every query is answered.

    positions: (n, 6) uint64; depths: (n,); arrivals: seconds from start.
    Batches of depth >= long_depth may hold at most long_slots slots at once
    (default: half of them), so short queries never queue behind a slot-full
    of long ones.  Among the depths that may start, the one whose oldest
    query arrived first goes next.
    Returns dict with per-query results, latencies (s) and counters."""
    if long_slots is None:
        long_slots = max(1, server.slots // 2)
    positions = positions_array(positions)
    n = len(positions)
    results = np.zeros(n, dtype=np.uint64)
    latency = np.zeros(n)
    done_t = np.zeros(n)
    waiting = collections.defaultdict(collections.deque)   # depth -> query ids
    nxt = 0
    finished = 0
    batches = 0
    overlap_samples = 0
    max_busy = 0
    batch_sizes = []
    running = {}   # slot -> depth of its batch
    t0 = time.perf_counter()
    while finished < n:
        now = time.perf_counter() - t0
        while nxt < n and arrivals[nxt] <= now:
            waiting[int(depths[nxt])].append(nxt)
            nxt += 1
        for s in server.busy_slots():
            r = server.poll(s)
            if r is None:
                continue
            counts, st, ids = r
            running.pop(s, None)
            tdone = time.perf_counter() - t0
            for q, c in zip(ids, counts):
                results[q] = c
                done_t[q] = tdone
                latency[q] = tdone - arrivals[q]
            finished += len(ids)
        for s in server.idle_slots():
            nlong = sum(1 for d in running.values() if d >= long_depth)
            live = [d for d in waiting if waiting[d] and (d < long_depth or nlong < long_slots)]
            if not live:
                break
            d = min(live, key=lambda k: arrivals[waiting[k][0]])
            budget = server.slot_bytes
            ids = []
            while waiting[d] and len(ids) < min(max_batch, server.max_roots):
                need = estimate_bytes(d, len(ids) + 1)
                if ids and need > budget:
                    break
                ids.append(waiting[d].popleft())
            server.submit(s, positions[ids], d, ids)
            running[s] = d
            batches += 1
            batch_sizes.append(len(ids))
        busy = len(server.busy_slots())
        max_busy = max(max_busy, busy)
        if busy >= 2:
            overlap_samples += 1
        if not busy and nxt < n and not any(waiting.values()):
            wait = arrivals[nxt] - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(min(wait, 0.001))
    elapsed = time.perf_counter() - t0
    return {"results": results, "latency": latency, "elapsed": elapsed, "batches": batches,
            "batch_sizes": batch_sizes, "max_busy_slots": max_busy, "overlap_samples": overlap_samples,
            "completion": done_t}


def latency_summary(lat):
    lat = np.asarray(lat) * 1e3
    return {"p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(lat.mean()), "max_ms": float(lat.max())}
