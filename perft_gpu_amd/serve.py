"""Served perft: a thin Python face of the library's serving loop (pg_serve_*).

Queries (position, depth) arrive over time; the loop in the library groups
the waiting ones by depth, admits a group into an idle slot as one batched
graph launch when its estimated level buffers fit the slot's HBM region,
runs batches of different slots concurrently on the device (sharing one
optional transposition table) and returns each query's result and latency.
The policy is described in include/perft_gpu.h (pg_serve_config); `serve`
below only feeds a stream of queries at their arrival times and collects the
answers.  Server wraps the slot-level pg_server_* calls the loop is built on.
"""
import ctypes
import time

import numpy as np

from . import BatchStats, SLIDER_KOGGE_STONE, _check, _ptr, lib, positions_array, u64, Pos48

P = ctypes.POINTER


class ServeConfig(ctypes.Structure):
    _fields_ = [("slots", ctypes.c_uint32), ("bytes_per_slot", u64), ("max_batch", ctypes.c_uint32),
                ("long_depth", ctypes.c_uint32), ("long_slots", ctypes.c_uint32), ("long_wg_per_cu", ctypes.c_uint32),
                ("slider", ctypes.c_int), ("dedup", ctypes.c_int), ("light_wg_per_cu", ctypes.c_uint32),
                ("light_queue", ctypes.c_uint32)]


class ServeResult(ctypes.Structure):
    _fields_ = [("id", u64), ("tag", u64), ("perft", u64), ("depth", ctypes.c_uint32), ("status", ctypes.c_int),
                ("latency_ms", ctypes.c_double), ("queued_ms", ctypes.c_double)]


class ServeStats(ctypes.Structure):
    _fields_ = [("submitted", u64), ("completed", u64), ("batches", u64), ("launches", u64), ("splits", u64),
                ("leaves", u64), ("waiting", ctypes.c_uint32), ("busy_slots", ctypes.c_uint32),
                ("max_busy_slots", ctypes.c_uint32), ("max_long_slots", ctypes.c_uint32),
                ("p50_ms", ctypes.c_double), ("p99_ms", ctypes.c_double), ("mean_ms", ctypes.c_double),
                ("max_ms", ctypes.c_double), ("elapsed_ms", ctypes.c_double),
                ("est_bytes_per_query", ctypes.c_double * 16), ("submit_ms_max", ctypes.c_double),
                ("submit_ms_total", ctypes.c_double), ("graphs", u64), ("device_ms", ctypes.c_double)]

    def as_dict(self):
        d = {n: getattr(self, n) for n, _ in self._fields_ if n != "est_bytes_per_query"}
        d["est_bytes_per_query"] = list(self.est_bytes_per_query)
        return d


def _sigs():
    L = lib()
    if getattr(L, "_serve_sigs", False):
        return L
    for name, res, args in [
        ("pg_server_create", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, u64, ctypes.c_uint32, ctypes.c_void_p,
                                            P(ctypes.c_void_p)]),
        ("pg_server_destroy", None, [ctypes.c_void_p]),
        ("pg_server_submit", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint32,
                                            ctypes.c_uint32, ctypes.c_int, ctypes.c_int]),
        ("pg_server_submit_capped", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint32,
                                                   ctypes.c_uint32, ctypes.c_int, ctypes.c_int, ctypes.c_uint32]),
        ("pg_server_poll", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, P(BatchStats)]),
        ("pg_device_memory", ctypes.c_int, [ctypes.c_void_p, P(u64), P(u64)]),
        ("pg_serve_create", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, P(ServeConfig), P(ctypes.c_void_p)]),
        ("pg_serve_destroy", None, [ctypes.c_void_p]),
        ("pg_serve_submit", ctypes.c_int, [ctypes.c_void_p, P(Pos48), ctypes.c_uint32, u64, P(u64)]),
        ("pg_serve_poll", ctypes.c_int, [ctypes.c_void_p, P(ServeResult), ctypes.c_uint32, P(ctypes.c_uint32)]),
        ("pg_serve_get_stats", ctypes.c_int, [ctypes.c_void_p, P(ServeStats)]),
        ("pg_serve_policy", ctypes.c_int, [P(ctypes.c_uint32), P(ctypes.c_double), ctypes.c_uint32, ctypes.c_uint32,
                                           P(ServeConfig), u64, P(ctypes.c_double), P(ctypes.c_uint32),
                                           P(ctypes.c_uint32)]),
        ("pg_serve_wg_cap", ctypes.c_uint32, [P(ServeConfig), ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int]),
    ]:
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    L._serve_sigs = True
    return L


def policy(waiting, oldest, busy_long, idle_slots, region_bytes, est_bytes, **cfg):
    """pg_serve_policy: (depth, count) of the next batch, or None."""
    L = _sigs()
    c = ServeConfig(**cfg)
    w = (ctypes.c_uint32 * 16)(*waiting)
    o = (ctypes.c_double * 16)(*oldest)
    e = (ctypes.c_double * 16)(*est_bytes)
    d, k = ctypes.c_uint32(), ctypes.c_uint32()
    rc = L.pg_serve_policy(w, o, busy_long, idle_slots, ctypes.byref(c), region_bytes, e, ctypes.byref(d),
                           ctypes.byref(k))
    _check(rc if rc < 0 else 0, "pg_serve_policy")
    return (int(d.value), int(k.value)) if rc == 1 else None


def wg_cap(depth, queued, shorts_about, **cfg):
    """pg_serve_wg_cap: resident work-groups per CU for the next batch (0 = uncapped)."""
    c = ServeConfig(**cfg)
    return int(_sigs().pg_serve_wg_cap(ctypes.byref(c), depth, queued, 1 if shorts_about else 0))


class ServeLoop:
    """pg_serve: the library's serving loop on one device context."""
    _close_rank = 0

    def __init__(self, dev, tt=None, slots=0, slot_bytes=0, max_batch=1024, long_depth=6, long_slots=0,
                 long_wg_per_cu=2, slider=SLIDER_KOGGE_STONE, dedup=True, light_wg_per_cu=0, light_queue=0):
        L = _sigs()
        self.cfg = ServeConfig(slots, slot_bytes, max_batch, long_depth, long_slots, long_wg_per_cu, slider,
                               1 if dedup else 0, light_wg_per_cu, light_queue)
        h = ctypes.c_void_p()
        _check(L.pg_serve_create(dev._h, tt._h if tt is not None else None, ctypes.byref(self.cfg), ctypes.byref(h)),
               "pg_serve_create")
        self._h = h
        dev._adopt(self)
        self._buf = (ServeResult * 4096)()

    def close(self):
        if self._h:
            lib().pg_serve_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def submit(self, pos, depth, tag=0):
        if not isinstance(pos, Pos48):
            pos = Pos48(*[int(x) for x in pos])
        qid = u64()
        _check(lib().pg_serve_submit(self._h, ctypes.byref(pos), int(depth), tag, ctypes.byref(qid)),
               "pg_serve_submit")
        return int(qid.value)

    def poll(self):
        """Runs the loop once; returns the queries finished since the last poll
        as (id, tag, perft, depth, status, latency_ms, queued_ms) tuples."""
        out = []
        while True:
            n = ctypes.c_uint32()
            _check(lib().pg_serve_poll(self._h, self._buf, len(self._buf), ctypes.byref(n)), "pg_serve_poll")
            out.extend((r.id, r.tag, r.perft, r.depth, r.status, r.latency_ms, r.queued_ms)
                       for r in self._buf[:n.value])
            if n.value < len(self._buf):
                return out

    def stats(self):
        st = ServeStats()
        _check(lib().pg_serve_get_stats(self._h, ctypes.byref(st)), "pg_serve_get_stats")
        return st.as_dict()


def poisson_arrivals(n, rate, seed=0):
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, size=n))


def serve(loop, positions, depths, arrivals):
    """Feeds query i to the loop at arrivals[i] seconds from the start and
    polls until every query is answered.  Latency counts from the scheduled
    arrival (the loop's own submit-to-done latency plus any delay before the
    submit); `queued` is the part of it before the query's batch was
    launched.  Returns results, latencies (s), elapsed time, this stream's
    batches, their summed device time and its ratio to the elapsed time (the
    mean number of batches running at once), and the loop's stats
    (max_busy_slots and max_long_slots are maxima over the loop's lifetime)."""
    positions = positions_array(positions)
    n = len(positions)
    results = np.zeros(n, dtype=np.uint64)
    latency = np.zeros(n)
    queued = np.zeros(n)
    late = np.zeros(n)
    pos = [Pos48(*[int(x) for x in w]) for w in positions]
    st0 = loop.stats()
    finished, nxt = 0, 0
    t0 = time.perf_counter()
    while finished < n:
        now = time.perf_counter() - t0
        while nxt < n and arrivals[nxt] <= now:
            late[nxt] = now - arrivals[nxt]
            loop.submit(pos[nxt], int(depths[nxt]), nxt)
            nxt += 1
        for qid, tag, perft, depth, status, lat_ms, q_ms in loop.poll():
            _check(status, "served query %d" % tag)
            results[tag] = perft
            latency[tag] = late[tag] + lat_ms * 1e-3
            queued[tag] = late[tag] + q_ms * 1e-3
            finished += 1
        if finished == nxt and nxt < n:
            wait = arrivals[nxt] - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(min(wait, 0.0005))
    elapsed = time.perf_counter() - t0
    st = loop.stats()
    # the loop's counters run over its lifetime: this stream's share
    batches = st["batches"] - st0["batches"]
    device_ms = st["device_ms"] - st0["device_ms"]
    return {"results": results, "latency": latency, "queued": queued, "elapsed": elapsed, "batches": batches,
            "device_ms": device_ms, "batches_in_flight": device_ms / max(elapsed * 1e3, 1e-9),
            "mean_batch": (st["completed"] - st0["completed"]) / max(1, batches), "max_busy_slots": st["max_busy_slots"],
            "max_long_slots": st["max_long_slots"], "splits": st["splits"], "stats": st}


def latency_summary(lat):
    lat = np.asarray(lat) * 1e3
    return {"p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(lat.mean()), "max_ms": float(lat.max())}


class Server:
    """pg_server: `slots` slots, each with its own stream, graphs and HBM region
    (the slot-level calls the serving loop is built on)."""
    _close_rank = 1

    def __init__(self, dev, slots=4, slot_bytes=None, max_roots=4096, tt=None, slider=SLIDER_KOGGE_STONE, dedup=True):
        L = _sigs()
        free, total = u64(), u64()
        _check(L.pg_device_memory(dev._h, ctypes.byref(free), ctypes.byref(total)), "pg_device_memory")
        if slot_bytes is None:
            slot_bytes = min(24 << 30, int(free.value * 0.6) // slots)
        self.slot_bytes = slot_bytes
        h = ctypes.c_void_p()
        _check(L.pg_server_create(dev._h, slots, slot_bytes, max_roots, tt._h if tt is not None else None,
                                  ctypes.byref(h)), "pg_server_create")
        self._h = h
        dev._adopt(self)
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

    def submit(self, slot, positions, depth, tag=None, wg_per_cu=0):
        arr = positions_array(positions)
        _check(lib().pg_server_submit_capped(self._h, slot, _ptr(arr), len(arr), depth, self.slider,
                                             1 if self.dedup else 0, wg_per_cu), "pg_server_submit")
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
