"""The serving loop's host logic, against a stand-in server that answers each
query with a value derived from its position and depth after a delay that
grows with the depth: every query gets its own answer, batches hold one
depth, long batches never hold more than their share of the slots, and a
stream of long queries alone still drains."""
import time

import numpy as np

from perft_gpu_amd.serve import serve, poisson_arrivals, latency_summary


class FakeServer:
    def __init__(self, slots=4, slot_bytes=16 << 30, max_roots=64):
        self.slots, self.slot_bytes, self.max_roots = slots, slot_bytes, max_roots
        self._busy = [None] * slots
        self.log = []          # (time, event, slot, depth, n)
        self.peak_long = 0

    def submit(self, slot, positions, depth, tag=None):
        assert self._busy[slot] is None
        arr = np.asarray(positions, np.uint64).reshape(-1, 6)
        done = time.perf_counter() + 0.0002 * (1 << max(0, depth - 3))
        self._busy[slot] = (done, depth, arr[:, 1] * 10 + depth, tag)
        long = sum(1 for b in self._busy if b is not None and b[1] >= 6)
        self.peak_long = max(self.peak_long, long)
        self.log.append(("submit", slot, depth, len(arr)))

    def poll(self, slot):
        done, depth, counts, tag = self._busy[slot]
        if time.perf_counter() < done:
            return None
        self._busy[slot] = None
        return counts.astype(np.uint64), {}, tag

    def idle_slots(self):
        return [i for i, b in enumerate(self._busy) if b is None]

    def busy_slots(self):
        return [i for i, b in enumerate(self._busy) if b is not None]


def queries(n, seed):
    rng = np.random.default_rng(seed)
    pos = np.zeros((n, 6), np.uint64)
    pos[:, 1] = np.arange(n, dtype=np.uint64) + 1000
    depths = rng.choice([3, 4, 5, 6, 7], size=n, p=[0.4, 0.3, 0.2, 0.08, 0.02]).astype(np.uint32)
    return pos, depths


def test_every_query_gets_its_answer_and_long_batches_keep_to_their_slots():
    pos, depths = queries(400, 3)
    srv = FakeServer(slots=4)
    out = serve(srv, pos, depths, poisson_arrivals(len(pos), 4000.0, seed=1))
    want = pos[:, 1] * 10 + depths
    assert np.array_equal(out["results"], want.astype(np.uint64))
    assert srv.peak_long <= 2
    assert sum(n for _, _, _, n in srv.log) == len(pos)
    assert max(out["batch_sizes"]) > 1
    s = latency_summary(out["latency"])
    assert 0 <= s["p50_ms"] <= s["p99_ms"] <= s["max_ms"]


def test_only_long_queries_still_drain():
    pos, _ = queries(40, 4)
    depths = np.full(len(pos), 7, np.uint32)
    srv = FakeServer(slots=4)
    out = serve(srv, pos, depths, np.zeros(len(pos)), long_slots=1)
    assert np.array_equal(out["results"], (pos[:, 1] * 10 + 7).astype(np.uint64))
    assert srv.peak_long == 1


def test_batches_respect_max_roots():
    pos, depths = queries(300, 5)
    srv = FakeServer(slots=2, max_roots=16)
    out = serve(srv, pos, depths, np.zeros(len(pos)))
    assert max(out["batch_sizes"]) <= 16
    assert np.array_equal(out["results"], (pos[:, 1] * 10 + depths).astype(np.uint64))
