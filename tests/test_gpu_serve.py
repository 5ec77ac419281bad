"""Served perft with continuous batching: every served result equals the
reference's value and the result of the same query alone; queries overlap
on several slots and share batches."""
import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu


def c5():
    g = golden("c5.json")
    return np.array(g["words"], np.uint64), np.array(g["depths"], np.uint32), g["perft"]


def test_serve_mixed_stream_exact(dev, pg):
    from perft_gpu_amd.serve import Server, serve, poisson_arrivals
    pos, depths, want = c5()
    n = 256
    with pg.TranspositionTable(dev, 2 << 30) as tt, Server(dev, slots=4, slot_bytes=4 << 30, tt=tt) as srv:
        arr = poisson_arrivals(n, 2000.0, seed=7)
        print("serve", n, "queries", flush=True)
        out = serve(srv, pos[:n], depths[:n], arr)
    assert [int(x) for x in out["results"]] == want[:n]
    assert max(out["batch_sizes"]) > 1          # queries shared a batch
    assert out["max_busy_slots"] >= 2           # batches overlapped on the device
    # the same query alone, on the plain path
    for i in range(0, n, 37):
        assert dev.perft(pg.pos_from_words(pos[i]), int(depths[i]), 0) == want[i]


def test_serve_burst_all_at_once(dev, pg):
    from perft_gpu_amd.serve import Server, serve
    pos, depths, want = c5()
    n = 128
    with Server(dev, slots=2, slot_bytes=4 << 30, tt=None) as srv:
        out = serve(srv, pos[:n], depths[:n], np.zeros(n))
    assert [int(x) for x in out["results"]] == want[:n]


def test_served_batch_overflow_finishes_in_halves(dev, pg):
    from perft_gpu_amd.serve import Server
    g = golden("c2.json")["positions"]
    arr = np.array([x["words"] for x in g], np.uint64)
    with Server(dev, slots=1, slot_bytes=16 << 20, tt=None) as srv:
        srv.submit(0, arr, 5, list(range(len(arr))))
        r = None
        while r is None:
            r = srv.poll(0)
        counts, st, _ = r
    assert [int(x) for x in counts] == [x["perft"][5] for x in g]
    assert st["splits"] > 0
