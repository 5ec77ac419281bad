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


def test_server_rejects_misuse(dev, pg):
    """Host-side argument checks of pg_server_submit and of the perft entry
    points: nothing invalid reaches a kernel, and the slot stays usable."""
    from perft_gpu_amd.serve import Server
    g = golden("c2.json")["positions"]
    arr = np.array([x["words"] for x in g], np.uint64)
    bad = np.zeros((1, 6), np.uint64)          # no kings
    with Server(dev, slots=2, slot_bytes=64 << 20, max_roots=8, tt=None) as srv:
        with pytest.raises(pg.PerftError):
            srv.submit(2, arr[:1], 3)           # no such slot
        with pytest.raises(pg.PerftError):
            srv.submit(0, arr[:9], 3)           # more roots than the slot holds
        with pytest.raises(pg.PerftError):
            srv.submit(0, bad, 3)
        srv.submit(0, arr[:4], 3, list(range(4)))
        with pytest.raises(pg.PerftError):
            srv.submit(0, arr[:4], 3)           # slot still busy
        r = None
        while r is None:
            r = srv.poll(0)
        assert [int(x) for x in r[0]] == [x["perft"][3] for x in g[:4]]
    with pytest.raises(pg.PerftError):
        dev.perft_batch(bad, 3, 0, False)
    with pytest.raises(pg.PerftError):
        dev.perft(pg.pos_from_words(bad[0]), 3, 0)
    with pg.TranspositionTable(dev, 1 << 30) as tt:
        with pytest.raises(pg.PerftError):
            tt.perft(pg.pos_from_words(bad[0]), 3)
