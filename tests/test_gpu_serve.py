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
    """256 queries arriving at 2000/s through the library's serving loop with a
    shared table: every answer equals the reference's value and the same query
    alone; queries share batches, batches overlap on several slots, long
    batches keep to their share of the slots."""
    from perft_gpu_amd.serve import ServeLoop, serve, poisson_arrivals
    pos, depths, want = c5()
    n = 256
    with pg.TranspositionTable(dev, 2 << 30) as tt, ServeLoop(dev, tt, slots=4, slot_bytes=4 << 30) as loop:
        arr = poisson_arrivals(n, 2000.0, seed=7)
        print("serve", n, "queries", flush=True)
        out = serve(loop, pos[:n], depths[:n], arr)
        st = loop.stats()
    assert [int(x) for x in out["results"]] == want[:n]
    assert st["completed"] == n and st["batches"] < n   # queries shared a batch
    assert st["max_busy_slots"] >= 2                     # batches overlapped on the device
    assert st["max_long_slots"] <= 2
    assert 0 < st["p50_ms"] <= st["p99_ms"] <= st["max_ms"]
    # device time of the batches: positive, and at most one batch per slot at a time
    assert 0 < out["device_ms"] <= 4 * out["elapsed"] * 1e3 * 1.05
    assert out["batches_in_flight"] > 0
    # the same query alone, on the plain path
    for i in range(0, n, 37):
        assert dev.perft(pg.pos_from_words(pos[i]), int(depths[i]), 0) == want[i]


def test_serve_burst_all_at_once_without_table(dev, pg):
    from perft_gpu_amd.serve import ServeLoop, serve
    pos, depths, want = c5()
    n = 128
    with ServeLoop(dev, None, slots=2, slot_bytes=4 << 30) as loop:
        out = serve(loop, pos[:n], depths[:n], np.zeros(n))
    assert [int(x) for x in out["results"]] == want[:n]
    assert out["mean_batch"] > 1


def test_serve_learns_region_bytes_and_splits_small_regions(dev, pg):
    """With 64 MiB regions the loop sizes depth-5 batches by its region-bytes
    estimate, learns the estimate from the finished batches, and every answer
    stays exact."""
    from perft_gpu_amd.serve import ServeLoop, serve
    g = golden("c2.json")["positions"]
    arr = np.array([x["words"] for x in g] * 4, np.uint64)
    want = [x["perft"][5] for x in g] * 4
    with ServeLoop(dev, None, slots=1, slot_bytes=64 << 20, max_batch=64) as loop:
        out = serve(loop, arr, np.full(len(arr), 5, np.uint32), np.zeros(len(arr)))
        st = loop.stats()
    assert [int(x) for x in out["results"]] == want
    assert st["batches"] >= 2 and st["est_bytes_per_query"][5] > 0


def test_serve_rejects_bad_queries(dev, pg):
    from perft_gpu_amd.serve import ServeLoop
    g = golden("c2.json")["positions"]
    with ServeLoop(dev, None, slots=1, slot_bytes=64 << 20) as loop:
        with pytest.raises(pg.PerftError):
            loop.submit(g[0]["words"], 0)                 # depth 0
        with pytest.raises(pg.PerftError):
            loop.submit(g[0]["words"], 16)                # beyond the deepest table
        with pytest.raises(pg.PerftError):
            loop.submit([0] * 6, 3, 77)                   # no kings: refused before it can join a batch
        good = loop.submit(g[0]["words"], 3, 78)
        got = []
        while not got:
            got = loop.poll()
        assert got == [(good, 78, g[0]["perft"][3], 3, pg.PG_OK, got[0][5], got[0][6])]
        assert 0 <= got[0][6] <= got[0][5]
        assert loop.stats()["submitted"] == 1


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
    assert st["splits"] > 0 and st["graph_ms"] > 0


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


def test_c_host_serves_a_stream(pg):
    """perft_gpu_amd/perft_serve (plain C over the C ABI) serves 300 Poisson
    queries with a shared table and checks each answer against pg_perft."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(pg.LIB_PATH), "perft_serve")
    assert os.path.exists(exe), "perft_serve is built with the library (make -C perft_gpu_amd)"
    out = subprocess.run([exe, "300", "2000", "1", "7"], capture_output=True, text=True, timeout=180)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "served 300 queries" in out.stdout and "mismatches 0" in out.stdout


def test_handle_lifetimes(pg):
    """A table shared by a slot server cannot be destroyed until the server
    is; closing a context releases its serving loops, servers and tables
    (through the C ABI: pg_shutdown releases what is still alive)."""
    import ctypes
    from perft_gpu_amd.serve import Server, ServeLoop
    with pg.Device(0, arena_bytes=1 << 30) as d:
        tt = pg.TranspositionTable(d, 1 << 30)
        srv = Server(d, slots=1, slot_bytes=64 << 20, tt=tt)
        with pytest.raises(pg.PerftError) as e:
            tt.close()
        assert e.value.code == pg.PG_EINVAL and tt._h
        srv.close()
        tt.close()
        # left open on purpose: the context closes them in order
        tt2 = pg.TranspositionTable(d, 1 << 30)
        ServeLoop(d, tt2, slots=1, slot_bytes=64 << 20)
        Server(d, slots=1, slot_bytes=64 << 20, tt=tt2)
    assert tt2._h is None
    # C ABI: a context shut down with a table and a server still alive
    L = pg.lib()
    ctx, t, s = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    assert L.pg_init(0, None, ctypes.byref(ctx)) == 0
    assert L.pg_tt_create(ctx, 1 << 30, ctypes.byref(t)) == 0
    assert L.pg_server_create(ctx, 1, 64 << 20, 8, t, ctypes.byref(s)) == 0
    assert L.pg_tt_destroy(t) == pg.PG_EINVAL
    L.pg_shutdown(ctx)
