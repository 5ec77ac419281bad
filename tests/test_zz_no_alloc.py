"""pg_perft_noalloc (no counterpart in the reference): the plain perft with
its tree arena and per-call words in the library's 64 MiB static device
scratch.  Called through the C ABI, it must return what pg_perft returns,
on two calls in a row (the second with a smaller input, over the first
call's leftovers in the shared scratch), for every slider variant, and
reject bad arguments with an error instead of launching anything."""
import ctypes

import pytest

from perft_gpu_amd import workloads as W

pytestmark = pytest.mark.gpu


def noalloc(pg, dev, fen, depth, slider=0):
    pos = pg.read_fen_string(fen)
    out = ctypes.c_uint64(0)
    rc = pg.lib().pg_perft_noalloc(dev._h, ctypes.byref(pos), depth, slider, ctypes.byref(out))
    return rc, int(out.value)


def test_noalloc_equals_pg_perft_two_calls(pg):
    with pg.Device(0, arena_bytes=1 << 30) as dev:
        # first call: start position depth 7; its ply-5 level (233 MB of boards)
        # only fits the 64 MiB scratch in split halves
        print("noalloc start d7", flush=True)
        rc, a = noalloc(pg, dev, W.CPW["start"], 7)
        assert rc == pg.PG_OK
        assert a == dev.perft(W.CPW["start"], 7, 0) == W.CPW_COUNTS["start"][7]
        assert dev.last_stats()["splits"] == 0          # pg_perft had its own 1 GiB arena
        # second call, smaller input, over the first call's scratch contents
        print("noalloc kiwipete d4", flush=True)
        rc, b = noalloc(pg, dev, W.CPW["kiwipete"], 4)
        assert rc == pg.PG_OK
        assert b == dev.perft(W.CPW["kiwipete"], 4, 0) == W.CPW_COUNTS["kiwipete"][4]


@pytest.mark.parametrize("v", [0, 1, 2, 3])
def test_noalloc_every_slider_variant(pg, v):
    with pg.Device(0, arena_bytes=1 << 30) as dev:
        for name, depth in (("pos4", 5), ("pos3", 4)):
            print("noalloc %s d%d variant %d" % (name, depth, v), flush=True)
            rc, got = noalloc(pg, dev, W.CPW[name], depth, v)
            assert rc == pg.PG_OK
            assert got == dev.perft(W.CPW[name], depth, v) == W.CPW_COUNTS[name][depth]


def test_noalloc_rejects_bad_arguments(pg):
    with pg.Device(0, arena_bytes=1 << 30) as dev:
        rc, _ = noalloc(pg, dev, W.CPW["start"], 16)
        assert rc == pg.PG_EINVAL and "15" in pg.last_error()
        bad = pg.Pos48(0, 0, 0, 0, 0, 0)          # no kings
        out = ctypes.c_uint64(0)
        assert pg.lib().pg_perft_noalloc(dev._h, ctypes.byref(bad), 3, 0, ctypes.byref(out)) == pg.PG_EINVAL
        assert pg.lib().pg_perft_noalloc(dev._h, ctypes.byref(bad), 3, 7, ctypes.byref(out)) == pg.PG_EINVAL
        # the context still works after the rejected calls
        rc, n = noalloc(pg, dev, W.CPW["start"], 5)
        assert rc == pg.PG_OK and n == W.CPW_COUNTS["start"][5]
