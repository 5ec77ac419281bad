"""Transposition-table perft: exact against the no-TT path, the reference's
values and published CPW counts; garbage in the table is never taken for a
hit; levels split when the arena is small."""
import numpy as np
import pytest

from conftest import golden
from perft_gpu_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tt(dev, pg):
    t = pg.TranspositionTable(dev, 2 << 30)
    yield t
    t.close()


def test_tt_layout(tt):
    nbytes, bits = tt.info()
    assert nbytes <= 2 << 30
    assert bits[2] == -1 and bits[3] == bits[4] > 0 and all(b > 0 for b in bits[5:])


@pytest.mark.parametrize("v", [0, 1, 2, 3])
def test_tt_start_position(dev, tt, v):
    tt.clear()
    for d in range(1, 9):
        print("tt start d%d variant %d" % (d, v), flush=True)
        assert tt.perft(W.CPW["start"], d, v) == W.CPW_COUNTS["start"][d]
    c = tt.counters()
    assert c[2:, 1].sum() > 0          # hits
    assert c[2:, 3].sum() > 0          # in-level duplicates


def test_tt_without_dedup(dev, tt):
    tt.clear()
    dev.set_dedup(False)
    try:
        assert tt.perft(W.CPW["start"], 7, 0) == W.CPW_COUNTS["start"][7]
        assert tt.counters()[:, 3].sum() == 0
    finally:
        dev.set_dedup(True)


def test_tt_cpw_and_random(dev, tt, pg):
    tt.clear()
    for name, item in golden("cpw.json").items():
        for d, want in enumerate(item["perft"]):
            assert tt.perft(item["fen"], d, 0) == want, (name, d)
    for item in golden("c2.json")["positions"]:
        print("tt c2", item["fen"], flush=True)
        assert tt.perft(pg.pos_from_words(item["words"]), 6, 0) == item["perft"][6]


def test_tt_reused_across_calls_and_poison_rejected(dev, tt, pg):
    tt.poison(12345)
    assert tt.perft(W.CPW["kiwipete"], 5, 0) == W.CPW_COUNTS["kiwipete"][5]
    tt.poison(999)
    for item in golden("random.json")[:64]:
        assert tt.perft(pg.pos_from_words(item["words"]), 4, 0) == item["perft"][4]
    # warm table: the second call is served from it
    tt.clear()
    a = tt.perft(W.CPW["pos3"], 7, 0)
    hits_before = tt.counters()[:, 1].sum()
    b = tt.perft(W.CPW["pos3"], 7, 0)
    assert a == b == W.CPW_COUNTS["pos3"][7]
    assert tt.counters()[:, 1].sum() >= hits_before


def test_tt_small_arena_splits(pg):
    with pg.Device(0, arena_bytes=96 << 20) as small:
        with pg.TranspositionTable(small, 256 << 20) as t:
            print("tt split arena start d8", flush=True)
            assert t.perft(W.CPW["start"], 8, 0) == W.CPW_COUNTS["start"][8]
            assert small.last_stats()["splits"] > 0


@pytest.mark.slow
def test_tt_depth9_equals_no_tt_path(dev, pg):
    # C4: TT perft(9) bit-exact to the plain breadth-first path and to CPW
    with pg.TranspositionTable(dev, 16 << 30) as big:
        print("tt d9", flush=True)
        a = big.perft(W.CPW["start"], 9, 0)
    print("no-tt d9", flush=True)
    b = dev.perft(W.CPW["start"], 9, 0)
    assert a == b == W.CPW_COUNTS["start"][9]


@pytest.mark.slow
@pytest.mark.parametrize("gb", [16, 64])
def test_tt_depth10_with_16_and_64_gb(pg, gb):
    """TT perft(10) of the start position with 16 GB and 64 GB tables equals
    the published 69,352,859,712,417 (perft(9) on the same table first).  The
    16 GB table runs on a 48 GiB arena (14 level splits), the 64 GB one on
    the arena the table leaves free (pg.table_search_arena: one or two)."""
    arena = 48 << 30 if gb == 16 else (pg.table_search_arena(gb << 30) or pg.DEFAULT_ARENA)
    with pg.Device(0, arena_bytes=arena) as d, pg.TranspositionTable(d, gb << 30) as big:
        for depth in (9, 10):
            print("tt %d GB d%d arena %.0f GiB" % (gb, depth, arena / (1 << 30)), flush=True)
            assert big.perft(W.CPW["start"], depth, 0) == W.CPW_COUNTS["start"][depth]


def test_table_search_arena(pg):
    a = pg.table_search_arena(1 << 30)
    assert a == 0 or pg.DEFAULT_ARENA < a <= pg.TABLE_ARENA_MAX
    assert pg.table_search_arena(1 << 50) == 0  # a table larger than the device: the library default


def test_tt_checkpoint_resume(dev, pg, tmp_path):
    """pg_tt_save / pg_tt_load: a restored table answers the root from the
    checkpoint; a file of another layout or with damaged data is refused and
    leaves a cleared, still exact table."""
    path = str(tmp_path / "tt.ckpt")
    with pg.TranspositionTable(dev, 1 << 30) as a:
        assert a.perft(W.CPW["start"], 6) == W.CPW_COUNTS["start"][6]
        a.save(path)
    with pg.TranspositionTable(dev, 1 << 30) as b:
        b.load(path)
        assert b.perft(W.CPW["start"], 6) == W.CPW_COUNTS["start"][6]
        # answered by the root entry: no level below was searched
        assert int(b.counters().sum()) == 0 and dev.last_stats()["leaf_launches"] == 0
    with pg.TranspositionTable(dev, 2 << 30) as other:
        with pytest.raises(pg.PerftError):
            other.load(path)
    with open(path, "r+b") as f:                      # damage one entry word
        f.seek(128 + 4096)
        w = f.read(8)
        f.seek(128 + 4096)
        f.write(bytes(x ^ 0xFF for x in w))
    with pg.TranspositionTable(dev, 1 << 30) as d:
        with pytest.raises(pg.PerftError):
            d.load(path)
        assert d.perft(W.CPW["start"], 6) == W.CPW_COUNTS["start"][6]
