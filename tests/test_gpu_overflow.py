"""64-bit totals never wrap silently: with the root total preloaded near
2^64 (pg_test_count_bias), a sum that passes 2^64 makes every perft path
return PG_EOVERFLOW, and one that ends exactly at 2^64 - 1 is returned
exactly.  The reference keeps deep totals in InfInt (launcher.h:1181)."""
import numpy as np
import pytest

from conftest import golden
from perft_gpu_amd import workloads as W

pytestmark = pytest.mark.gpu
TOP = (1 << 64) - 1


@pytest.fixture
def biased(dev):
    yield dev
    dev.set_count_bias(0)


@pytest.mark.parametrize("depth", [1, 2, 3, 5])
def test_plain_path_overflow(biased, pg, depth):
    want = W.CPW_COUNTS["start"][depth]
    print("plain overflow d%d" % depth, flush=True)
    biased.set_count_bias(TOP - want)          # ends exactly at 2^64 - 1
    assert biased.perft(W.CPW["start"], depth, 0) == TOP
    biased.set_count_bias(TOP - want + 1)      # one past: wraps
    with pytest.raises(pg.PerftError) as e:
        biased.perft(W.CPW["start"], depth, 0)
    assert e.value.code == pg.PG_EOVERFLOW
    biased.set_count_bias(0)
    assert biased.perft(W.CPW["start"], depth, 0) == want


def test_materialized_leaf_overflow(pg):
    with pg.Device(0, arena_bytes=1 << 30, leaf_mode=1) as d:
        want = W.CPW_COUNTS["kiwipete"][4]
        d.set_count_bias(TOP - want + 1)
        with pytest.raises(pg.PerftError) as e:
            d.perft(W.CPW["kiwipete"], 4, 0)
        assert e.value.code == pg.PG_EOVERFLOW
        d.set_count_bias(TOP - want)
        assert d.perft(W.CPW["kiwipete"], 4, 0) == TOP
        assert d.perft(W.CPW["kiwipete"], 4, 3) == TOP   # byte-lookup variant: materialized leaf


@pytest.mark.parametrize("depth", [1, 2, 3, 6])
def test_tt_path_overflow(biased, pg, depth):
    want = W.CPW_COUNTS["start"][depth]
    with pg.TranspositionTable(biased, 1 << 30) as tt:
        print("tt overflow d%d" % depth, flush=True)
        biased.set_count_bias(TOP - want + 1)
        with pytest.raises(pg.PerftError) as e:
            tt.perft(W.CPW["start"], depth, 0)
        assert e.value.code == pg.PG_EOVERFLOW
        biased.set_count_bias(TOP - want)
        assert tt.perft(W.CPW["start"], depth, 0) == TOP
        biased.set_count_bias(0)
        # biased totals were not stored: the table still answers exactly
        assert tt.perft(W.CPW["start"], depth, 0) == want


@pytest.mark.parametrize("depth", [1, 2, 3, 4])
def test_batch_path_overflow(biased, pg, depth):
    items = golden("random.json")[:16]
    arr = np.array([it["words"] for it in items], np.uint64)
    wants = [it["perft"][depth] for it in items]
    big = max(wants)
    print("batch overflow d%d" % depth, flush=True)
    biased.set_count_bias(TOP - big + 1)      # the largest root wraps
    with pytest.raises(pg.PerftError) as e:
        biased.perft_batch(arr, depth, 0, True)
    assert e.value.code == pg.PG_EOVERFLOW
    biased.set_count_bias(TOP - big)
    got, _ = biased.perft_batch(arr, depth, 0, True)
    assert [int(x) for x in got] == [TOP - big + w for w in wants]
    biased.set_count_bias(0)
    got, _ = biased.perft_batch(arr, depth, 0, True)
    assert [int(x) for x in got] == wants
