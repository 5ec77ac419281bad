"""Batched perft (one graph launch over many roots): each result equals the
single-position result and the reference's value; graph replays are exact
with eager work in between; an arena too small for the batch splits it."""
import numpy as np
import pytest

from conftest import golden
from perft_gpu_amd import workloads as W

pytestmark = pytest.mark.gpu


def random_set(n=256):
    items = golden("random.json")[:n]
    return np.array([it["words"] for it in items], np.uint64), items


@pytest.mark.parametrize("v", [0, 1, 2, 3])
@pytest.mark.parametrize("dedup", [False, True])
def test_batch_depth4_matches_reference(dev, v, dedup):
    arr, items = random_set()
    print("batch d4 variant", v, "dedup", dedup, flush=True)
    got, st = dev.perft_batch(arr, 4, v, dedup)
    assert [int(x) for x in got] == [it["perft"][4] for it in items]
    assert st["launches"] == 1


@pytest.mark.parametrize("depth", [0, 1, 2, 3])
def test_batch_shallow_depths(dev, depth):
    arr, items = random_set(64)
    got, _ = dev.perft_batch(arr, depth, 0)
    assert [int(x) for x in got] == [it["perft"][depth] for it in items]


def test_batch_mixed_sides_and_duplicates(dev, pg):
    # the same position several times, both colours interleaved
    arr, items = random_set(32)
    idx = np.array([3, 3, 0, 7, 3, 1, 0, 5, 6, 2, 2], np.int64)
    got, _ = dev.perft_batch(arr[idx], 4, 0, True)
    assert [int(x) for x in got] == [items[i]["perft"][4] for i in idx]


@pytest.mark.parametrize("with_tt", [False, True])
def test_batch_of_mixed_depths(dev, pg, with_tt):
    """pg_perft_batch_mixed: a depth per root (0-4, shuffled, the same board
    at several depths), against the reference's values; then the c2
    positions at depth 6 in the same call as random ones at depth 3."""
    arr, items = random_set(200)
    rng = np.random.default_rng(7)
    depths = rng.integers(0, 5, size=len(arr)).astype(np.uint32)
    idx = np.concatenate([np.arange(len(arr)), [5, 5, 5]])
    depths = np.concatenate([depths, [0, 2, 4]]).astype(np.uint32)
    tt = pg.TranspositionTable(dev, 1 << 30) if with_tt else None
    try:
        print("mixed-depth batch tt", with_tt, flush=True)
        got, st = dev.perft_batch_mixed(arr[idx], depths, 0, True, tt=tt)
        assert [int(x) for x in got] == [items[i]["perft"][d] for i, d in zip(idx, depths)]
        assert st["launches"] == len(set(depths.tolist()) - {0})  # depth 0 needs no launch
        c2 = golden("c2.json")["positions"]
        both = np.concatenate([np.array([x["words"] for x in c2], np.uint64), arr[:32]])
        d2 = np.array([6] * len(c2) + [3] * 32, np.uint32)
        got, _ = dev.perft_batch_mixed(both, d2, 0, True, tt=tt)
        assert [int(x) for x in got] == [x["perft"][6] for x in c2] + [it["perft"][3] for it in items[:32]]
    finally:
        if tt:
            tt.close()
    got, _ = dev.perft_batch_mixed(arr[:0], np.zeros(0, np.uint32))
    assert len(got) == 0
    with pytest.raises(pg.PerftError) as e:
        dev.perft_batch_mixed(arr[:2], np.array([3, 16], np.uint32))
    assert e.value.code == pg.PG_EINVAL
    with pytest.raises(ValueError):
        dev.perft_batch_mixed(arr[:2], [3])


def test_graph_replays_with_eager_work_between(dev, pg):
    c2 = golden("c2.json")["positions"]
    arr = np.array([x["words"] for x in c2], np.uint64)
    want = [x["perft"][5] for x in c2]
    for rep in range(3):
        print("replay", rep, flush=True)
        got, st = dev.perft_batch(arr, 5, 0, True)
        assert [int(x) for x in got] == want
        # eager single-position work on the same context between replays
        assert dev.perft(W.CPW["kiwipete"], 4, 0) == W.CPW_COUNTS["kiwipete"][4]


def test_batch_equals_single_position_path(dev, pg):
    c2 = golden("c2.json")["positions"]
    arr = np.array([x["words"] for x in c2], np.uint64)
    got, _ = dev.perft_batch(arr, 6, 0, True)
    for g, x in zip(got, c2):
        assert int(g) == x["perft"][6] == dev.perft(pg.pos_from_words(x["words"]), 6, 0)


def test_batch_splits_when_region_small(pg):
    # 128 roots at depth 4: about 7 MB for the deepest level (keys, perfts,
    # owners, states, moves, parents and the collapse set); 2 MiB forces splits
    arr, items = random_set(128)
    with pg.Device(0, arena_bytes=2 << 20) as small:
        got, st = small.perft_batch(arr, 4, 0, True)
        assert [int(x) for x in got] == [it["perft"][4] for it in items]
        assert st["splits"] > 0


def test_batch_with_shared_tt(dev, pg):
    arr, items = random_set(128)
    with pg.TranspositionTable(dev, 1 << 30) as tt:
        for rep in range(2):
            got, _ = dev.perft_batch(arr, 4, 0, True, tt=tt)
            assert [int(x) for x in got] == [it["perft"][4] for it in items]
        assert tt.counters()[4, 1] > 0   # second pass hits the roots


@pytest.mark.slow
def test_c3_batch_4096_positions_depth5(dev):
    c3 = golden("c3.json")
    arr = np.array(c3["words"], np.uint64)
    print("c3 batch", len(arr), flush=True)
    got, st = dev.perft_batch(arr, 5, 0, True)
    assert [int(x) for x in got] == c3["perft"]
    assert st["launches"] == 1


def test_batch_grid_sizes_never_change_counts(dev):
    arr, items = random_set(256)
    want = [it["perft"][4] for it in items]
    try:
        for g in (1, 64, 1024, 0, -1):
            print("batch grid", g, flush=True)
            dev.set_batch_grid(g)
            got, _ = dev.perft_batch(arr, 4, 0, True)
            assert [int(x) for x in got] == want
    finally:
        dev.set_batch_grid(-1)


def test_split_depth_and_autotuner(dev, pg):
    """A batch expanded one or two plies on the host first (split depth)
    folds back to the same per-root counts; the autotuner's choice, run on
    the whole set, reproduces the reference's values."""
    from perft_gpu_amd import autotune
    arr, items = random_set(128)
    want = np.array([it["perft"][4] for it in items], np.uint64)
    for split in (0, 1, 2):
        for grid in (-1, 0):
            cfg = {"slider": 0, "dedup": True, "blocks_per_cu": 8, "split": split, "grid": grid, "chunk": None}
            assert np.array_equal(autotune.run(dev, arr, 4, cfg), want), cfg
    cfg = {"slider": 2, "dedup": False, "blocks_per_cu": 4, "split": 1, "grid": 0, "chunk": 32}
    assert np.array_equal(autotune.run(dev, arr, 4, cfg), want)
    for tpp in (16, 64, 4096):   # threads per root: 16 puts the 128 roots on 8 work-groups
        cfg = {"slider": 0, "dedup": True, "blocks_per_cu": 8, "split": 0, "grid": -1, "chunk": None, "tpp": tpp}
        assert np.array_equal(autotune.run(dev, arr, 4, cfg), want), cfg
    space = {"slider": [0, 2], "dedup": [True], "blocks_per_cu": [8], "split": [0, 1], "grid": [-1, 0],
             "tpp": [256, 1024], "split_deeper": [2]}
    best, table = autotune.tune(dev, arr, 4, space=space, sample=64)
    assert len(table) == 8 + 1 + 2 and best in [c for c, _, _ in table]
    assert table[8][0]["split"] == 2 and [c.get("tpp") for c, _, _ in table[9:]] == [256, 1024]
    assert np.array_equal(autotune.run(dev, arr, 4, best), want)
