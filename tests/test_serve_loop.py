"""The serving loop's admission decision (pg_serve_policy, host only, no
device): the depth whose oldest query arrived first goes next; long depths
keep to their share of the slots; a batch takes as many queries as
max_batch and the slot's region (by the per-depth estimate) allow, and at
least one; and the work-group cap a long batch runs under
(pg_serve_wg_cap)."""
from perft_gpu_amd.serve import policy

CFG = dict(slots=4, max_batch=1024, long_depth=6, long_slots=2, long_wg_per_cu=2)
EST = [4096.0] * 16


def pick(waiting, oldest, busy_long=0, idle=4, region=16 << 30, est=EST, **cfg):
    w, o = [0] * 16, [0.0] * 16
    for d, n in waiting.items():
        w[d] = n
        o[d] = oldest[d]
    c = dict(CFG)
    c.update(cfg)
    return policy(w, o, busy_long, idle, region, est, **c)


def test_oldest_depth_first():
    assert pick({3: 5, 5: 2, 7: 1}, {3: 2.0, 5: 1.0, 7: 0.5}) == (7, 1)
    assert pick({3: 5, 5: 2}, {3: 2.0, 5: 1.0}) == (5, 2)
    assert pick({3: 5, 4: 9}, {3: 0.25, 4: 1.0}) == (3, 5)


def test_long_depths_keep_to_their_slots():
    waiting, oldest = {4: 3, 6: 8, 7: 1}, {4: 3.0, 6: 1.0, 7: 0.5}
    assert pick(waiting, oldest, busy_long=1) == (7, 1)
    assert pick(waiting, oldest, busy_long=2) == (4, 3)           # both long slots taken
    assert pick({6: 4}, {6: 0.0}, busy_long=2) is None            # only long queries wait: they wait
    assert pick({6: 4}, {6: 0.0}, busy_long=2, long_slots=3) == (6, 4)


def test_batch_size_bounded_by_region_and_max_batch():
    est = list(EST)
    est[5] = 1e6
    assert pick({5: 50}, {5: 0.0}, region=10_000_000, est=est) == (5, 10)
    assert pick({5: 50}, {5: 0.0}, region=10_000_000, est=est, max_batch=4) == (5, 4)
    est[5] = 1e12                                                  # one query exceeds the region estimate
    assert pick({5: 50}, {5: 0.0}, region=10_000_000, est=est) == (5, 1)


def test_nothing_to_start():
    assert pick({3: 5}, {3: 0.0}, idle=0) is None
    assert pick({}, {}) is None


def test_long_batch_work_group_cap():
    from perft_gpu_amd.serve import wg_cap
    c = dict(long_depth=6, long_wg_per_cu=2, light_wg_per_cu=1, light_queue=16)
    assert wg_cap(5, 0, True, **c) == 0                   # short batches run uncapped
    assert wg_cap(6, 3, True, **c) == 1                   # light load with short queries about
    assert wg_cap(6, 16, True, **c) == 1
    assert wg_cap(6, 17, True, **c) == 2                  # a queue past light_queue: throughput
    assert wg_cap(7, 0, False, **c) == 2                  # no short queries about
    assert wg_cap(9, 0, True, **dict(c, light_wg_per_cu=3)) == 3
