"""Autotuner for batched perft.

Picks, for a batch workload (positions, depth), the slider variant, whether
to collapse duplicates within levels, the persistent grid size (work-groups
per CU, i.e. how many threads share the batch's positions) and how many
positions go into one graph launch.  Candidates are timed on a seeded sample
of the batch; every timed run is also checked against the first candidate's
counts, so a configuration that changed a result could never win.
"""
import time

import numpy as np

from . import SLIDER_NAMES

DEFAULT_SPACE = {
    "slider": [0, 1, 2, 3],
    "dedup": [True, False],
    "blocks_per_cu": [4, 8, 16],
    "chunk": [None],             # None = whole batch in one launch
}


def _run(dev, arr, depth, cfg):
    chunk = cfg["chunk"] or len(arr)
    out = []
    for i in range(0, len(arr), chunk):
        c, _ = dev.perft_batch(arr[i:i + chunk], depth, cfg["slider"], cfg["dedup"], cfg["blocks_per_cu"])
        out.append(c)
    return np.concatenate(out)


def tune(dev, positions, depth, space=None, sample=512, seed=1, repeats=1, log=None):
    """Returns (best_config, table of (config, seconds))."""
    space = space or DEFAULT_SPACE
    arr = np.ascontiguousarray(positions, dtype=np.uint64).reshape(-1, 6)
    rng = np.random.default_rng(seed)
    if len(arr) > sample:
        arr = arr[np.sort(rng.choice(len(arr), sample, replace=False))]
    ref = None
    table = []
    for s in space["slider"]:
        for dd in space["dedup"]:
            for b in space["blocks_per_cu"]:
                for ch in space["chunk"]:
                    cfg = {"slider": s, "dedup": dd, "blocks_per_cu": b, "chunk": ch}
                    _run(dev, arr, depth, cfg)      # graph capture + warm caches
                    best = float("inf")
                    for _ in range(repeats):
                        t = time.perf_counter()
                        got = _run(dev, arr, depth, cfg)
                        best = min(best, time.perf_counter() - t)
                    if ref is None:
                        ref = got
                    elif not np.array_equal(ref, got):
                        raise RuntimeError("autotune: counts differ for %s" % cfg)
                    table.append((cfg, best))
                    if log:
                        log("autotune %s dedup=%d bpc=%d chunk=%s: %.2f ms" % (SLIDER_NAMES[s], dd, b, ch, best * 1e3))
    best_cfg = min(table, key=lambda x: x[1])[0]
    return best_cfg, table
