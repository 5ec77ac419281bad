"""Autotuner for batched perft.

For a batch workload (positions, depth) it times, on a seeded sample of the
positions, every combination of
  slider         slider-attack variant (Kogge-Stone, plain / fixed-shift /
                 byte-lookup magics)
  dedup          collapse identical positions within a level
  blocks_per_cu  work-groups per CU of the persistent grid-stride kernels
  split          split depth: the positions are first expanded `split` plies
                 on the host (generateBoards) and perft(depth - split) of all
                 the resulting positions runs as one batch, folded back per
                 root -- the batched analogue of the reference's launch depth
                 (estimateLaunchDepth, launcher.h:1535-1563; perft.cu:117-120)
  grid           work-groups per kernel, i.e. how many threads share the
                 batch's positions: -1 sized to the batch's largest level, 0
                 the whole chip (pg_set_batch_grid)
  chunk          positions per graph launch (None: the whole batch); the
                 default space holds only None (one launch per batch), and
                 callers may pass candidates
and then, on the fastest of those, every value of
  tpp            threads per root position: the grid sized so that this many
                 threads share each root (work-groups = tpp x roots / 256)
  split_deeper   split depths beyond the first search's (two host plies: a
                 root's ~1000 grandchildren each become a batch root)
and returns the fastest of all.  Every timed run is checked against the first
configuration's counts, so a configuration that changed a result could never
win.  The reference has no autotuner; its choices are compile-time switches
(switches.h) and a fixed launch depth.
"""
import time

import numpy as np

from . import SLIDER_NAMES, children_host

DEFAULT_SPACE = {
    "slider": [0, 1, 2, 3],
    "dedup": [True, False],
    "blocks_per_cu": [4, 8],
    "split": [0, 1],
    "grid": [-1, 0],
    "chunk": [None],
    "tpp": [64, 256, 1024, 4096],
    "split_deeper": [2],
}


def grid_for_tpp(tpp, roots):
    """Work-groups of 256 threads for tpp threads per root."""
    return max(1, -(-tpp * roots // 256))


def work_groups(dev_cus, n, depth, grid, blocks_per_cu):
    """Work-groups of the batch's grid-stride kernels (mirrors pg_set_batch_grid)."""
    full = dev_cus * 64
    if grid > 0:
        return grid
    if grid == 0:
        return dev_cus * blocks_per_cu
    wgs = n / 256.0
    for _ in range(2, depth):
        if wgs >= full:
            break
        wgs *= 48.0
    if wgs >= full:
        return dev_cus * blocks_per_cu
    g = 64
    while g < wgs:
        g <<= 1
    return min(g, dev_cus * blocks_per_cu)


def run(dev, arr, depth, cfg):
    """perft(depth) of every row of arr under one configuration."""
    split = min(cfg.get("split", 0), depth)
    work, owner = (children_host(arr, split) if split else (arr, np.arange(len(arr))))
    d = depth - split
    chunk = cfg.get("chunk") or len(work)
    dev.set_batch_grid(grid_for_tpp(cfg["tpp"], len(arr)) if cfg.get("tpp") else cfg.get("grid", -1))
    try:
        parts = []
        for i in range(0, len(work), chunk):
            c, _ = dev.perft_batch(work[i:i + chunk], d, cfg["slider"], cfg["dedup"], cfg["blocks_per_cu"])
            parts.append(c)
    finally:
        dev.set_batch_grid(-1)
    counts = np.concatenate(parts) if parts else np.zeros(0, np.uint64)
    out = np.zeros(len(arr), dtype=np.uint64)
    np.add.at(out, owner, counts)
    return out


def tune(dev, positions, depth, space=None, sample=512, seed=1, repeats=1, log=None, cus=256):
    """Returns (best_config, table of (config, seconds, threads_per_position))."""
    space = dict(DEFAULT_SPACE, **(space or {}))
    arr = np.ascontiguousarray(positions, dtype=np.uint64).reshape(-1, 6)
    rng = np.random.default_rng(seed)
    if len(arr) > sample:
        arr = arr[np.sort(rng.choice(len(arr), sample, replace=False))]
    ref = None
    table = []

    def timed(cfg):
        nonlocal ref
        run(dev, arr, depth, cfg)      # graph capture + warm caches
        best = float("inf")
        for _ in range(repeats):
            t = time.perf_counter()
            got = run(dev, arr, depth, cfg)
            best = min(best, time.perf_counter() - t)
        if ref is None:
            ref = got
        elif not np.array_equal(ref, got):
            raise RuntimeError("autotune: counts differ for %s" % cfg)
        return best

    for s in space["slider"]:
        for dd in space["dedup"]:
            for b in space["blocks_per_cu"]:
                for sp in space["split"]:
                    for g in space["grid"]:
                        for ch in space["chunk"]:
                            cfg = {"slider": s, "dedup": dd, "blocks_per_cu": b, "split": sp, "grid": g, "chunk": ch}
                            best = timed(cfg)
                            n_work = len(arr) * (30 ** sp)
                            tpp = work_groups(cus, n_work, depth - sp, g, b) * 256 / max(1, len(arr))
                            table.append((cfg, best, tpp))
                            if log:
                                log("autotune %s dedup=%d bpc=%d split=%d grid=%s chunk=%s: %.2f ms"
                                    % (SLIDER_NAMES[s], dd, b, sp, "batch" if g < 0 else ("chip" if g == 0 else g),
                                       ch, best * 1e3))
    # threads per root position and deeper splits, on the fastest configuration so far
    base = min(table, key=lambda x: x[1])[0]
    for sp in space.get("split_deeper") or []:
        cfg = dict(base, split=sp)
        best = timed(cfg)
        n_work = len(arr) * (30 ** sp)
        table.append((cfg, best, work_groups(cus, n_work, depth - sp, base["grid"], base["blocks_per_cu"]) * 256
                      / max(1, len(arr))))
        if log:
            log("autotune %s dedup=%d bpc=%d split=%d grid=%s: %.2f ms"
                % (SLIDER_NAMES[base["slider"]], base["dedup"], base["blocks_per_cu"], sp,
                   "batch" if base["grid"] < 0 else "chip", best * 1e3))
    for tpp in space.get("tpp") or []:
        cfg = dict(base, tpp=tpp)
        best = timed(cfg)
        table.append((cfg, best, float(grid_for_tpp(tpp, len(arr)) * 256) / max(1, len(arr))))
        if log:
            log("autotune %s dedup=%d bpc=%d split=%d %d threads per position: %.2f ms"
                % (SLIDER_NAMES[base["slider"]], base["dedup"], base["blocks_per_cu"], base["split"], tpp, best * 1e3))
    best_cfg = min(table, key=lambda x: x[1])[0]
    return best_cfg, table
