"""Timing sweeps of the start position that the measurement notes quote.

    python3 -m perft_gpu_amd.sweep leafmodes   # perft(7)/(8) per slider variant, fused vs materialized leaf
    python3 -m perft_gpu_amd.sweep tt [depth]  # TT perft(7..depth) with 16 GB and 64 GB tables

Every count is checked against the published start-position values.
"""
import sys
import time

import perft_gpu_amd as pg
from perft_gpu_amd import workloads as W


def _check(depth, n):
    want = W.CPW_COUNTS["start"][depth] if depth < len(W.CPW_COUNTS["start"]) else None
    if want is not None and n != want:
        raise SystemExit("MISMATCH: start perft(%d) = %d, expected %d" % (depth, n, want))
    return "exact" if want is not None else "unchecked"


def leafmodes():
    for mode in (0, 1):
        with pg.Device(0, arena_bytes=48 << 30, timing=True, leaf_mode=mode) as d:
            for v in range(4):
                for depth in (7, 8):
                    t = time.perf_counter()
                    n = d.perft(pg.START_FEN, depth, v)
                    dt = time.perf_counter() - t
                    st = d.last_stats()
                    print("leaf_mode %d %-14s d%d %d %s %.4f s %7.1f G leaves/s  leaf kernel %.2f ms  leaf items %d"
                          % (mode, pg.SLIDER_NAMES[v], depth, n, _check(depth, n), dt, n / dt / 1e9,
                             st["leaf_kernel_ms"], st["leaf_items"]), flush=True)


def tt(top=10):
    with pg.Device(0, arena_bytes=96 << 30, timing=True) as d:
        for gb in (16, 64):
            with pg.TranspositionTable(d, gb << 30) as table:
                for depth in range(7, top + 1):
                    t = time.perf_counter()
                    n = table.perft(pg.START_FEN, depth, 0)
                    dt = time.perf_counter() - t
                    st = d.last_stats()
                    print("TT %2d GB d%d %d %s %.3f s  peak arena %.1f GB" % (gb, depth, n, _check(depth, n), dt,
                                                                           st["peak_arena"] / 1e9), flush=True)
                c = table.counters()
                print("TT %d GB counters per remaining depth r (probes, hits, stores, in-level duplicates):" % gb)
                for r in range(2, top + 1):
                    print("  r=%d %s" % (r, " ".join(str(int(x)) for x in c[r])))


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    if not argv or argv[0] not in ("leafmodes", "tt"):
        raise SystemExit(__doc__)
    if argv[0] == "leafmodes":
        leafmodes()
    else:
        tt(int(argv[1]) if len(argv) > 1 else 10)


if __name__ == "__main__":
    main()
