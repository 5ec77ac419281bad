"""Command line of the reference's perft.cu (perft.cu:14-164):

    python -m perft_gpu_amd.cli <fen> <depth> [numGPUs] [launchDepth] [--tt GB] [--slider N]
    python -m perft_gpu_amd.cli --records <file> [--depth D] [--tt GB]

Runs perft for every depth from 3 (minDepth, perft.cu:72) up to <depth>.

Without a table each depth is one GPU perft and prints the reference's
no-TT line (launcher.h:1496-1498).  With --tt (the reference's default
build, switches.h:39) each depth is a divided perft (dividedPerft,
launcher.h:1392-1450): nodes deeper than the launch depth (argv[4], default
10) are walked on the host with 128-bit totals, every move of a node deeper
than 10 (DIVIDED_PERFT_DEPTH) prints "<path>   <count>", and the depth ends
with "Perft(NN): <total>, time: .. s, gpuTime: .. s" (launcher.h:1439).

argv[3] is the GPU count (the reference's usage text calls it launchdepth,
perft.cu:82); only one GPU is used here.

--disk-hash FILE keeps the totals of the nodes two plies below the root
in a perfthash.dat-layout file (checkAndCreateDiskHash, launcher.h:1698;
DISK_HASH_LEVEL 2 for depths of DISK_HASH_MIN_DEPTH 11 and up, :38-40), so
a deep run that stops part way resumes from the subtrees it finished.

--records runs the reference's records mode (processPerftRecords,
launcher.h:1578-1695): perft of every FEN line of <file> (its last number
is an occurrence count) into <file>.op, resuming after the lines already
there; the reference hard-codes depth 7.
"""
import argparse
import os
import sys
import time

import perft_gpu_amd as pg

DIVIDED_PERFT_DEPTH = 10
DEFAULT_LAUNCH_DEPTH = 10
DISK_HASH_MIN_DEPTH = 11


def main(argv=None):
    ap = argparse.ArgumentParser(prog="perft_gpu_amd.cli")
    ap.add_argument("fen", nargs="?", default=None)
    ap.add_argument("depth", nargs="?", type=int, default=3)
    ap.add_argument("num_gpus", nargs="?", type=int, default=1)
    ap.add_argument("launch_depth", nargs="?", type=int, default=None)
    ap.add_argument("--tt", type=float, default=0.0, help="transposition table size in GB (0 = no TT)")
    ap.add_argument("--arena", type=float, default=0.0,
                    help="tree arena in GiB (0 = the library default; with --tt, as much as the table leaves free, "
                         "up to 160 GiB)")
    ap.add_argument("--slider", type=int, default=pg.SLIDER_AUTO)
    ap.add_argument("--min-depth", type=int, default=3)
    ap.add_argument("--divide", action="store_true", help="print perft(depth-1) per root move at the last depth")
    ap.add_argument("--tt-file", default=None,
                    help="with --tt: resume the table from this checkpoint if it exists, and save it at the end")
    ap.add_argument("--disk-hash", default=None, help="with --tt: disk table file for the nodes two plies below the root")
    ap.add_argument("--disk-min-depth", type=int, default=DISK_HASH_MIN_DEPTH,
                    help="smallest depth that uses the disk table")
    ap.add_argument("--records", default=None, help="records mode: FEN file (last number of a line = count)")
    ap.add_argument("--records-depth", type=int, default=7)
    args = ap.parse_args(argv)
    if args.records:
        return records(args)
    if args.fen is None or len(args.fen) <= 5:
        print("\nUsage perft_gpu_amd.cli <fen> <depth> [<numGPUs>] [<launchdepth>]")
        print("\nAs no paramaters were provided... running default test")
        fen = pg.START_FEN
    else:
        fen = args.fen
    if args.num_gpus != 1:
        print("\nUsing 1 GPU (this engine runs on a single MI355X)")
    pos = pg.read_fen_string(fen)
    print("No of GPUs detected: 1")
    print(pg.pos_to_fen(pos))
    launch = args.launch_depth if args.launch_depth else DEFAULT_LAUNCH_DEPTH
    with pg.Device(0, arena_bytes=arena_bytes(args)) as dev:
        tt = pg.TranspositionTable(dev, int(args.tt * (1 << 30))) if args.tt > 0 else None
        if tt and args.tt_file and os.path.exists(args.tt_file):
            tt.load(args.tt_file)
            print("resumed the transposition table from %s" % args.tt_file)
        try:
            for d in range(min(args.min_depth, args.depth), args.depth + 1):
                if tt:
                    disk = args.disk_hash if args.disk_hash and d >= args.disk_min_depth else None
                    divided_perft(dev, tt, pos, d, launch, args.slider, disk)
                    continue
                t0 = time.perf_counter()
                res = dev.perft(pos, d, args.slider)
                dt = time.perf_counter() - t0
                print("\nGPU Perft %d: %d,   Time taken: %g seconds, nps: %d" % (d, res, dt, int(res / dt) if dt else 0))
                sys.stdout.flush()
            if args.divide:
                div = pg.divide(dev, pos, args.depth, args.slider, tt)
                for m, c in div.items():
                    print("%s: %d" % (pg.move_to_str(m), c))
                print("total: %d" % sum(div.values()))
            if tt and args.tt_file:
                tt.save(args.tt_file)
                print("saved the transposition table to %s" % args.tt_file)
        finally:
            if tt:
                tt.close()
    return 0


def arena_bytes(args):
    """--arena, or for a table search the arena the table leaves room for."""
    if args.arena > 0:
        return int(args.arena * (1 << 30))
    return pg.table_search_arena(int(args.tt * (1 << 30))) if args.tt > 0 else 0


def divided_perft(dev, tt, pos, depth, launch, slider, disk=None):
    """dividedPerft (launcher.h:1392-1450): the divided lines of every host
    node deeper than DIVIDED_PERFT_DEPTH, then the total with wall and GPU
    time.  disk: a disk table for the nodes of remaining depth depth - 2."""
    def line(path, count):
        print("%s   %20s" % (path, count))
        sys.stdout.flush()

    print()
    t0 = time.perf_counter()
    total, st = pg.perft_deep(dev, pos, depth, min(launch, 15), DIVIDED_PERFT_DEPTH, tt, slider, line, disk)
    dt = time.perf_counter() - t0
    print("Perft(%02d):%20s, time: %8g s, gpuTime: %8g s" % (depth, total, dt, st["device_ms"] / 1e3))
    if disk:
        print("disk table %s: %d nodes read, %d written" % (disk, st["disk_hits"], st["disk_stores"]))
    sys.stdout.flush()
    return total


def records(args):
    print("filename of op: %s.op" % args.records)
    with pg.Device(0, arena_bytes=int(args.arena * (1 << 30))) as dev:
        tt = pg.TranspositionTable(dev, int(args.tt * (1 << 30))) if args.tt > 0 else None
        try:
            t0 = time.perf_counter()
            done, skipped = pg.perft_records(dev, args.records, args.records_depth, tt, args.slider)
            dt = time.perf_counter() - t0
        finally:
            if tt:
                tt.close()
    print("\nRecords done: %d (skipped %d already in the output), Total: %g seconds" % (done, skipped, dt))
    return 0


if __name__ == "__main__":
    sys.exit(main())
