"""Command line of the reference's perft.cu (perft.cu:14-164):

    python -m perft_gpu_amd.cli <fen> <depth> [numGPUs] [launchDepth] [--tt GB] [--slider N]

Runs perft for every depth from 3 (minDepth, perft.cu:72) up to <depth> and
prints the reference's result line for each (launcher.h:1496-1498).  argv[3]
is the GPU count (the reference's usage text calls it launchdepth,
perft.cu:82); only one GPU is used here.  argv[4], the launch depth, is
accepted and ignored: the level loop sizes its own splits.  With --tt the
transposition-table path runs (the reference's default build,
switches.h:39).
"""
import argparse
import os
import sys
import time

import perft_gpu_amd as pg


def main(argv=None):
    ap = argparse.ArgumentParser(prog="perft_gpu_amd.cli")
    ap.add_argument("fen", nargs="?", default=None)
    ap.add_argument("depth", nargs="?", type=int, default=3)
    ap.add_argument("num_gpus", nargs="?", type=int, default=1)
    ap.add_argument("launch_depth", nargs="?", type=int, default=None)
    ap.add_argument("--tt", type=float, default=0.0, help="transposition table size in GB (0 = no TT)")
    ap.add_argument("--slider", type=int, default=pg.SLIDER_AUTO)
    ap.add_argument("--min-depth", type=int, default=3)
    ap.add_argument("--divide", action="store_true", help="print perft(depth-1) per root move at the last depth")
    ap.add_argument("--tt-file", default=None,
                    help="with --tt: resume the table from this checkpoint if it exists, and save it at the end")
    args = ap.parse_args(argv)
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
    with pg.Device(0) as dev:
        tt = pg.TranspositionTable(dev, int(args.tt * (1 << 30))) if args.tt > 0 else None
        if tt and args.tt_file and os.path.exists(args.tt_file):
            tt.load(args.tt_file)
            print("resumed the transposition table from %s" % args.tt_file)
        try:
            for d in range(min(args.min_depth, args.depth), args.depth + 1):
                t0 = time.perf_counter()
                res = tt.perft(pos, d, args.slider) if tt else dev.perft(pos, d, args.slider)
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


if __name__ == "__main__":
    sys.exit(main())
