#!/usr/bin/env python3
"""Writes the golden fixtures in tests/golden/ from the reference's own CPU
code (oracle/_ref/libperft_ref.so, compiled from the reference checkout by
oracle/Makefile).  Run in a container that has that library; the JSON files it
writes are committed.  Positions for the random configurations come from the
engine's seeded generator and are stored as FEN and as the 48-byte words.

  cpw.json        perft(d) of the CPW positions listed in perft.cu:62-70
  zobrist.json    computeZobristKey128b of positions, and makeMove<.., true>
                  (128-bit incremental key + child board) for every legal move
  movegen.json    countMoves / generateMoves (sorted) of positions
  random.json     perft(1..4) of 256 random positions (ply 0-80)
  c2.json         the 16 C2 positions, perft(1..6)
  c3.json         the 4096 C3 positions, perft(5)
  c5.json         served-query positions/depths and their perft
  edge.json       special-rule positions (en passant that uncovers check,
                  castling through/into check, promotions out of check,
                  double check, stalemate), perft(1..d)
"""
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

import perft_gpu_amd as pg  # noqa: E402
from perft_gpu_amd import workloads as W  # noqa: E402
from oracle import ref  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))
THREADS = os.cpu_count() or 4


def words(p):
    return [int(x) for x in p]


def fen_of(w):
    return pg.pos_to_fen(pg.pos_from_words(w))


def dump(name, obj):
    with open(os.path.join(OUT, name), "w") as f:
        json.dump(obj, f, indent=0, separators=(",", ":"))
    print("wrote", name)


def main(which):
    R = ref()
    assert R is not None, "oracle/_ref/libperft_ref.so not built (make -C oracle ref)"
    t0 = time.time()

    if "cpw" in which:
        depth_cap = {"start": 7, "kiwipete": 5, "pos3": 7, "pos4": 6, "pos4_mirror": 6, "pos5_old": 5, "max218": 4}
        out = {}
        for name, fen in W.CPW.items():
            p = R.fen(fen)
            counts = [1] + [R.perft_mt(p, d, THREADS) for d in range(1, depth_cap[name] + 1)]
            out[name] = {"fen": fen, "words": words(p), "perft": counts}
        dump("cpw.json", out)

    if "edge" in which:
        # well-known special-rule test positions (community "perft tests" set)
        edge = [("3k4/3p4/8/K1P4r/8/8/8/8 b - - 0 1", 6), ("8/8/4k3/8/2p5/8/B2P2K1/8 w - - 0 1", 6),
                ("8/8/1k6/2b5/2pP4/8/5K2/8 b - d3 0 1", 6), ("5k2/8/8/8/8/8/8/4K2R w K - 0 1", 6),
                ("3k4/8/8/8/8/8/8/R3K3 w Q - 0 1", 6), ("r3k2r/1b4bq/8/8/8/8/7B/R3K2R w KQkq - 0 1", 4),
                ("r3k2r/8/3Q4/8/8/5q2/8/R3K2R b KQkq - 0 1", 4), ("2K2r2/4P3/8/8/8/8/8/3k4 w - - 0 1", 6),
                ("8/8/1P2K3/8/2n5/1q6/8/5k2 b - - 0 1", 5), ("4k3/1P6/8/8/8/8/K7/8 w - - 0 1", 6),
                ("8/P1k5/K7/8/8/8/8/8 w - - 0 1", 6), ("K1k5/8/P7/8/8/8/8/8 w - - 0 1", 6),
                ("8/k1P5/8/1K6/8/8/8/8 w - - 0 1", 7), ("8/8/2k5/5q2/5n2/8/5K2/8 b - - 0 1", 4),
                ("rnbqkbnr/pp1ppppp/8/2pP4/8/8/PPP1PPPP/RNBQKBNR w KQkq c6 0 2", 5),
                ("8/2p5/8/KP5r/8/8/8/7k w - - 0 1", 6)]
        items = []
        for fen, d in edge:
            p = R.fen(fen)
            items.append({"fen": fen, "words": words(p), "perft": [1] + [R.perft_mt(p, k, THREADS) for k in range(1, d + 1)]})
        dump("edge.json", items)

    rnd, plies = pg.random_positions(0x5EED0001, 256, 0, 80)
    rnd = [tuple(int(x) for x in w) for w in rnd]

    if "zobrist" in which:
        items = []
        for w in rnd[:64] + [R.fen(f) for f in W.CPW.values()]:
            key = R.zobrist128(w)
            children = []
            for m in R.generate_moves(w):
                child, ck = R.make_move(w, m, key)
                children.append({"move": m, "child": words(child), "key": [ck[0], ck[1]],
                                 "full": list(R.zobrist128(child))})
            items.append({"words": words(w), "key": [key[0], key[1]], "children": children})
        dump("zobrist.json", items)

    if "movegen" in which:
        items = []
        for w in rnd:
            ms = R.generate_moves(w)
            items.append({"words": words(w), "count": R.count_moves(w), "moves": sorted(m & 0xFFFF for m in ms)})
        dump("movegen.json", items)

    if "random" in which:
        items = []
        for w, ply in zip(rnd, plies):
            items.append({"words": words(w), "fen": fen_of(w), "ply": int(ply),
                          "perft": [1] + [R.perft(w, d) for d in range(1, 5)]})
        dump("random.json", items)

    if "c2" in which:
        pos, pl = W.c2_positions()
        items = []
        for w, ply in zip(pos, pl):
            w = tuple(int(x) for x in w)
            items.append({"words": words(w), "fen": fen_of(w), "ply": int(ply),
                          "perft": [1] + [R.perft_mt(w, d, THREADS) for d in range(1, W.C2_DEPTH + 1)]})
        dump("c2.json", {"seed": W.C2_SEED, "ply": list(W.C2_PLY), "depth": W.C2_DEPTH, "positions": items})

    if "c3" in which:
        pos, pl = W.c3_positions()
        counts = []
        for i, w in enumerate(pos):
            counts.append(R.perft_mt(tuple(int(x) for x in w), W.C3_DEPTH, THREADS))
            if i % 512 == 0:
                print("c3", i, time.time() - t0, flush=True)
        dump("c3.json", {"seed": W.C3_SEED, "ply": list(W.C3_PLY), "depth": W.C3_DEPTH,
                         "words": [[int(x) for x in w] for w in pos], "plies": [int(x) for x in pl],
                         "perft": counts})

    if "c5" in which:
        pos, depths = W.c5_queries(512)
        counts = []
        for w, d in zip(pos, depths):
            counts.append(R.perft_mt(tuple(int(x) for x in w), int(d), THREADS))
        dump("c5.json", {"seed": W.C5_SEED, "n": 512, "words": [[int(x) for x in w] for w in pos],
                         "depths": [int(d) for d in depths], "perft": counts})
    print("done in %.1f s" % (time.time() - t0))


if __name__ == "__main__":
    main(sys.argv[1:] or ["cpw", "edge", "zobrist", "movegen", "random", "c2", "c3", "c5"])
