"""Seeded workloads of the benchmark configurations.

Positions come from pg_random_positions (start position + uniform-random
legal moves, ply drawn uniformly from a range, terminal positions redrawn).
"""
import numpy as np

from . import START_FEN, random_positions, read_fen_string

# CPW test positions the reference lists in perft.cu:62-70
CPW = {
    "start": START_FEN,
    "kiwipete": "r3k2r/p1ppqpb1/bn2pnp1/3PN3/1p2P3/2N2Q1p/PPPBBPPP/R3K2R w KQkq - 0 1",
    "pos3": "8/2p5/3p4/KP5r/1R3p1k/8/4P1P1/8 w - - 0 1",
    "pos4": "r2q1rk1/pP1p2pp/Q4n2/bbp1p3/Np6/1B3NBn/pPPP1PPP/R3K2R b KQ - 0 1",
    "pos4_mirror": "r3k2r/Pppp1ppp/1b3nbN/nP6/BBP1P3/q4N2/Pp1P2PP/R2Q1RK1 w kq - 0 1",
    "pos5_old": "rnbqkb1r/pp1p1ppp/2p5/4P3/2B5/8/PPP1NnPP/RNBQK2R w KQkq - 0 6",
    "max218": "3Q4/1Q4Q1/4Q3/2Q4R/Q4Q2/3Q4/1Q4Rp/1K1BBNNk w - - 0 1",
}

# published CPW perft values (chessprogramming.org "Perft Results")
CPW_COUNTS = {
    "start": [1, 20, 400, 8902, 197281, 4865609, 119060324, 3195901860, 84998978956, 2439530234167,
              69352859712417, 2097651003696806],
    "kiwipete": [1, 48, 2039, 97862, 4085603, 193690690, 8031647685],
    "pos3": [1, 14, 191, 2812, 43238, 674624, 11030083, 178633661],
    "pos4": [1, 6, 264, 9467, 422333, 15833292, 706045033],
    "pos4_mirror": [1, 6, 264, 9467, 422333, 15833292, 706045033],
    "max218": [1, 218],
}

# plain GPU BFS: start position d7/d8 and 16 random positions at d6
C2_SEED, C2_N, C2_PLY, C2_DEPTH = 0x5EED0002, 16, (10, 60), 6
# batched: 4096 random positions at d5
C3_SEED, C3_N, C3_PLY, C3_DEPTH = 0x5EED0003, 4096, (5, 80), 5
# served queries: positions with ply in [0, 60], depth drawn from C5_DEPTHS
C5_SEED, C5_PLY = 0x5EED0005, (0, 60)
C5_DEPTHS = {3: 0.40, 4: 0.30, 5: 0.20, 6: 0.08, 7: 0.02}


def c2_positions():
    return random_positions(C2_SEED, C2_N, *C2_PLY)


def c3_positions():
    return random_positions(C3_SEED, C3_N, *C3_PLY)


def c5_queries(n, seed=C5_SEED):
    """n served queries: (positions array, depths array)."""
    pos, _ = random_positions(seed, n, *C5_PLY)
    rng = np.random.default_rng(seed)
    ds = np.array(sorted(C5_DEPTHS))
    p = np.array([C5_DEPTHS[d] for d in ds])
    depths = rng.choice(ds, size=n, p=p / p.sum()).astype(np.uint32)
    return pos, depths


def start_position():
    return read_fen_string(START_FEN)
