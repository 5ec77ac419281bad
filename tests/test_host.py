"""CPU tests of the product library: every declared symbol is exported, and
the host half of the move generator (same source the kernels compile) agrees
with the oracle and the reference's own outputs.  No compute on a GPU."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT, golden
from perft_gpu_amd import workloads as W


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "perft_gpu.h")).read()
    return sorted(set(re.findall(r"\b(pg_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol(pg):
    L = ctypes.CDLL(pg.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 15
    for n in names:
        assert hasattr(L, n), n


def test_no_gpu_is_reported_not_faked(pg):
    import torch  # noqa: F401  (only to see whether this box has a GPU)
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(pg.PerftError) as e:
        pg.Device(0)
    assert e.value.code == pg.PG_ENODEV


def test_fen_roundtrip(pg, orc):
    for item in golden("random.json"):
        p = pg.read_fen_string(item["fen"])
        assert list(p.words()) == item["words"]
        assert pg.pos_to_fen(p) == item["fen"]
    for fen in W.CPW.values():
        assert list(pg.read_fen_string(fen).words()) == list(orc.fen(fen))


def test_bad_fen_rejected(pg):
    with pytest.raises(pg.PerftError):
        pg.read_fen_string("8/8/8/8/8/8/8/8 w - - 0 1")   # no kings


def test_host_movegen_matches_reference(pg):
    for item in golden("movegen.json"):
        p = pg.pos_from_words(item["words"])
        mv = pg.generate_moves_host(p)
        assert len(mv) == item["count"]
        # flags on non-pawn captures differ by design (device convention)
        key = lambda m: (m & 0xFFF) | ((((m >> 12) & 3) + 1 if m & 0x8000 else 0) << 12)
        assert sorted(key(m) for m in mv) == sorted(key(m) for m in item["moves"])


@pytest.mark.parametrize("v", [0, 1, 2, 3])
def test_host_count_moves_every_variant(pg, orc, v):
    # positions of all kinds: random games, their children and grandchildren
    items = golden("movegen.json")
    boards = [tuple(it["words"]) for it in items]
    for w in list(boards[:96]):
        kids = orc.generate_boards(w)
        boards.extend(kids)
        for kid in kids[:3]:
            boards.extend(orc.generate_boards(kid))
    for fen in W.CPW.values():
        boards.append(orc.fen(fen))
    arr = np.array(boards, np.uint64)
    got = pg.count_moves_host(arr, v)
    want = np.array([orc.count_moves(b) for b in boards], np.uint32)
    assert np.array_equal(got, want), np.nonzero(got != want)[0][:5]


def test_host_kogge_stone_random_occupancies(pg):
    # per-square Kogge-Stone attacks against a ray walk at random occupancies
    rng = np.random.default_rng(5)
    sq = np.arange(64, dtype=np.uint8)
    for trial in range(300):
        occ = int(rng.integers(0, 2**63, dtype=np.int64)) & int(rng.integers(0, 2**63, dtype=np.int64))
        occ |= 1 << int(rng.integers(0, 64))
        # rook attacks from each occupied square, via the host KS lookup
        rooks = [s for s in range(64) if occ >> s & 1]
        got = pg.slider_attacks_host(0, 1, np.array(rooks, np.uint8), np.full(len(rooks), occ, np.uint64))
        for s, a in zip(rooks, got):
            assert int(a) == _ks(s, occ, True)


def test_host_make_move_matches_reference(pg):
    for item in golden("zobrist.json"):
        p = pg.pos_from_words(item["words"])
        assert list(pg.zobrist128(p)) == item["key"]
        for ch in item["children"]:
            child, key = pg.make_move_host(p, ch["move"], tuple(item["key"]))
            assert list(child.words()) == ch["child"]
            assert list(key) == ch["key"]


def _subsets(mask):
    out, s = [], 0
    while True:
        out.append(s)
        s = (s - mask) & mask
        if s == 0:
            return out


def _ks(sq, occ, rook):
    # plain ray walk: the textbook definition the four variants must reproduce
    dirs = [(1, 0), (-1, 0), (0, 1), (0, -1)] if rook else [(1, 1), (1, -1), (-1, 1), (-1, -1)]
    f0, r0 = sq & 7, sq >> 3
    a = 0
    for df, dr in dirs:
        f, r = f0 + df, r0 + dr
        while 0 <= f < 8 and 0 <= r < 8:
            a |= 1 << (r * 8 + f)
            if occ >> (r * 8 + f) & 1:
                break
            f, r = f + df, r + dr
    return a


def slider_cases(rook, stride=1):
    sqs, occs, want = [], [], []
    for sq in range(64):
        f0, r0 = sq & 7, sq >> 3
        mask = 0
        for s in range(64):
            f, r = s & 7, s >> 3
            if s == sq:
                continue
            line = (f == f0 or r == r0) if rook else (abs(f - f0) == abs(r - r0))
            edge = (rook and ((f in (0, 7) and f != f0) or (r in (0, 7) and r != r0))) or \
                   (not rook and (f in (0, 7) or r in (0, 7)))
            if line and not edge:
                mask |= 1 << s
        for occ in _subsets(mask)[::stride]:
            sqs.append(sq)
            occs.append(occ)
            want.append(_ks(sq, occ, rook))
    return np.array(sqs, np.uint8), np.array(occs, np.uint64), np.array(want, np.uint64)


@pytest.mark.parametrize("rook", [0, 1])
def test_host_slider_variants_equal_ray_walk(pg, rook):
    sq, occ, want = slider_cases(rook, stride=3 if rook else 1)
    for v in range(4):
        got = pg.slider_attacks_host(v, rook, sq, occ)
        assert np.array_equal(got, want), (v, rook)


def test_random_position_generator_matches_oracle_regeneration(pg, orc):
    # the same spec re-implemented over the oracle's move list must give the
    # same positions (pins the generator independently of its move order)
    def splitmix(s):
        s = (s + 0x9E3779B97F4A7C15) & (2**64 - 1)
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & (2**64 - 1)
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & (2**64 - 1)
        return s, z ^ (z >> 31)

    def key(m):
        fl = m >> 12
        return (m & 0xFFF) | ((((fl & 3) + 1) if fl & 8 else 0) << 12)

    seed, n, lo, hi = 12345, 24, 0, 70
    got, plies = pg.random_positions(seed, n, lo, hi)
    start = orc.fen(W.START_FEN if hasattr(W, "START_FEN") else pg.START_FEN)
    s = seed
    for i in range(n):
        while True:
            s, r = splitmix(s)
            ply = lo + r % (hi - lo + 1)
            p, ok = start, True
            for _ in range(ply):
                mv = sorted(orc.generate_moves(p), key=key)
                if not mv:
                    ok = False
                    break
                s, r = splitmix(s)
                p, _ = orc.make_move(p, mv[r % len(mv)])
            if ok and orc.count_moves(p) == 0:
                ok = False
            if ok:
                break
        assert tuple(int(x) for x in got[i]) == p, i
        assert plies[i] == ply


def test_workload_fixtures_are_reproduced(pg):
    c2 = golden("c2.json")
    pos, plies = W.c2_positions()
    assert [list(map(int, w)) for w in pos] == [x["words"] for x in c2["positions"]]


@pytest.mark.parametrize("v", [0, 1, 2, 3])
def test_host_count_special_rule_trees(pg, orc, v):
    # every node of the special-rule positions' trees to depth 3
    boards = []
    frontier = [tuple(it["words"]) for it in golden("edge.json")]
    for _ in range(3):
        boards.extend(frontier)
        nxt = []
        for b in frontier:
            nxt.extend(orc.generate_boards(b))
        frontier = nxt
    arr = np.array(boards, np.uint64)
    got = pg.count_moves_host(arr, v)
    want = np.array([orc.count_moves(b) for b in boards], np.uint32)
    assert np.array_equal(got, want), np.nonzero(got != want)[0][:5]


def test_cpu_perft_entry_point(pg):
    """pg_cpu_perft (perft_bb, perft_bb.h:79-101, one host thread) against the
    published counts and the reference's own outputs for random and edge
    positions (fixtures written by the reference compiled here)."""
    for name, depth in (("start", 4), ("kiwipete", 3), ("pos3", 4), ("pos4", 3), ("pos4_mirror", 3)):
        assert pg.cpu_perft(W.CPW[name], depth) == W.CPW_COUNTS[name][depth]
    for item in golden("random.json")[:64] + golden("edge.json"):
        p = pg.pos_from_words(item["words"])
        for d, want in enumerate(item["perft"][:4]):
            assert pg.cpu_perft(p, d) == want, (item["fen"], d)
    with pytest.raises(pg.PerftError):
        pg.cpu_perft(W.CPW["start"], 16)


def test_cli_arena_option():
    """--arena in GiB wins; without a table the library default (0) stands
    (a table search sizes its arena on the device, covered on the GPU)."""
    import argparse
    from perft_gpu_amd import cli
    ns = lambda **k: argparse.Namespace(**dict(dict(arena=0.0, tt=0.0), **k))
    assert cli.arena_bytes(ns()) == 0
    assert cli.arena_bytes(ns(arena=2.0)) == 2 << 30
    assert cli.arena_bytes(ns(arena=1.5, tt=16.0)) == 3 << 29
