"""Deep perft with 128-bit totals and divided output (perft_bb_cpu_launcher /
dividedPerft, launcher.h:1181-1450), and the records mode with .op resume
(processPerftRecords, launcher.h:1578-1695)."""
import collections
import os

import pytest

from perft_gpu_amd import workloads as W

pytestmark = pytest.mark.gpu


def test_divided_lines_sum_to_published_total(dev, pg):
    """start position depth 6 walked on the host down to depth 4: one line per
    move of the root (perft 5 each) and of each depth-5 node (perft 4 each);
    the root lines sum to the published total and each equals the child's own
    device perft; each child's lines sum to its line."""
    lines = []
    total, st = pg.perft_deep(dev, W.CPW["start"], 6, launch_depth=4, divided_depth=4,
                              on_divide=lambda p, c: lines.append((p, c)))
    assert total == W.CPW_COUNTS["start"][6]
    root = [(p, c) for p, c in lines if p.count(" ") == 1]
    second = [(p, c) for p, c in lines if p.count(" ") == 2]
    assert len(root) == 20 and len(second) == 400
    assert sum(c for _, c in root) == total
    assert all(p.startswith("..") and p[4] in "-*" for p, _ in root)
    by_parent = collections.defaultdict(int)
    for p, c in second:
        by_parent[p[:8]] += c
    assert {p: c for p, c in root} == dict(by_parent)
    # each root line equals the device perft of that child
    start = pg.read_fen_string(W.CPW["start"])
    for m in pg.generate_moves_host(start):
        child, _ = pg.make_move_host(start, m)
        fr, to = m & 63, (m >> 6) & 63
        name = "..%s%d-%s%d " % ("abcdefgh"[fr & 7], fr // 8 + 1, "abcdefgh"[to & 7], to // 8 + 1)
        assert dict(root)[name] == dev.perft(child, 5, 0)
    assert st["device_calls"] + st["host_tt_hits"] >= 400 and st["host_nodes"] >= 21


def test_deep_with_table_and_host_memo(dev, pg):
    """Kiwipete depth 5 (captures print '*') and start depth 9 with a table:
    the published totals; transpositions among host nodes hit the host table."""
    with pg.TranspositionTable(dev, 4 << 30) as tt:
        lines = []
        total, _ = pg.perft_deep(dev, W.CPW["kiwipete"], 5, 3, 4, tt, on_divide=lambda p, c: lines.append((p, c)))
        assert total == W.CPW_COUNTS["kiwipete"][5]
        assert len(lines) == 48 and any("*" in p for p, _ in lines)
        print("deep start d9, launch depth 6", flush=True)
        total, st = pg.perft_deep(dev, W.CPW["start"], 9, 6, 8, tt)
        assert total == W.CPW_COUNTS["start"][9] and st["device_calls"] == 8902
    # transpositions start at ply 3 (1.Rb3 Rh6 2.e3 = 1.e3 Rh6 2.Rb3): with the
    # launch depth at 1 the ply-3 nodes of position 3 are host nodes
    print("deep pos3 d5, launch depth 1", flush=True)
    total, st = pg.perft_deep(dev, W.CPW["pos3"], 5, 1, 10)
    assert total == W.CPW_COUNTS["pos3"][5]
    assert st["host_tt_hits"] > 0 and st["host_nodes"] == 1 + 14 + 191 + 2812
    total, st = pg.perft_deep(dev, W.CPW["pos3"], 5, 10, 10)   # depth <= launch depth: one device call
    assert total == W.CPW_COUNTS["pos3"][5] and st["device_calls"] == 1 and st["host_nodes"] == 0


def test_totals_past_2_pow_64_are_exact(dev, pg):
    """With every device result biased by 2^63 (pg_test_count_bias), twenty
    root moves sum past 2^64: the 128-bit total is exact."""
    want = W.CPW_COUNTS["start"][5]
    try:
        dev.set_count_bias(1 << 63)
        total, st = pg.perft_deep(dev, W.CPW["start"], 5, 4, 10)
    finally:
        dev.set_count_bias(0)
    assert st["device_calls"] == 20
    assert total == want + 20 * (1 << 63) and total >= 1 << 64


def test_disk_table_resume(dev, pg, tmp_path):
    """The disk table (perfthash.dat layout) of the nodes two plies below the
    root: a table that holds one root move's subtree -- a run stopped after
    that move -- answers those nodes and the rest are computed and written; a
    full table answers every node with no device call.  Each total is the
    published value."""
    path = str(tmp_path / "perfthash.dat")
    want = W.CPW_COUNTS["kiwipete"][6]
    root = pg.read_fen_string(W.CPW["kiwipete"])
    child, _ = pg.make_move_host(root, pg.generate_moves_host(root)[0])
    first = {pg.zobrist128(pg.make_move_host(child, m)[0]) for m in pg.generate_moves_host(child)}
    ply2 = {pg.zobrist128(pg.make_move_host(pg.make_move_host(root, a)[0], b)[0])
            for a in pg.generate_moves_host(root) for b in pg.generate_moves_host(pg.make_move_host(root, a)[0])}
    sub, st = pg.perft_deep(dev, child, 5, 4, 10, hash_path=path, disk_depth=4)
    assert sub == dev.perft(child, 5, 0)
    assert (st["disk_stores"], st["disk_hits"]) == (len(first), 0)
    assert os.path.getsize(path) >= 4096 * 32 and os.path.getsize(path) % 32 == 0
    print("kiwipete d6 with the partial table", flush=True)
    total, st = pg.perft_deep(dev, W.CPW["kiwipete"], 6, 4, 10, hash_path=path, disk_depth=4)
    assert total == want
    # 2039 ply-2 nodes; a position met again (a transposition) is read back
    # from the table too
    assert st["disk_stores"] == len(ply2 - first) and st["disk_hits"] == 2039 - len(ply2 - first)
    total, st = pg.perft_deep(dev, W.CPW["kiwipete"], 6, 4, 10, hash_path=path, disk_depth=4)
    assert total == want
    assert (st["disk_hits"], st["disk_stores"], st["device_calls"]) == (2039, 0, 0)


def test_cli_divided_output(pg, capsys):
    from perft_gpu_amd import cli
    cli.main([W.CPW["pos3"], "5", "1", "3", "--tt", "1", "--min-depth", "4"])
    out = capsys.readouterr().out
    for d in (4, 5):
        assert "Perft(%02d):%20s, time:" % (d, W.CPW_COUNTS["pos3"][d]) in out


def unique_positions_after(pg, plies):
    """{fen: occurrences} of the positions after `plies` moves from the start."""
    level = {pg.zobrist128(pg.read_fen_string(W.CPW["start"])): (pg.read_fen_string(W.CPW["start"]), 1)}
    for _ in range(plies):
        nxt = {}
        for p, n in level.values():
            for m in pg.generate_moves_host(p):
                c, _ = pg.make_move_host(p, m)
                k = pg.zobrist128(c)
                nxt[k] = (c, nxt[k][1] + n) if k in nxt else (c, n)
        level = nxt
    return [(pg.pos_to_fen(p), n) for p, n in level.values()]


def test_records_mode_and_resume(dev, pg, tmp_path):
    """The 5362 distinct positions after three plies with their occurrence
    counts: sum of count x perft(2) = start perft(5).  A run cut short resumes
    after the lines it wrote and ends with the same file; a finished file is
    left alone."""
    recs = unique_positions_after(pg, 3)
    assert len(recs) == 5362 and sum(n for _, n in recs) == 8902
    path = str(tmp_path / "ply3.fen")
    with open(path, "w") as f:
        for fen, n in recs:
            f.write("%s %d\n" % (fen, n))
    done, skipped = pg.perft_records(dev, path, 2, chunk=1000)
    assert (done, skipped) == (5362, 0)
    full = open(path + ".op").read().splitlines()
    assert len(full) == 5362
    assert sum(int(l.split()[-1]) for l in full) == W.CPW_COUNTS["start"][5]
    for l, (fen, n) in zip(full, recs):
        parts = l.split()
        assert l.startswith("%s %d " % (fen, n)) and int(parts[-1]) == int(parts[-2]) * n
    # interrupted after 2000 lines (and half a line): resume completes it
    with open(path + ".op", "w") as f:
        f.write("\n".join(full[:2000]) + "\n" + full[2000][:20])
    with pg.TranspositionTable(dev, 1 << 30) as tt:
        done, skipped = pg.perft_records(dev, path, 2, tt, chunk=1000)
    assert (done, skipped) == (3362, 2000)
    assert open(path + ".op").read().splitlines() == full
    done, skipped = pg.perft_records(dev, path, 2)
    assert (done, skipped) == (0, 5362)
    assert open(path + ".op").read().splitlines() == full
