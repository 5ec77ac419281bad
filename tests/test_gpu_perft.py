"""GPU perft parity: breadth-first perft on the device against published CPW
values, the reference's own CPU perft (golden fixtures) and the oracle."""
import pytest

from conftest import golden
from perft_gpu_amd import workloads as W

pytestmark = pytest.mark.gpu
VARIANTS = [0, 1, 2, 3]


@pytest.mark.parametrize("v", VARIANTS)
def test_start_position_to_depth7(dev, v):
    for d in range(0, 8):
        print("start d%d variant %d" % (d, v), flush=True)
        assert dev.perft(W.CPW["start"], d, v) == W.CPW_COUNTS["start"][d]


@pytest.mark.parametrize("name", ["kiwipete", "pos3", "pos4", "pos4_mirror", "pos5_old", "max218"])
def test_cpw_positions(dev, name):
    item = golden("cpw.json")[name]
    for v in VARIANTS:
        for d, want in enumerate(item["perft"]):
            if d > 5 and v != 0:
                continue
            print(name, "d%d variant %d" % (d, v), flush=True)
            assert dev.perft(item["fen"], d, v) == want


def test_random_positions_depth4(dev, pg):
    for item in golden("random.json")[:128]:
        p = pg.pos_from_words(item["words"])
        for v in VARIANTS:
            assert dev.perft(p, 4, v) == item["perft"][4], (item["fen"], v)


@pytest.mark.parametrize("v", VARIANTS)
def test_c2_random_positions_depth6(dev, pg, v):
    c2 = golden("c2.json")
    for item in c2["positions"]:
        print("c2", item["fen"], "variant", v, flush=True)
        assert dev.perft(pg.pos_from_words(item["words"]), 6, v) == item["perft"][6]


@pytest.mark.parametrize("v", VARIANTS)
def test_start_depth8(dev, v):
    print("start d8 variant %d" % v, flush=True)
    assert dev.perft(W.CPW["start"], 8, v) == W.CPW_COUNTS["start"][8]
    st = dev.last_stats()
    assert st["leaves"] == W.CPW_COUNTS["start"][8]
    assert st["nodes"][1:8] == W.CPW_COUNTS["start"][1:8]


def test_small_arena_forces_splits_and_stays_exact(pg):
    # perft(7) stores ~12 MB for ply 4 and 29 MB of moves and parents for the
    # 4.9e6 ply-5 boards the leaf makes itself (the byte-lookup variant: 250 MB
    # of ply-5 boards and 700 MB of ply-6 items); a 16 MiB arena forces the
    # levels to split (the graph path finds that its levels do not fit and
    # hands over to the host-driven loop)
    with pg.Device(0, arena_bytes=16 << 20) as small:
        for v in (0, 3):
            print("split arena start d7 variant", v, flush=True)
            assert small.perft(W.CPW["start"], 7, v) == W.CPW_COUNTS["start"][7]
            st = small.last_stats()
            assert st["splits"] > 0 and st["peak_arena"] <= 16 << 20
            assert st["nodes"][1:7] == W.CPW_COUNTS["start"][1:7]


def test_divide_matches_oracle(dev, pg, orc):
    p = orc.fen(W.CPW["kiwipete"])
    want = orc.divide(p, 4)
    got = pg.divide(dev, W.CPW["kiwipete"], 4)
    key = lambda m: (m & 0xFFF) | ((((m >> 12) & 3) + 1 if m & 0x8000 else 0) << 12)
    assert {key(m): c for m, c in got.items()} == {key(m): c for m, c in want.items()}
    assert sum(got.values()) == W.CPW_COUNTS["kiwipete"][4]


def test_cli_prints_reference_lines(pg, capsys):
    from perft_gpu_amd import cli
    cli.main([W.CPW["pos3"], "5"])
    out = capsys.readouterr().out
    for d in (3, 4, 5):
        assert "GPU Perft %d: %d," % (d, W.CPW_COUNTS["pos3"][d]) in out


@pytest.mark.parametrize("v", [0, 1, 2, 3])
def test_special_rule_positions(dev, pg, v):
    for item in golden("edge.json"):
        p = pg.pos_from_words(item["words"])
        for d, want in enumerate(item["perft"]):
            assert dev.perft(p, d, v) == want, (item["fen"], d, v)


# positions without a legal move: checkmate (fool's mate, white to move) and
# stalemate (black to move)
TERMINAL = ["rnb1kbnr/pppp1ppp/8/4p3/6Pq/5P2/PPPPP2P/RNBQKBNR w KQkq - 1 3",
            "7k/5Q2/6K1/8/8/8/8/8 b - - 0 1"]


def test_terminal_roots_on_every_path(dev, pg, orc):
    import numpy as np
    boards = [pg.read_fen_string(f) for f in TERMINAL]
    for b in boards:
        assert orc.count_moves(pg.positions_array([b])[0].tolist()) == 0
    for d in range(0, 5):
        want = 1 if d == 0 else 0
        for b in boards:
            print("terminal plain/tt d%d" % d, flush=True)
            assert dev.perft(b, d, 0) == want
    with pg.TranspositionTable(dev, 1 << 30) as tt:
        for d in range(0, 5):
            for b in boards:
                assert tt.perft(b, d) == (1 if d == 0 else 0)
    # terminal roots inside a batch of ordinary ones, and an empty batch
    items = golden("random.json")[:64]
    arr = np.concatenate([pg.positions_array(boards), np.array([it["words"] for it in items], np.uint64)])
    for dedup in (False, True):
        print("terminal batch dedup %d" % dedup, flush=True)
        got, _ = dev.perft_batch(arr, 3, 0, dedup)
        assert [int(x) for x in got] == [0, 0] + [it["perft"][3] for it in items]
        got, _ = dev.perft_batch(pg.positions_array(boards), 3, 0, dedup)
        assert [int(x) for x in got] == [0, 0]
    got, _ = dev.perft_batch(np.zeros((0, 6), np.uint64), 3, 0, True)
    assert len(got) == 0


def test_c_caller_of_the_abi(pg):
    """perft_gpu_amd/perft_c (plain C, linked against the library) prints the
    reference's lines with the published counts, with and without a TT."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(pg.LIB_PATH), "perft_c")
    assert os.path.exists(exe), "perft_c is built with the library (make -C perft_gpu_amd)"
    out = subprocess.run([exe, W.CPW["start"], "6"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    for d in range(1, 7):
        assert "GPU Perft %d: %d," % (d, W.CPW_COUNTS["start"][d]) in out.stdout
    out = subprocess.run([exe, W.CPW["kiwipete"], "4", "1"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert "GPU Perft 4: %d," % W.CPW_COUNTS["kiwipete"][4] in out.stdout


@pytest.mark.parametrize("v", VARIANTS)
def test_graph_path_start_position_and_node_counts(gdev, v):
    """pg_perft as one graph launch: start perft(0..8) and the nodes per ply."""
    for d in range(0, 9):
        print("graph start d%d variant %d" % (d, v), flush=True)
        assert gdev.perft(W.CPW["start"], d, v) == W.CPW_COUNTS["start"][d]
    st = gdev.last_stats()
    assert st["nodes"][1:8] == W.CPW_COUNTS["start"][1:8]
    # the byte-lookup variant materializes ply 7 (166 GB at depth 8): it does
    # not fit the 16 GiB arena, and the host-driven loop splits it instead
    assert (st["splits"] == 0) == (v != 3)


def test_graph_path_reference_positions(gdev, pg):
    for name, item in golden("cpw.json").items():
        for d, want in enumerate(item["perft"]):
            for v in VARIANTS if d <= 5 else [0]:
                assert gdev.perft(item["fen"], d, v) == want, (name, d, v)
    for item in golden("c2.json")["positions"]:
        for v in VARIANTS:
            assert gdev.perft(pg.pos_from_words(item["words"]), 6, v) == item["perft"][6]
    for item in golden("random.json")[:128]:
        assert gdev.perft(pg.pos_from_words(item["words"]), 4, 0) == item["perft"][4]
    for item in golden("edge.json"):
        for d, want in enumerate(item["perft"]):
            assert gdev.perft(pg.pos_from_words(item["words"]), d, 0) == want
    for fen in TERMINAL:
        for d in range(0, 4):
            assert gdev.perft(fen, d, 0) == (1 if d == 0 else 0)


def test_graph_path_overflow_is_reported(gdev, pg):
    want = W.CPW_COUNTS["start"][6]
    try:
        gdev.set_count_bias((1 << 64) - want)
        with pytest.raises(pg.PerftError) as e:
            gdev.perft(W.CPW["start"], 6, 0)
        assert e.value.code == pg.PG_EOVERFLOW
        gdev.set_count_bias((1 << 64) - 1 - want)
        assert gdev.perft(W.CPW["start"], 6, 0) == (1 << 64) - 1
    finally:
        gdev.set_count_bias(0)


def test_divide_falls_back_when_a_child_does_not_fit_a_batch(dev, pg):
    """Kiwipete divided at depth 7 on a 64 MiB arena: a depth-6 child's ply-4
    level (millions of boards) does not fit one batch region, so divide()
    runs one perft per move (the host-driven loop splits levels) and still
    sums to the published total."""
    with pg.Device(0, arena_bytes=64 << 20) as small:
        pos = pg.read_fen_string(W.CPW["kiwipete"])
        kid, _ = pg.make_move_host(pos, pg.generate_moves_host(pos)[0])
        with pytest.raises(pg.PerftError) as e:
            small.perft_batch(pg.positions_array([kid]), 6, 0, True)
        assert e.value.code == pg.PG_ENOMEM
        print("divide kiwipete d7 on 64 MiB", flush=True)
        div = pg.divide(small, W.CPW["kiwipete"], 7)
    assert len(div) == 48 and sum(div.values()) == 374190009323 == dev.perft(W.CPW["kiwipete"], 7, 0)
