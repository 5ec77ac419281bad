"""GPU parity of the device move generator (one thread per board) against
the oracle and the reference's golden outputs, for each slider variant."""
import numpy as np
import pytest

from conftest import golden
from test_host import slider_cases

pytestmark = pytest.mark.gpu
VARIANTS = [0, 1, 2, 3]


def positions_and_children(orc, n_parents=256):
    items = golden("movegen.json")[:n_parents]
    words = [tuple(it["words"]) for it in items]
    kids = []
    for w in words:
        kids.extend(orc.generate_boards(w))
    return words + kids


@pytest.mark.parametrize("v", VARIANTS)
def test_device_count_moves(dev, orc, v):
    boards = positions_and_children(orc)
    arr = np.array(boards, dtype=np.uint64)
    print("count_moves variant", v, "boards", len(arr), flush=True)
    got = dev.count_moves(arr, v)
    want = np.array([orc.count_moves(b) for b in boards], dtype=np.uint32)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("v", VARIANTS)
def test_device_generate_moves(dev, orc, v):
    boards = positions_and_children(orc, 96)
    print("generate_moves variant", v, "boards", len(boards), flush=True)
    got = dev.generate_moves(np.array(boards, dtype=np.uint64), v)
    key = lambda m: (m & 0xFFF) | ((((m >> 12) & 3) + 1 if m & 0x8000 else 0) << 12)
    for b, g in zip(boards, got):
        assert sorted(key(int(m)) for m in g) == sorted(key(m) for m in orc.generate_moves(b))


def test_device_make_move_with_keys(dev, orc):
    items = golden("zobrist.json")
    parents, moves, keys, want_child, want_key = [], [], [], [], []
    for it in items:
        for ch in it["children"]:
            parents.append(it["words"])
            moves.append(ch["move"])
            keys.append(it["key"])
            want_child.append(ch["child"])
            want_key.append(ch["key"])
    print("make_move items", len(moves), flush=True)
    out, k = dev.make_moves(np.array(parents, np.uint64), np.array(moves, np.uint16), np.array(keys, np.uint64))
    assert np.array_equal(out, np.array(want_child, np.uint64))
    assert np.array_equal(k, np.array(want_key, np.uint64))


@pytest.mark.parametrize("v", VARIANTS)
@pytest.mark.parametrize("rook", [0, 1])
def test_device_slider_lookups_every_occupancy(dev, v, rook):
    # all relevant occupancies of all squares (rook: 102400, bishop: 5248)
    sq, occ, want = slider_cases(rook)
    print("slider variant", v, "rook", rook, "cases", len(sq), flush=True)
    got = dev.slider_attacks(v, rook, sq, occ)
    assert np.array_equal(got, want)


def test_device_rejects_kingless_board(dev, pg):
    bad = np.zeros((1, 6), np.uint64)
    with pytest.raises(pg.PerftError):
        dev.count_moves(bad, 0)
