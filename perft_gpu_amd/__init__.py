"""perft_gpu_amd -- chess perft on AMD Instinct MI355X (gfx950).

Python face of the C ABI in include/perft_gpu.h (libperft_gpu_amd.so, built
in-tree).  The functions mirror the reference's host entry points:

  read_fen_string        Utils::readFENString + board088ToHexBB (util.cpp:189, :92)
  init_gpu               initGPU + MoveGeneratorBitboard::init (launcher.h:1506,
                         MoveGeneratorBitboard.h:982)
  perft_launcher         perftLauncher, no-TT branch (launcher.h:1455-1503)
  zobrist128             computeZobristKey128b (MoveGeneratorBitboard.h:3914)

Perft itself only runs on the GPU: every device entry point raises if the
library or a GPU is missing -- there is no CPU fallback.
"""
import ctypes
import os
import time

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libperft_gpu_amd.so")

START_FEN = "rnbqkbnr/pppppppp/8/8/8/8/PPPPPPPP/RNBQKBNR w KQkq - 0 1"

SLIDER_AUTO = -1
SLIDER_KOGGE_STONE = 0
SLIDER_MAGIC_PLAIN = 1
SLIDER_MAGIC_FANCY = 2
SLIDER_MAGIC_BYTE = 3
SLIDER_NAMES = {0: "kogge_stone", 1: "magic_plain", 2: "magic_fancy", 3: "magic_byte_lds"}

PG_OK, PG_EINVAL, PG_ENOMEM, PG_EHIP, PG_ENODEV, PG_ETABLE, PG_EOVERFLOW = 0, -1, -2, -3, -4, -5, -6
_ERR_NAMES = {-1: "EINVAL", -2: "ENOMEM", -3: "EHIP", -4: "ENODEV", -5: "ETABLE", -6: "EOVERFLOW"}

u64 = ctypes.c_uint64


class Pos48(ctypes.Structure):
    """HexaBitBoardPosition (chess.h:208-230), 48 bytes."""
    _fields_ = [(n, u64) for n in ("whitePieces", "pawns", "knights", "bishopQueens", "rookQueens", "kings")]

    def words(self):
        return tuple(getattr(self, n) for n, _ in self._fields_)

    def side(self):
        return int(self.pawns >> 63)


class Key128(ctypes.Structure):
    _fields_ = [("lo", u64), ("hi", u64)]


class Config(ctypes.Structure):
    _fields_ = [("arena_bytes", u64), ("slider", ctypes.c_int), ("block_size", ctypes.c_int),
                ("blocks_per_cu", ctypes.c_int), ("timing", ctypes.c_int), ("leaf_mode", ctypes.c_int)]


class BatchStats(ctypes.Structure):
    _fields_ = [("total_ms", ctypes.c_double), ("graph_ms", ctypes.c_double), ("launches", u64), ("splits", u64),
                ("leaf_items", u64), ("graphs_cached", u64), ("peak_bytes", u64)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class Stats(ctypes.Structure):
    _fields_ = [("total_ms", ctypes.c_double), ("leaf_kernel_ms", ctypes.c_double),
                ("leaf_launches", u64), ("leaf_items", u64), ("leaf_parents", u64), ("leaves", u64),
                ("splits", u64), ("peak_arena", u64), ("nodes", u64 * 16)]

    def as_dict(self):
        d = {n: getattr(self, n) for n, _ in self._fields_ if n != "nodes"}
        d["nodes"] = list(self.nodes)
        return d


class DeepStats(ctypes.Structure):
    _fields_ = [("total_ms", ctypes.c_double), ("device_ms", ctypes.c_double), ("device_calls", u64),
                ("host_nodes", u64), ("host_tt_hits", u64), ("disk_hits", u64), ("disk_stores", u64)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


DIVIDE_CB = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_char_p, u64, u64)


class PerftError(RuntimeError):
    def __init__(self, code, what):
        msg = "%s failed: %s" % (what, _ERR_NAMES.get(code, code))
        detail = last_error()
        if detail:
            msg += " (%s)" % detail
        super().__init__(msg)
        self.code = code


_lib = None


def lib():
    """Load the in-tree library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libperft_gpu_amd.so is not built (run __graft_entry__.build())")
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.POINTER
    sig = {
        "pg_fen_to_pos": (ctypes.c_int, [ctypes.c_char_p, P(Pos48)]),
        "pg_pos_to_fen": (ctypes.c_int, [P(Pos48), ctypes.c_char_p, ctypes.c_int]),
        "pg_zobrist128": (ctypes.c_int, [P(Pos48), P(Key128)]),
        "pg_host_generate_moves": (ctypes.c_int, [P(Pos48), P(ctypes.c_uint16), P(ctypes.c_uint32)]),
        "pg_host_make_move": (ctypes.c_int, [P(Pos48), ctypes.c_uint16, P(Key128)]),
        "pg_host_children": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, u64, ctypes.c_void_p]),
        "pg_host_count_moves": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int, ctypes.c_void_p]),
        "pg_cpu_perft": (ctypes.c_int, [P(Pos48), ctypes.c_uint32, P(u64)]),
        "pg_host_slider_attacks": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                                  ctypes.c_uint32, ctypes.c_void_p]),
        "pg_random_positions": (ctypes.c_int, [u64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                               ctypes.c_void_p, ctypes.c_void_p]),
        "pg_init": (ctypes.c_int, [ctypes.c_int, P(Config), P(ctypes.c_void_p)]),
        "pg_shutdown": (None, [ctypes.c_void_p]),
        "pg_get_config": (ctypes.c_int, [ctypes.c_void_p, P(Config)]),
        "pg_set_slider": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
        "pg_perft": (ctypes.c_int, [ctypes.c_void_p, P(Pos48), ctypes.c_uint32, ctypes.c_int, P(u64)]),
        "pg_perft_noalloc": (ctypes.c_int, [ctypes.c_void_p, P(Pos48), ctypes.c_uint32, ctypes.c_int, P(u64)]),
        "pg_last_stats": (ctypes.c_int, [ctypes.c_void_p, P(Stats)]),
        "pg_test_count_bias": (ctypes.c_int, [ctypes.c_void_p, u64]),
        "pg_last_error": (ctypes.c_char_p, []),
        "pg_device_count_moves": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int,
                                                 ctypes.c_void_p]),
        "pg_device_generate_moves": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int,
                                                    ctypes.c_void_p, ctypes.c_void_p]),
        "pg_device_make_move": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32,
                                               ctypes.c_void_p, ctypes.c_void_p]),
        "pg_device_slider_attacks": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                                    ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p]),
        "pg_device_copy": (ctypes.c_int, [ctypes.c_void_p, u64, ctypes.c_uint32, P(ctypes.c_double)]),
        "pg_device_memory": (ctypes.c_int, [ctypes.c_void_p, P(u64), P(u64)]),
        "pg_device_scan": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, u64, ctypes.c_uint32, ctypes.c_void_p]),
        "pg_device_issue_probe": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32, P(ctypes.c_double),
                                                 P(u64), P(ctypes.c_uint32), P(ctypes.c_uint32)]),
        "pg_tt_create": (ctypes.c_int, [ctypes.c_void_p, u64, P(ctypes.c_void_p)]),
        "pg_tt_clear": (ctypes.c_int, [ctypes.c_void_p]),
        "pg_tt_destroy": (ctypes.c_int, [ctypes.c_void_p]),
        "pg_tt_info": (ctypes.c_int, [ctypes.c_void_p, P(u64), ctypes.c_void_p]),
        "pg_tt_counters": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
        "pg_tt_poison": (ctypes.c_int, [ctypes.c_void_p, u64]),
        "pg_tt_save": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p]),
        "pg_tt_load": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p]),
        "pg_perft_tt": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, P(Pos48), ctypes.c_uint32, ctypes.c_int,
                                       P(u64)]),
        "pg_set_dedup": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
        "pg_set_batch_grid": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
        "pg_perft_deep": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, P(Pos48), ctypes.c_uint32, ctypes.c_uint32,
                                         ctypes.c_uint32, ctypes.c_int, DIVIDE_CB, ctypes.c_void_p, P(u64),
                                         P(DeepStats)]),
        "pg_perft_deep_disk": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, P(Pos48), ctypes.c_uint32,
                                              ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, DIVIDE_CB, ctypes.c_void_p,
                                              ctypes.c_char_p, ctypes.c_uint32, P(u64), P(DeepStats)]),
        "pg_perft_records": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_char_p, ctypes.c_uint32,
                                            ctypes.c_int, ctypes.c_uint32, P(ctypes.c_uint32), P(ctypes.c_uint32)]),
        "pg_perft_batch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32,
                                          ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                          P(BatchStats)]),
        "pg_perft_batch_mixed": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32,
                                                ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                                ctypes.c_void_p, P(BatchStats)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def last_error():
    if _lib is None:
        return ""
    s = _lib.pg_last_error()
    return s.decode() if s else ""


def _check(rc, what):
    if rc != PG_OK:
        raise PerftError(rc, what)


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


# ------------------------------------------------------------------ host side
def read_fen_string(fen):
    """FEN -> Pos48 (Utils::readFENString + board088ToHexBB)."""
    p = Pos48()
    _check(lib().pg_fen_to_pos(fen.encode(), ctypes.byref(p)), "pg_fen_to_pos")
    return p


def pos_to_fen(pos):
    buf = ctypes.create_string_buffer(128)
    _check(lib().pg_pos_to_fen(ctypes.byref(pos), buf, 128), "pg_pos_to_fen")
    return buf.value.decode()


def zobrist128(pos):
    k = Key128()
    _check(lib().pg_zobrist128(ctypes.byref(pos), ctypes.byref(k)), "pg_zobrist128")
    return int(k.lo), int(k.hi)


def generate_moves_host(pos):
    buf = (ctypes.c_uint16 * 256)()
    n = ctypes.c_uint32()
    _check(lib().pg_host_generate_moves(ctypes.byref(pos), buf, ctypes.byref(n)), "pg_host_generate_moves")
    return list(buf[:n.value])


def count_moves_host(positions, slider):
    arr = positions_array(positions)
    out = np.zeros(len(arr), dtype=np.uint32)
    _check(lib().pg_host_count_moves(_ptr(arr), len(arr), slider, _ptr(out)), "pg_host_count_moves")
    return out


def cpu_perft(pos, depth):
    """perft_bb (perft_bb.h:79-101) on one host thread (pg_cpu_perft): the
    reference's CPU path as an entry point of its own."""
    if isinstance(pos, str):
        pos = read_fen_string(pos)
    out = u64()
    _check(lib().pg_cpu_perft(ctypes.byref(pos), depth, ctypes.byref(out)), "pg_cpu_perft")
    return int(out.value)


def children_host(positions, plies=1):
    """generateBoards on the host, `plies` times: (children (m,6) array, owner
    index of each child in the input)."""
    arr = positions_array(positions)
    owner = np.arange(len(arr), dtype=np.int64)
    for _ in range(plies):
        cap = max(1, len(arr) * 218)
        out = np.zeros((cap, 6), dtype=np.uint64)
        cnt = np.zeros(len(arr), dtype=np.uint32)
        _check(lib().pg_host_children(_ptr(arr), len(arr), _ptr(out), cap, _ptr(cnt)), "pg_host_children")
        owner = np.repeat(owner, cnt.astype(np.int64))
        arr = np.ascontiguousarray(out[:int(cnt.sum())])
    return arr, owner


def make_move_host(pos, move, key=None):
    """Returns (child, child_key or None)."""
    child = Pos48(*pos.words())
    k = None
    if key is not None:
        k = Key128(key[0], key[1])
    _check(lib().pg_host_make_move(ctypes.byref(child), move, ctypes.byref(k) if k else None), "pg_host_make_move")
    return child, ((int(k.lo), int(k.hi)) if k else None)


def slider_attacks_host(slider, rook, squares, occupancy):
    sq = np.ascontiguousarray(squares, dtype=np.uint8)
    occ = np.ascontiguousarray(occupancy, dtype=np.uint64)
    out = np.zeros(len(sq), dtype=np.uint64)
    _check(lib().pg_host_slider_attacks(slider, int(rook), _ptr(sq), _ptr(occ), len(sq), _ptr(out)),
           "pg_host_slider_attacks")
    return out


def random_positions(seed, n, min_ply, max_ply):
    """Seeded workload generator; returns (positions as (n,6) uint64 array, plies)."""
    arr = np.zeros((n, 6), dtype=np.uint64)
    plies = np.zeros(n, dtype=np.uint32)
    _check(lib().pg_random_positions(seed, n, min_ply, max_ply, _ptr(arr), _ptr(plies)), "pg_random_positions")
    return arr, plies


def positions_array(positions):
    """list of Pos48 / FEN strings / (n,6) array -> contiguous (n,6) uint64 array."""
    if isinstance(positions, np.ndarray):
        return np.ascontiguousarray(positions, dtype=np.uint64).reshape(-1, 6)
    out = np.zeros((len(positions), 6), dtype=np.uint64)
    for i, p in enumerate(positions):
        if isinstance(p, str):
            p = read_fen_string(p)
        out[i] = p.words()
    return out


def pos_from_words(w):
    return Pos48(*[int(x) for x in w])


# ---------------------------------------------------------------- device side
class Device:
    """One GPU context (pg_ctx): tables uploaded, tree arena reserved."""

    def __init__(self, device=0, arena_bytes=0, slider=SLIDER_AUTO, timing=False, blocks_per_cu=0, leaf_mode=0):
        L = lib()
        cfg = Config(arena_bytes, slider, 0, blocks_per_cu, 1 if timing else 0, leaf_mode)
        h = ctypes.c_void_p()
        _check(L.pg_init(device, ctypes.byref(cfg), ctypes.byref(h)), "pg_init")
        self._h = h
        self.device = device
        self._children = []   # tables, slot servers and serving loops made on this context

    def _adopt(self, obj):
        # the context keeps what was made on it until it closes it
        self._children.append(obj)

    def close(self):
        if self._h:
            # serving loops, then slot servers, then tables, then the context
            for rank in (0, 1, 2):
                for obj in self._children:
                    if getattr(obj, "_close_rank", 2) == rank:
                        obj.close()
            self._children = []
            lib().pg_shutdown(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def config(self):
        c = Config()
        _check(lib().pg_get_config(self._h, ctypes.byref(c)), "pg_get_config")
        return {n: getattr(c, n) for n, _ in c._fields_}

    def set_dedup(self, on):
        _check(lib().pg_set_dedup(self._h, 1 if on else 0), "pg_set_dedup")

    def set_count_bias(self, bias):
        """Test hook pg_test_count_bias: later perft totals start at `bias`."""
        _check(lib().pg_test_count_bias(self._h, bias), "pg_test_count_bias")

    def set_batch_grid(self, max_groups):
        """pg_set_batch_grid: -1 sized to each batch, 0 whole chip, > 0 at most that many work-groups."""
        _check(lib().pg_set_batch_grid(self._h, max_groups), "pg_set_batch_grid")

    def set_slider(self, slider):
        _check(lib().pg_set_slider(self._h, slider), "pg_set_slider")

    def perft(self, pos, depth, slider=SLIDER_AUTO):
        if isinstance(pos, str):
            pos = read_fen_string(pos)
        out = u64()
        _check(lib().pg_perft(self._h, ctypes.byref(pos), depth, slider, ctypes.byref(out)), "pg_perft")
        return int(out.value)

    def last_stats(self):
        s = Stats()
        _check(lib().pg_last_stats(self._h, ctypes.byref(s)), "pg_last_stats")
        return s.as_dict()

    def perft_batch(self, positions, depth, slider=SLIDER_AUTO, dedup=True, blocks_per_cu=0, tt=None):
        """perft(depth) of every position in one graph launch; returns (counts, stats)."""
        arr = positions_array(positions)
        out = np.zeros(len(arr), dtype=np.uint64)
        st = BatchStats()
        _check(lib().pg_perft_batch(self._h, _ptr(arr), len(arr), depth, slider, 1 if dedup else 0, blocks_per_cu,
                                    tt._h if tt is not None else None, _ptr(out), ctypes.byref(st)), "pg_perft_batch")
        return out, st.as_dict()

    def perft_batch_mixed(self, positions, depths, slider=SLIDER_AUTO, dedup=True, blocks_per_cu=0, tt=None):
        """perft(depths[i]) of every position i, one graph launch per depth
        present (pg_perft_batch_mixed); returns (counts, stats)."""
        arr = positions_array(positions)
        d = np.ascontiguousarray(depths, dtype=np.uint32)
        if len(d) != len(arr):
            raise ValueError("one depth per position")
        out = np.zeros(len(arr), dtype=np.uint64)
        st = BatchStats()
        _check(lib().pg_perft_batch_mixed(self._h, _ptr(arr), _ptr(d), len(arr), slider, 1 if dedup else 0,
                                          blocks_per_cu, tt._h if tt is not None else None, _ptr(out),
                                          ctypes.byref(st)), "pg_perft_batch_mixed")
        return out, st.as_dict()

    def count_moves(self, positions, slider=SLIDER_AUTO):
        arr = positions_array(positions)
        out = np.zeros(len(arr), dtype=np.uint32)
        _check(lib().pg_device_count_moves(self._h, _ptr(arr), len(arr), slider, _ptr(out)), "pg_device_count_moves")
        return out

    def generate_moves(self, positions, slider=SLIDER_AUTO):
        arr = positions_array(positions)
        moves = np.zeros((len(arr), 256), dtype=np.uint16)
        cnt = np.zeros(len(arr), dtype=np.uint32)
        _check(lib().pg_device_generate_moves(self._h, _ptr(arr), len(arr), slider, _ptr(moves), _ptr(cnt)),
               "pg_device_generate_moves")
        return [list(moves[i, :cnt[i]]) for i in range(len(arr))]

    def make_moves(self, positions, moves, keys=None):
        arr = positions_array(positions)
        mv = np.ascontiguousarray(moves, dtype=np.uint16)
        out = np.zeros_like(arr)
        k = None if keys is None else np.ascontiguousarray(keys, dtype=np.uint64).reshape(-1, 2).copy()
        _check(lib().pg_device_make_move(self._h, _ptr(arr), _ptr(mv), len(arr), _ptr(out),
                                         _ptr(k) if k is not None else None), "pg_device_make_move")
        return out, k

    def slider_attacks(self, slider, rook, squares, occupancy):
        sq = np.ascontiguousarray(squares, dtype=np.uint8)
        occ = np.ascontiguousarray(occupancy, dtype=np.uint64)
        out = np.zeros(len(sq), dtype=np.uint64)
        _check(lib().pg_device_slider_attacks(self._h, slider, int(rook), _ptr(sq), _ptr(occ), len(sq), _ptr(out)),
               "pg_device_slider_attacks")
        return out


    def scan(self, counts, misalign=0):
        """The level scan on the device: exclusive prefix sums of u32 counts
        (offsets[n] = total), with the counts `misalign` elements past a
        16-byte boundary."""
        c = np.ascontiguousarray(counts, dtype=np.uint32)
        off = np.zeros(len(c) + 1, dtype=np.uint64)
        _check(lib().pg_device_scan(self._h, _ptr(c), len(c), misalign, _ptr(off)), "pg_device_scan")
        return off

    def memory(self):
        """(free, total) bytes of the device's HBM (hipMemGetInfo)."""
        free, total = u64(), u64()
        _check(lib().pg_device_memory(self._h, ctypes.byref(free), ctypes.byref(total)), "pg_device_memory")
        return int(free.value), int(total.value)

    def copy_bandwidth(self, nbytes, iters=10):
        """Mean time (ms) and GB/s (read + write) of a plain nbytes copy."""
        ms = ctypes.c_double()
        _check(lib().pg_device_copy(self._h, nbytes, iters, ctypes.byref(ms)), "pg_device_copy")
        return ms.value, 2 * nbytes / (ms.value * 1e-3) / 1e9

    ISSUE_KINDS = {0: "v_bitop3_b32", 1: "v_xor_b32 (two-operand)", 2: "v_lshlrev_b64", 3: "v_bcnt_u32_b32",
                   4: "v_and_b32", 5: "v_or3_b32", 6: "v_lshlrev_b32", 7: "v_alignbit_b32", 8: "v_add_u32",
                   9: "v_lshl_add_u64", 10: "v_bfrev_b32", 11: "v_ffbl_b32", 12: "v_mov_b32", 13: "v_lshl_or_b32",
                   14: "v_and_or_b32", 15: "v_not_b32"}

    def issue_rate(self, kind, iters=4096):
        """VALU issue probe: (wave instructions per CU per clock, ms, instructions, clock kHz, CUs)."""
        ms, n, clk, cus = ctypes.c_double(), u64(), ctypes.c_uint32(), ctypes.c_uint32()
        _check(lib().pg_device_issue_probe(self._h, kind, iters, ctypes.byref(ms), ctypes.byref(n), ctypes.byref(clk),
                                           ctypes.byref(cus)), "pg_device_issue_probe")
        rate = n.value / (cus.value * ms.value * 1e-3 * clk.value * 1e3)
        return rate, ms.value, int(n.value), int(clk.value), int(cus.value)


class TranspositionTable:
    """Device TT (pg_tt): setupHashTables128b (launcher.h:293), zeroed."""

    _close_rank = 2

    def __init__(self, dev, nbytes):
        h = ctypes.c_void_p()
        _check(lib().pg_tt_create(dev._h, nbytes, ctypes.byref(h)), "pg_tt_create")
        self._h = h
        self.dev = dev
        dev._adopt(self)

    def close(self):
        if self._h:
            _check(lib().pg_tt_destroy(self._h), "pg_tt_destroy")
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def clear(self):
        _check(lib().pg_tt_clear(self._h), "pg_tt_clear")

    def poison(self, seed=1):
        _check(lib().pg_tt_poison(self._h, seed), "pg_tt_poison")

    def save(self, path):
        """pg_tt_save: write the table's entries (with layout and checksum) to path."""
        _check(lib().pg_tt_save(self._h, os.fsencode(path)), "pg_tt_save")

    def load(self, path):
        """pg_tt_load: restore entries saved from a table of the same size and layout."""
        _check(lib().pg_tt_load(self._h, os.fsencode(path)), "pg_tt_load")

    def info(self):
        b = u64()
        bits = np.zeros(16, dtype=np.int32)
        _check(lib().pg_tt_info(self._h, ctypes.byref(b), _ptr(bits)), "pg_tt_info")
        return int(b.value), [int(x) for x in bits]

    def counters(self):
        """per remaining depth: (probes, hits, stores, duplicates)"""
        c = np.zeros((16, 4), dtype=np.uint64)
        _check(lib().pg_tt_counters(self._h, _ptr(c)), "pg_tt_counters")
        return c

    def perft(self, pos, depth, slider=SLIDER_AUTO):
        if isinstance(pos, str):
            pos = read_fen_string(pos)
        out = u64()
        _check(lib().pg_perft_tt(self.dev._h, self._h, ctypes.byref(pos), depth, slider, ctypes.byref(out)),
               "pg_perft_tt")
        return int(out.value)


def divide(dev, pos, depth, slider=SLIDER_AUTO, tt=None):
    """Divided perft (launcher.h:1344-1351, 1392-1450): perft(depth-1) of every
    root move, all children in one batched GPU search.  Returns {move: count}."""
    if isinstance(pos, str):
        pos = read_fen_string(pos)
    moves = generate_moves_host(pos)
    if depth <= 1:
        return {m: 1 for m in moves}
    kids = np.array([make_move_host(pos, m)[0].words() for m in moves], dtype=np.uint64).reshape(-1, 6)
    try:
        counts, _ = dev.perft_batch(kids, depth - 1, slider, True, 0, tt)
    except PerftError as e:
        if e.code != PG_ENOMEM:
            raise
        # one child's levels do not fit a batch region: one perft per move,
        # whose host-driven level loop splits levels to fit
        counts = [(tt.perft(pos_from_words(k), depth - 1, slider) if tt is not None
                   else dev.perft(pos_from_words(k), depth - 1, slider)) for k in kids]
    return {m: int(c) for m, c in zip(moves, counts)}


# Tree arena of a context that runs table searches.  A deep table search
# expands levels of billions of boards (start-position perft(10): 2.6e9 at
# ply 8, about 43 B each with key, perft, owner, state, move, parent and
# collapse-set slots); a level that does not fit is split, and equal boards
# in different parts are not collapsed, so a larger arena means less leaf
# work: start perft(10) with a 16 GB table takes 0.95 s on 48 GiB (6 splits)
# and 0.84 s on 160 GiB (none; its deepest level peaks at 159 GB).
DEFAULT_ARENA = 48 << 30     # the library's default (pg_config.arena_bytes = 0)
TABLE_ARENA_MAX = 160 << 30
TABLE_ARENA_RESERVE = 16 << 30


def table_search_arena(table_bytes, device=0):
    """Arena bytes for a context that will hold a table of table_bytes: the
    HBM free now less the table and TABLE_ARENA_RESERVE, at most
    TABLE_ARENA_MAX; 0 (the library default) when that is not larger."""
    with Device(device, arena_bytes=64 << 20) as probe:
        free, _ = probe.memory()
    a = min(TABLE_ARENA_MAX, free - int(table_bytes) - TABLE_ARENA_RESERVE)
    return a if a > DEFAULT_ARENA else 0


def perft_deep(dev, pos, depth, launch_depth=10, divided_depth=10, tt=None, slider=SLIDER_AUTO, on_divide=None,
               hash_path=None, disk_depth=0):
    """perft_bb_cpu_launcher + dividedPerft (launcher.h:1181-1450): perft of any
    depth with a 128-bit total.  Nodes deeper than launch_depth are walked on
    the host (with a host table of their totals), each node at launch_depth is
    one device call (TT perft if tt is given).  on_divide(path, count) is
    called for every move of a host node deeper than divided_depth, in the
    reference's print order.  hash_path: a perfthash.dat disk table for the
    nodes of remaining depth disk_depth (0: depth - 2), read before and
    written after each (pg_perft_deep_disk).  Returns (total, stats)."""
    if isinstance(pos, str):
        pos = read_fen_string(pos)
    out = (u64 * 2)()
    st = DeepStats()
    cb = DIVIDE_CB(0)
    if on_divide is not None:
        def _cb(_user, path, lo, hi):
            on_divide(path.decode(), int(lo) | (int(hi) << 64))
        cb = DIVIDE_CB(_cb)
    _check(lib().pg_perft_deep_disk(dev._h, tt._h if tt is not None else None, ctypes.byref(pos), depth, launch_depth,
                                    divided_depth, slider, cb, None,
                                    os.fsencode(hash_path) if hash_path is not None else None, disk_depth, out,
                                    ctypes.byref(st)), "pg_perft_deep_disk")
    return int(out[0]) | (int(out[1]) << 64), st.as_dict()


def perft_records(dev, path, depth, tt=None, slider=SLIDER_AUTO, chunk=1024):
    """processPerftRecords (launcher.h:1578-1695): perft(depth) of every FEN
    line of `path` (last number = occurrence count) into path + ".op", resuming
    after the lines the .op file already holds.  Returns (done, skipped)."""
    done, skipped = ctypes.c_uint32(), ctypes.c_uint32()
    _check(lib().pg_perft_records(dev._h, tt._h if tt is not None else None, os.fsencode(path), depth, slider, chunk,
                                  ctypes.byref(done), ctypes.byref(skipped)), "pg_perft_records")
    return int(done.value), int(skipped.value)


def move_to_str(m):
    """CMove -> long algebraic (e2e4, e7e8q)."""
    fr, to, fl = m & 63, (m >> 6) & 63, m >> 12
    s = "%s%d%s%d" % ("abcdefgh"[fr & 7], (fr >> 3) + 1, "abcdefgh"[to & 7], (to >> 3) + 1)
    if fl & 8:
        s += "nbrq"[fl & 3]
    return s


def init_gpu(device=0, **kw):
    """initGPU (launcher.h:1506): select the GPU, upload tables, reserve the arena."""
    return Device(device, **kw)


def perft_launcher(dev, pos, depth, launch_depth=None, slider=SLIDER_AUTO, verbose=True):
    """perftLauncher, no-TT branch (launcher.h:1455-1503): run the GPU perft and
    print the reference's result line.  launch_depth is accepted for call
    compatibility; the level loop sizes its own splits from the arena."""
    t0 = time.perf_counter()
    res = dev.perft(pos, depth, slider)
    dt = time.perf_counter() - t0
    if verbose:
        print("\nGPU Perft %d: %d,   Time taken: %g seconds, nps: %d" % (depth, res, dt, int(res / dt) if dt > 0 else 0))
    return res
