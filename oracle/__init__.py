"""ctypes binding of the oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this module.  It wraps
  liboracle.so           the C restatement of the reference's CPU perft
                         (perft_oracle.c, each function cites the reference
                         file:line it follows)
  _ref/libperft_ref.so   the reference's own CPU move generator compiled from
                         its sources (present where it was built; it travels
                         with the tree as a built library)
Positions are 6-word tuples in HexaBitBoardPosition order (chess.h:208-230).
"""
import ctypes
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libperft_ref.so")

P6 = ctypes.c_uint64 * 6
K2 = ctypes.c_uint64 * 2


def build(force=False):
    args = ["make", "-C", HERE] + (["-B"] if force else [])
    subprocess.run(args, check=True, stdout=subprocess.DEVNULL)


class _Lib:
    def __init__(self, path, prefix):
        self.L = ctypes.CDLL(path)
        self.p = prefix
        c = ctypes
        sig = {
            "init": (None, []),
            "perft": (c.c_ulonglong, [c.c_void_p, c.c_uint]),
            "perft_mt": (c.c_ulonglong, [c.c_void_p, c.c_uint, c.c_int]),
            "count_moves": (c.c_uint, [c.c_void_p]),
            "generate_boards": (c.c_uint, [c.c_void_p, c.c_void_p]),
            "generate_moves": (c.c_uint, [c.c_void_p, c.c_void_p]),
            "make_move": (None, [c.c_void_p, c.c_void_p, c.c_ushort]),
            "zobrist128": (None, [c.c_void_p, c.c_void_p]),
            "fen": (c.c_int, [c.c_char_p, c.c_void_p]),
        }
        for name, (res, args) in sig.items():
            f = getattr(self.L, prefix + name)
            f.restype = res
            f.argtypes = args
            setattr(self, "_" + name, f)
        self._init()

    def fen(self, fen):
        p = P6()
        rc = self._fen(fen.encode(), p)
        if rc != 0:
            raise ValueError("bad FEN: " + fen)
        return tuple(p)

    def perft(self, pos, depth):
        return int(self._perft(P6(*pos), depth))

    def perft_mt(self, pos, depth, threads):
        return int(self._perft_mt(P6(*pos), depth, threads))

    def count_moves(self, pos):
        return int(self._count_moves(P6(*pos)))

    def generate_boards(self, pos):
        out = (ctypes.c_uint64 * (6 * 256))()
        n = self._generate_boards(P6(*pos), out)
        return [tuple(out[6 * i:6 * i + 6]) for i in range(n)]

    def generate_moves(self, pos):
        out = (ctypes.c_uint16 * 256)()
        n = self._generate_moves(P6(*pos), out)
        return list(out[:n])

    def make_move(self, pos, move, key=(0, 0)):
        p = P6(*pos)
        k = K2(*key)
        self._make_move(p, k, move)
        return tuple(p), (int(k[0]), int(k[1]))

    def zobrist128(self, pos):
        k = K2()
        self._zobrist128(P6(*pos), k)
        return int(k[0]), int(k[1])


class _Oracle(_Lib):
    def __init__(self):
        super().__init__(ORACLE_SO, "or_")
        f = self.L.or_divide
        f.restype = ctypes.c_uint
        f.argtypes = [ctypes.c_void_p, ctypes.c_uint, ctypes.c_void_p, ctypes.c_void_p]

    def divide(self, pos, depth):
        mv = (ctypes.c_uint16 * 256)()
        cnt = (ctypes.c_uint64 * 256)()
        n = self.L.or_divide(P6(*pos), depth, mv, cnt)
        return {int(mv[i]): int(cnt[i]) for i in range(n)}


class _Ref(_Lib):
    def __init__(self):
        super().__init__(REF_SO, "ref_")
        # fancy/byte table accessors
        for name, res, args in [("ref_fancy_table", ctypes.c_ulonglong, [ctypes.c_int]),
                                ("ref_byte_table", ctypes.c_uint, [ctypes.c_int]),
                                ("ref_byte_rook", ctypes.c_ulonglong, [ctypes.c_int]),
                                ("ref_byte_bishop", ctypes.c_ulonglong, [ctypes.c_int]),
                                ("ref_fancy_offset", ctypes.c_int, [ctypes.c_int, ctypes.c_int])]:
            f = getattr(self.L, name)
            f.restype = res
            f.argtypes = args


_oracle = None
_ref = None


def oracle():
    """The C restatement (builds it if absent)."""
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_SO):
            build()
        _oracle = _Oracle()
    return _oracle


def ref_available():
    return os.path.exists(REF_SO)


def ref():
    """The reference's own CPU code, or None where it was not built."""
    global _ref
    if _ref is None and ref_available():
        _ref = _Ref()
    return _ref
