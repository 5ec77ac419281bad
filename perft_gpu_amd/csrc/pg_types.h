// pg_types.h -- data model shared by host code and the gfx950 kernels.
//
// The position is the reference's 48-byte, six-bitboard HexaBitBoardPosition
// (chess.h:208-230), kept byte-identical so boards cross the C ABI unchanged:
//   word 0 whitePieces   word 1 pawns (+ game state)   word 2 knights
//   word 3 bishopQueens  word 4 rookQueens             word 5 kings
// Queens sit in both slider words.  Game state lives in the pawn word's
// rank-1 and rank-8 bytes, which can never hold a pawn:
//   bits 0-1 white castle (1 = king side, 2 = queen side), bits 2-3 black,
//   bits 4-7 en-passant file + 1, bits 56-62 half-move counter, bit 63 side.
// Moves are the reference's 16-bit CMove (chess.h:234-293):
//   from | to << 6 | flags << 12.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define PG_HD __host__ __device__ __attribute__((always_inline)) inline
#else
#define PG_HD inline
#endif

namespace pg {

typedef unsigned long long u64;
typedef unsigned int u32;
typedef unsigned short u16;
typedef unsigned char u8;

struct Pos {
    u64 white, pawns, knights, bishopQueens, rookQueens, kings;
};
static_assert(sizeof(Pos) == 48, "position must be 48 bytes");

struct Key128 {
    u64 lo, hi;
};
static_assert(sizeof(Key128) == 16, "key must be 16 bytes");

enum { WHITE = 0, BLACK = 1 };

// CMove flag values (chess.h:269-293)
enum {
    MF_QUIET = 0, MF_DOUBLE_PUSH = 1, MF_KING_CASTLE = 2, MF_QUEEN_CASTLE = 3,
    MF_CAPTURE = 4, MF_EP_CAPTURE = 5, MF_PROMO = 8,   // 8..11: N, B, R, Q (+4: capture)
};

constexpr u64 RANKS2TO7 = 0x00FFFFFFFFFFFF00ULL;
constexpr u64 RANK1 = 0x00000000000000FFULL;
constexpr u64 RANK3 = 0x0000000000FF0000ULL;
constexpr u64 RANK6 = 0x0000FF0000000000ULL;
constexpr u64 RANK8 = 0xFF00000000000000ULL;
constexpr u64 FILEA = 0x0101010101010101ULL;
constexpr u64 FILEB = 0x0202020202020202ULL;
constexpr u64 FILEG = 0x4040404040404040ULL;
constexpr u64 FILEH = 0x8080808080808080ULL;
constexpr u64 DIAG_A1H8 = 0x8040201008040201ULL;
constexpr u64 DIAG_A8H1 = 0x0102040810204080ULL;

PG_HD int pos_side(const Pos &p) { return (int)(p.pawns >> 63); }
PG_HD int pos_wcastle(const Pos &p) { return (int)(p.pawns & 3); }
PG_HD int pos_bcastle(const Pos &p) { return (int)((p.pawns >> 2) & 3); }
PG_HD int pos_ep(const Pos &p) { return (int)((p.pawns >> 4) & 15); }
PG_HD u64 pos_pawns(const Pos &p) { return p.pawns & RANKS2TO7; }
PG_HD u64 pos_occ(const Pos &p) { return p.kings | (p.pawns & RANKS2TO7) | p.knights | p.bishopQueens | p.rookQueens; }

// A board whose white word is zero is a marker (no king can be absent from a
// real position): used for transposition hits and removed duplicates.
PG_HD bool pos_valid(const Pos &p) { return p.white != 0; }

}  // namespace pg
