/*
 * perft_oracle.c -- CPU restatement of the reference's CPU perft path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/, the
 * smoke() entry point and bench.py's cpu_baseline leg may load the library
 * built from it.  The product (perft_gpu_amd) never links or calls it.
 *
 * It restates, function by function, the host path of ankan-ban/perft_gpu
 * that the reference's own CPU perft takes:
 *   perft_bb                  perft_bb.h:79-101   (countMoves at depth 1,
 *                                                  generateBoards above)
 *   countMoves(+OutOfCheck)   MoveGeneratorBitboard.h:2825-3295
 *   generateBoards(+OOC)      MoveGeneratorBitboard.h:1741-2261, helpers
 *                             :1302-1738 (addSlidingMove, addKnightMove,
 *                             addKingMove, addCastleMove, addSinglePawnMove,
 *                             addEnPassentMove, addPawnMoves)
 *   generateMoves(+OOC)       MoveGeneratorBitboard.h:2266-2823
 *   makeMove (128-bit hash)   MoveGeneratorBitboard.h:3575-3836, with
 *                             EXACT_EN_PASSENT_FLAGGING (switches.h:48)
 *   computeZobristKey128b     MoveGeneratorBitboard.h:3914-3980
 *   findPinnedPieces          MoveGeneratorBitboard.h:1213-1247
 *   findAttackedSquares       MoveGeneratorBitboard.h:1254-1298
 *   Kogge-Stone fills/attacks MoveGeneratorBitboard.h:454-728
 *   squaresInBetween/InLine   MoveGeneratorBitboard.h:905-961
 *   init tables               MoveGeneratorBitboard.h:982-1021
 *   readFENString             util.cpp:189-281 + board088ToHexBB util.cpp:92-140
 *
 * Sliders use the reference's Kogge-Stone variant (USE_SLIDING_LUT 0); the
 * three magic variants are lookups of the same function and are pinned
 * against it separately.  Zobrist key sets come from the engine's data file,
 * which is pinned against the reference's own keys by golden vectors.
 */
#include <stdint.h>
#include <string.h>
#include <stdlib.h>
#include <pthread.h>

typedef unsigned long long u64;
typedef uint32_t u32;
typedef uint16_t u16;
typedef uint8_t u8;

#include "../perft_gpu_amd/csrc/zobrist_keys.inc"

typedef struct { u64 whitePieces, pawns, knights, bishopQueens, rookQueens, kings; } OPos;
typedef struct { u64 lo, hi; } OKey;

#define BIT(i) (1ULL << (i))
#define RANKS2TO7 0x00FFFFFFFFFFFF00ULL
#define RANK1 0x00000000000000FFULL
#define RANK3 0x0000000000FF0000ULL
#define RANK8 0xFF00000000000000ULL
#define FILEA 0x0101010101010101ULL
#define FILEH 0x8080808080808080ULL
#define DIAG_A1H8 0x8040201008040201ULL
#define DIAG_A8H1 0x0102040810204080ULL
#define CENTRAL 0x007E7E7E7E7E7E00ULL
#define F1G1 0x60ULL
#define C1D1 0x0CULL
#define B1D1 0x0EULL
#define F8G8 0x6000000000000000ULL
#define C8D8 0x0C00000000000000ULL
#define B8D8 0x0E00000000000000ULL
#define WKR 0x0000000000000080ULL
#define WQR 0x0000000000000001ULL
#define BKR 0x8000000000000000ULL
#define BQR 0x0100000000000000ULL
enum { WHITE = 0, BLACK = 1 };
enum { PAWN = 1, KNIGHT, BISHOP, ROOK, QUEEN, KING };
enum { A1 = 0, B1, C1, D1, E1, F1, G1, H1, A8 = 56, B8, C8, D8, E8, F8, G8, H8 };

/* game-state fields packed into the pawns word (chess.h:212-226) */
static inline int wc(const OPos *p) { return (int)(p->pawns & 3); }
static inline int bc(const OPos *p) { return (int)((p->pawns >> 2) & 3); }
static inline int ep(const OPos *p) { return (int)((p->pawns >> 4) & 15); }
static inline int side(const OPos *p) { return (int)(p->pawns >> 63); }
static inline void set_wc(OPos *p, int v) { p->pawns = (p->pawns & ~3ULL) | (u64)(v & 3); }
static inline void set_bc(OPos *p, int v) { p->pawns = (p->pawns & ~12ULL) | ((u64)(v & 3) << 2); }
static inline void set_ep(OPos *p, int v) { p->pawns = (p->pawns & ~0xF0ULL) | ((u64)(v & 15) << 4); }
static inline void set_side(OPos *p, int v) { p->pawns = (p->pawns & ~(1ULL << 63)) | ((u64)(v & 1) << 63); }
static inline void set_hmc(OPos *p, int v) { p->pawns = (p->pawns & ~(0x7FULL << 56)) | ((u64)(v & 0x7F) << 56); }

static inline int popc(u64 x) { return __builtin_popcountll(x); }
static inline int bsf(u64 x) { return __builtin_ctzll(x); }
static inline u64 one(u64 x) { return x & (0 - x); }
static inline int multiple(u64 x) { return (x ^ one(x)) != 0; }

/* MoveGeneratorBitboard.h:454-496 */
static inline u64 northOne(u64 x) { return x << 8; }
static inline u64 southOne(u64 x) { return x >> 8; }
static inline u64 eastOne(u64 x) { return (x << 1) & ~FILEA; }
static inline u64 westOne(u64 x) { return (x >> 1) & ~FILEH; }
static inline u64 northEastOne(u64 x) { return (x << 9) & ~FILEA; }
static inline u64 northWestOne(u64 x) { return (x << 7) & ~FILEH; }
static inline u64 southEastOne(u64 x) { return (x >> 7) & ~FILEA; }
static inline u64 southWestOne(u64 x) { return (x >> 9) & ~FILEH; }

/* Kogge-Stone directional attacks, MoveGeneratorBitboard.h:614-712 */
static inline u64 northAttacks(u64 g, u64 p) { g |= (g << 8) & p; p &= p << 8; g |= (g << 16) & p; p &= p << 16; g |= (g << 32) & p; return g << 8; }
static inline u64 southAttacks(u64 g, u64 p) { g |= (g >> 8) & p; p &= p >> 8; g |= (g >> 16) & p; p &= p >> 16; g |= (g >> 32) & p; return g >> 8; }
static inline u64 eastAttacks(u64 g, u64 p) { p &= ~FILEA; g |= (g << 1) & p; p &= p << 1; g |= (g << 2) & p; p &= p << 2; g |= (g << 3) & p; return (g << 1) & ~FILEA; }
static inline u64 westAttacks(u64 g, u64 p) { p &= ~FILEH; g |= (g >> 1) & p; p &= p >> 1; g |= (g >> 2) & p; p &= p >> 2; g |= (g >> 3) & p; return (g >> 1) & ~FILEH; }
static inline u64 northEastAttacks(u64 g, u64 p) { p &= ~FILEA; g |= (g << 9) & p; p &= p << 9; g |= (g << 18) & p; p &= p << 18; g |= (g << 36) & p; return (g << 9) & ~FILEA; }
static inline u64 northWestAttacks(u64 g, u64 p) { p &= ~FILEH; g |= (g << 7) & p; p &= p << 7; g |= (g << 14) & p; p &= p << 14; g |= (g << 28) & p; return (g << 7) & ~FILEH; }
static inline u64 southEastAttacks(u64 g, u64 p) { p &= ~FILEA; g |= (g >> 7) & p; p &= p >> 7; g |= (g >> 14) & p; p &= p >> 14; g |= (g >> 28) & p; return (g >> 7) & ~FILEA; }
static inline u64 southWestAttacks(u64 g, u64 p) { p &= ~FILEH; g |= (g >> 9) & p; p &= p >> 9; g |= (g >> 18) & p; p &= p >> 18; g |= (g >> 36) & p; return (g >> 9) & ~FILEH; }

/* MoveGeneratorBitboard.h:714-728 */
static inline u64 bishopAttacks(u64 b, u64 pro) { return northEastAttacks(b, pro) | northWestAttacks(b, pro) | southEastAttacks(b, pro) | southWestAttacks(b, pro); }
static inline u64 rookAttacks(u64 r, u64 pro) { return northAttacks(r, pro) | southAttacks(r, pro) | eastAttacks(r, pro) | westAttacks(r, pro); }

/* MoveGeneratorBitboard.h:857-878 */
static inline u64 kingAttacks(u64 k) { u64 a = eastOne(k) | westOne(k); k |= a; a |= northOne(k) | southOne(k); return a; }
static inline u64 knightAttacks(u64 n) {
    u64 l1 = (n >> 1) & 0x7f7f7f7f7f7f7f7fULL, l2 = (n >> 2) & 0x3f3f3f3f3f3f3f3fULL;
    u64 r1 = (n << 1) & 0xfefefefefefefefeULL, r2 = (n << 2) & 0xfcfcfcfcfcfcfcfcULL;
    u64 h1 = l1 | r1, h2 = l2 | r2;
    return (h1 << 16) | (h1 >> 16) | (h2 << 8) | (h2 >> 8);
}

static u64 Between[64][64], Line[64][64];
static u64 RookAtt[64], BishopAtt[64], KnightAtt[64], KingAtt[64];
static const u64 *ZLO = PG_ZOBRIST_LO, *ZHI = PG_ZOBRIST_HI;

/* MoveGeneratorBitboard.h:905-925 */
static u64 squaresInBetween(int sq1, int sq2) {
    const u64 m1 = ~0ULL, a2a7 = 0x0001010101010100ULL, b2g7 = 0x0040201008040200ULL, h1b7 = 0x0002040810204080ULL;
    u64 btwn = (m1 << sq1) ^ (m1 << sq2);
    u64 file = (u64)((sq2 & 7) - (sq1 & 7));
    u64 rank = (u64)(((sq2 | 7) - sq1) >> 3);
    u64 line = ((file & 7) - 1) & a2a7;
    line += 2 * (((rank & 7) - 1) >> 58);
    line += (((rank - file) & 15) - 1) & b2g7;
    line += (((rank + file) & 15) - 1) & h1b7;
    line *= btwn & (0 - btwn);
    return line & btwn;
}

/* MoveGeneratorBitboard.h:928-961 */
static u64 squaresInLine(int sq1, int sq2) {
    int fileDiff = (sq2 & 7) - (sq1 & 7);
    int rankDiff = ((sq2 | 7) - sq1) >> 3;
    int file = sq1 & 7, rank = sq1 >> 3;
    if (fileDiff == 0) return FILEA << file;
    if (rankDiff == 0) return RANK1 << (rank * 8);
    if (fileDiff - rankDiff == 0) return (rank - file >= 0) ? DIAG_A1H8 << ((rank - file) * 8) : DIAG_A1H8 >> ((file - rank) * 8);
    if (fileDiff + rankDiff == 0) {
        int s = (rank + file - 7) * 8;
        return s >= 0 ? DIAG_A8H1 << s : DIAG_A8H1 >> (-s);
    }
    return 0;
}

static int g_inited = 0;
/* MoveGeneratorBitboard.h:982-1021 (the part the Kogge-Stone build needs) */
void or_init(void) {
    if (g_inited) return;
    for (int i = 0; i < 64; i++) {
        u64 x = BIT(i);
        RookAtt[i] = northAttacks(x, ~0ULL) | southAttacks(x, ~0ULL) | eastAttacks(x, ~0ULL) | westAttacks(x, ~0ULL);
        BishopAtt[i] = northEastAttacks(x, ~0ULL) | northWestAttacks(x, ~0ULL) | southEastAttacks(x, ~0ULL) | southWestAttacks(x, ~0ULL);
        KnightAtt[i] = knightAttacks(x);
        KingAtt[i] = kingAttacks(x);
    }
    for (int i = 0; i < 64; i++)
        for (int j = 0; j < 64; j++) {
            if (i <= j) { Between[i][j] = squaresInBetween(i, j); Between[j][i] = Between[i][j]; }
            Line[i][j] = squaresInLine(i, j);
        }
    g_inited = 1;
}

/* MoveGeneratorBitboard.h:1213-1247 */
static u64 findPinnedPieces(u64 myPieces, u64 enemyBishops, u64 enemyRooks, u64 allPieces, int k) {
    (void)myPieces;
    u64 attackers = (BishopAtt[k] & enemyBishops) | (RookAtt[k] & enemyRooks);
    u64 pinned = 0;
    while (attackers) {
        u64 a = one(attackers);
        u64 between = Between[bsf(a)][k] & allPieces;
        if (!multiple(between)) pinned |= between;
        attackers ^= a;
    }
    return pinned;
}

/* MoveGeneratorBitboard.h:1254-1298 */
static u64 findAttackedSquares(u64 empty, u64 eB, u64 eR, u64 eP, u64 eN, u64 eK, u64 myKing, int enemyColor) {
    u64 a = (enemyColor == WHITE) ? (northEastOne(eP) | northWestOne(eP)) : (southEastOne(eP) | southWestOne(eP));
    a |= knightAttacks(eN);
    a |= bishopAttacks(eB, empty | myKing);
    a |= rookAttacks(eR, empty | myKing);
    a |= kingAttacks(eK);
    return a;
}

typedef struct {
    u64 allPawns, allPieces, myPieces, enemyPieces, enemyBishops, enemyRooks, myKing, pinned, threatened;
    int k;
} Ctx;

static void prologue(const OPos *p, int c, Ctx *x) {
    x->allPawns = p->pawns & RANKS2TO7;
    x->allPieces = p->kings | x->allPawns | p->knights | p->bishopQueens | p->rookQueens;
    u64 black = x->allPieces & ~p->whitePieces;
    x->myPieces = c == WHITE ? p->whitePieces : black;
    x->enemyPieces = c == WHITE ? black : p->whitePieces;
    x->enemyBishops = p->bishopQueens & x->enemyPieces;
    x->enemyRooks = p->rookQueens & x->enemyPieces;
    x->myKing = p->kings & x->myPieces;
    x->k = bsf(x->myKing);
    x->pinned = findPinnedPieces(x->myPieces, x->enemyBishops, x->enemyRooks, x->allPieces, x->k);
    x->threatened = findAttackedSquares(~x->allPieces, x->enemyBishops, x->enemyRooks, x->allPawns & x->enemyPieces,
                                        p->knights & x->enemyPieces, p->kings & x->enemyPieces, x->myKing, !c);
}

static u64 epTargetOf(const OPos *p, int c) {
    if (!ep(p)) return 0;
    return c == BLACK ? BIT(ep(p) - 1) << 16 : BIT(ep(p) - 1) << 40;
}

static u64 checkers(const OPos *p, int c, const Ctx *x) {
    u64 king = x->myKing, a = 0;
    u64 enemyPawns = x->allPawns & x->enemyPieces;
    a |= (c == WHITE ? (northEastOne(king) | northWestOne(king)) : (southEastOne(king) | southWestOne(king))) & enemyPawns;
    a |= knightAttacks(king) & (p->knights & x->enemyPieces);
    a |= bishopAttacks(king, ~x->allPieces) & x->enemyBishops;
    a |= rookAttacks(king, ~x->allPieces) & x->enemyRooks;
    return a;
}

/* ---------------------------------------------------------------- countMoves */
/* MoveGeneratorBitboard.h:2825-2994 */
static u32 countMovesOutOfCheck(const OPos *p, int c, const Ctx *x) {
    u32 n = 0;
    u64 attackers = checkers(p, c, x);
    n += popc(KingAtt[x->k] & ~(x->threatened | x->myPieces));
    if (multiple(attackers)) return n;
    u64 safe = attackers | Between[x->k][bsf(attackers)];
    u64 myPieces = x->myPieces & ~x->pinned;
    u64 myPawns = x->allPawns & myPieces;
    u64 dbl = RANK3 << (c * 24);
    u64 ept = epTargetOf(p, c);
    u64 epCap = c == WHITE ? southOne(ept) : northOne(ept);
    if (epCap != attackers) ept = 0;
    while (myPawns) {
        u64 pawn = one(myPawns);
        u64 dst = (c == WHITE ? northOne(pawn) : southOne(pawn)) & ~x->allPieces;
        if (dst) {
            if (dst & safe) n += (dst & (RANK1 | RANK8)) ? 4 : 1;
            else {
                dst = (c == WHITE ? northOne(dst & dbl) : southOne(dst & dbl)) & safe & ~x->allPieces;
                if (dst) n++;
            }
        }
        u64 wcap = c == WHITE ? northWestOne(pawn) : southWestOne(pawn);
        u64 ecap = c == WHITE ? northEastOne(pawn) : southEastOne(pawn);
        dst = (wcap | ecap) & x->enemyPieces & safe;
        if (dst) n += (dst & (RANK1 | RANK8)) ? 4 : 1;
        if ((wcap | ecap) & ept) n++;
        myPawns ^= pawn;
    }
    u64 kn = p->knights & myPieces;
    while (kn) { u64 b = one(kn); n += popc(KnightAtt[bsf(b)] & safe); kn ^= b; }
    u64 bs = p->bishopQueens & myPieces;
    while (bs) { u64 b = one(bs); n += popc(bishopAttacks(b, ~x->allPieces) & safe); bs ^= b; }
    u64 rs = p->rookQueens & myPieces;
    while (rs) { u64 b = one(rs); n += popc(rookAttacks(b, ~x->allPieces) & safe); rs ^= b; }
    return n;
}

/* MoveGeneratorBitboard.h:3004-3295 */
u32 or_count_moves(const OPos *p) {
    int c = side(p);
    Ctx x;
    prologue(p, c, &x);
    if (x.threatened & x.myKing) return countMovesOutOfCheck(p, c, &x);
    u32 n = 0;
    u64 myPawns = x.allPawns & x.myPieces;
    u64 ept = epTargetOf(p, c);
    if (ept) {
        u64 epCap = c == WHITE ? southOne(ept) : northOne(ept);
        u64 src = (eastOne(epCap) | westOne(epCap)) & myPawns;
        while (src) {
            u64 pawn = one(src);
            if (pawn & x.pinned) {
                if (ept & Line[bsf(pawn)][x.k]) n++;
            } else {
                u64 prop = ~x.allPieces | epCap | pawn;
                u64 chk = (eastAttacks(x.enemyRooks, prop) | westAttacks(x.enemyRooks, prop)) & x.myKing;
                if (!chk) n++;
            }
            src ^= pawn;
        }
    }
    u64 dbl = RANK3 << (c * 24);
    u64 pp = myPawns & x.pinned;
    while (pp) {
        u64 pawn = one(pp);
        u64 line = Line[bsf(pawn)][x.k];
        u64 dst = (c == WHITE ? northOne(pawn) : southOne(pawn)) & line & ~x.allPieces;
        if (dst) {
            n++;
            dst = (c == WHITE ? northOne(dst & dbl) : southOne(dst & dbl)) & ~x.allPieces;
            if (dst) n++;
        }
        dst = (c == WHITE ? northWestOne(pawn) : southWestOne(pawn)) & line;
        dst |= (c == WHITE ? northEastOne(pawn) : southEastOne(pawn)) & line;
        if (dst & x.enemyPieces) n += (dst & (RANK1 | RANK8)) ? 4 : 1;
        pp ^= pawn;
    }
    myPawns &= ~x.pinned;
    u64 d = (c == WHITE ? northOne(myPawns) : southOne(myPawns)) & ~x.allPieces;
    n += popc(d) + 3 * popc(d & (RANK1 | RANK8));
    d = (c == WHITE ? northOne(d & dbl) : southOne(d & dbl)) & ~x.allPieces;
    n += popc(d);
    d = (c == WHITE ? northWestOne(myPawns) : southWestOne(myPawns)) & x.enemyPieces;
    n += popc(d) + 3 * popc(d & (RANK1 | RANK8));
    d = (c == WHITE ? northEastOne(myPawns) : southEastOne(myPawns)) & x.enemyPieces;
    n += popc(d) + 3 * popc(d & (RANK1 | RANK8));
    if (c == WHITE) {
        if ((wc(p) & 1) && !(F1G1 & x.allPieces) && !(F1G1 & x.threatened)) n++;
        if ((wc(p) & 2) && !(B1D1 & x.allPieces) && !(C1D1 & x.threatened)) n++;
    } else {
        if ((bc(p) & 1) && !(F8G8 & x.allPieces) && !(F8G8 & x.threatened)) n++;
        if ((bc(p) & 2) && !(B8D8 & x.allPieces) && !(C8D8 & x.threatened)) n++;
    }
    n += popc(KingAtt[x.k] & ~(x.threatened | x.myPieces));
    u64 kn = p->knights & x.myPieces & ~x.pinned;
    while (kn) { u64 b = one(kn); n += popc(KnightAtt[bsf(b)] & ~x.myPieces); kn ^= b; }
    u64 myB = p->bishopQueens & x.myPieces;
    u64 bs = myB & x.pinned;
    while (bs) { u64 b = one(bs); n += popc(bishopAttacks(b, ~x.allPieces) & ~x.myPieces & Line[bsf(b)][x.k]); bs ^= b; }
    bs = myB & ~x.pinned;
    while (bs) { u64 b = one(bs); n += popc(bishopAttacks(b, ~x.allPieces) & ~x.myPieces); bs ^= b; }
    u64 myR = p->rookQueens & x.myPieces;
    u64 rs = myR & x.pinned;
    while (rs) { u64 b = one(rs); n += popc(rookAttacks(b, ~x.allPieces) & ~x.myPieces & Line[bsf(b)][x.k]); rs ^= b; }
    rs = myR & ~x.pinned;
    while (rs) { u64 b = one(rs); n += popc(rookAttacks(b, ~x.allPieces) & ~x.myPieces); rs ^= b; }
    return n;
}

/* ------------------------------------------------------------ generateBoards */
/* MoveGeneratorBitboard.h:1311-1338 (bitwise-magic switch off) */
static void updateCastleFlag(OPos *p, u64 dst, int c) {
    if (c == WHITE) {
        if (dst & BKR) set_bc(p, bc(p) & ~1);
        else if (dst & BQR) set_bc(p, bc(p) & ~2);
    } else {
        if (dst & WKR) set_wc(p, wc(p) & ~1);
        else if (dst & WQR) set_wc(p, wc(p) & ~2);
    }
}

typedef struct { OPos *out; u32 n; } Sink;
static void emit(Sink *s, const OPos *b) { s->out[s->n++] = *b; }

/* :1340-1397 */
static void addSlidingMove(Sink *s, const OPos *p, u64 src, u64 dst, int c) {
    OPos b = *p;
    b.bishopQueens = p->bishopQueens & ~dst;
    b.rookQueens = p->rookQueens & ~dst;
    b.kings = p->kings & ~dst;
    b.knights = p->knights & ~dst;
    b.pawns = p->pawns & ~(dst & RANKS2TO7);
    u64 isB = b.bishopQueens & src, isR = b.rookQueens & src;
    b.bishopQueens ^= isB; b.rookQueens ^= isR;
    b.bishopQueens |= isB ? dst : 0; b.rookQueens |= isR ? dst : 0;
    b.whitePieces = c == WHITE ? ((p->whitePieces ^ src) | dst) : (p->whitePieces & ~dst);
    set_side(&b, !c); set_ep(&b, 0);
    updateCastleFlag(&b, dst, c);
    updateCastleFlag(&b, src, !c);
    emit(s, &b);
}

/* :1400-1443 */
static void addKnightMove(Sink *s, const OPos *p, u64 src, u64 dst, int c) {
    OPos b = *p;
    b.bishopQueens = p->bishopQueens & ~dst;
    b.rookQueens = p->rookQueens & ~dst;
    b.kings = p->kings & ~dst;
    b.pawns = p->pawns & ~(dst & RANKS2TO7);
    b.knights = (p->knights ^ src) | dst;
    b.whitePieces = c == WHITE ? ((p->whitePieces ^ src) | dst) : (p->whitePieces & ~dst);
    set_side(&b, !c); set_ep(&b, 0);
    updateCastleFlag(&b, dst, c);
    emit(s, &b);
}

/* :1446-1488 */
static void addKingMove(Sink *s, const OPos *p, u64 src, u64 dst, int c) {
    OPos b = *p;
    b.bishopQueens = p->bishopQueens & ~dst;
    b.rookQueens = p->rookQueens & ~dst;
    b.knights = p->knights & ~dst;
    b.pawns = p->pawns & ~(dst & RANKS2TO7);
    b.kings = (p->kings ^ src) | dst;
    if (c == WHITE) { b.whitePieces = (p->whitePieces ^ src) | dst; set_wc(&b, 0); }
    else { b.whitePieces = p->whitePieces & ~dst; set_bc(&b, 0); }
    set_side(&b, !c); set_ep(&b, 0);
    updateCastleFlag(&b, dst, c);
    emit(s, &b);
}

/* :1491-1529 */
static void addCastleMove(Sink *s, const OPos *p, u64 kf, u64 kt, u64 rf, u64 rt, int c) {
    OPos b = *p;
    b.kings = (p->kings ^ kf) | kt;
    b.rookQueens = (p->rookQueens ^ rf) | rt;
    set_side(&b, !c); set_ep(&b, 0); set_hmc(&b, 0);
    if (c == WHITE) { b.whitePieces = (p->whitePieces ^ (kf | rf)) | (kt | rt); set_wc(&b, 0); }
    else { set_bc(&b, 0); b.whitePieces = p->whitePieces; }
    emit(s, &b);
}

/* :1534-1583 */
static void addSinglePawnMove(Sink *s, const OPos *p, u64 src, u64 dst, int c, int dbl, int idx) {
    OPos b = *p;
    b.bishopQueens = p->bishopQueens & ~dst;
    b.rookQueens = p->rookQueens & ~dst;
    b.knights = p->knights & ~dst;
    b.kings = p->kings;
    b.pawns = (p->pawns ^ src) | dst;
    b.whitePieces = c == WHITE ? ((p->whitePieces ^ src) | dst) : (p->whitePieces & ~dst);
    set_side(&b, !c);
    set_ep(&b, dbl ? (idx & 7) + 1 : 0);
    set_hmc(&b, 0);
    emit(s, &b);
}

/* :1585-1626 */
static void addEnPassentMove(Sink *s, const OPos *p, u64 src, u64 dst, int c) {
    OPos b = *p;
    u64 cap = c == WHITE ? southOne(dst) : northOne(dst);
    b.pawns = (p->pawns ^ (cap | src)) | dst;
    b.whitePieces = c == WHITE ? ((p->whitePieces ^ src) | dst) : (p->whitePieces ^ cap);
    set_side(&b, !c); set_hmc(&b, 0); set_ep(&b, 0);
    emit(s, &b);
}

/* :1630-1709 (promotions N, B, Q, R) */
static void addPawnMoves(Sink *s, const OPos *p, u64 src, u64 dst, int c) {
    if (dst & (RANK1 | RANK8)) {
        OPos b = *p;
        b.kings = p->kings;
        b.whitePieces = c == WHITE ? ((p->whitePieces ^ src) | dst) : (p->whitePieces & ~dst);
        b.pawns = p->pawns ^ src;
        set_side(&b, !c); set_ep(&b, 0); set_hmc(&b, 0);
        updateCastleFlag(&b, dst, c);
        b.knights = p->knights | dst; b.bishopQueens = p->bishopQueens & ~dst; b.rookQueens = p->rookQueens & ~dst;
        emit(s, &b);
        b.knights = p->knights & ~dst; b.bishopQueens = p->bishopQueens | dst; b.rookQueens = p->rookQueens & ~dst;
        emit(s, &b);
        b.rookQueens = p->rookQueens | dst;
        emit(s, &b);
        b.bishopQueens = p->bishopQueens & ~dst;
        emit(s, &b);
    } else {
        addSinglePawnMove(s, p, src, dst, c, 0, 0);
    }
}

/* :1741-1940 */
static void generateBoardsOutOfCheck(const OPos *p, int c, const Ctx *x, Sink *s) {
    u64 attackers = checkers(p, c, x);
    u64 km = KingAtt[x->k] & ~(x->threatened | x->myPieces);
    while (km) { u64 d = one(km); addKingMove(s, p, x->myKing, d, c); km ^= d; }
    if (multiple(attackers)) return;
    u64 safe = attackers | Between[x->k][bsf(attackers)];
    u64 myPieces = x->myPieces & ~x->pinned;
    u64 myPawns = x->allPawns & myPieces;
    u64 dbl = RANK3 << (c * 24);
    u64 ept = epTargetOf(p, c);
    u64 epCap = c == WHITE ? southOne(ept) : northOne(ept);
    if (epCap != attackers) ept = 0;
    while (myPawns) {
        u64 pawn = one(myPawns);
        u64 dst = (c == WHITE ? northOne(pawn) : southOne(pawn)) & ~x->allPieces;
        if (dst) {
            if (dst & safe) addPawnMoves(s, p, pawn, dst, c);
            else {
                dst = (c == WHITE ? northOne(dst & dbl) : southOne(dst & dbl)) & safe & ~x->allPieces;
                if (dst) addSinglePawnMove(s, p, pawn, dst, c, 1, bsf(pawn));
            }
        }
        u64 wcap = c == WHITE ? northWestOne(pawn) : southWestOne(pawn);
        u64 ecap = c == WHITE ? northEastOne(pawn) : southEastOne(pawn);
        dst = (wcap | ecap) & x->enemyPieces & safe;
        if (dst) addPawnMoves(s, p, pawn, dst, c);
        dst = (wcap | ecap) & ept;
        if (dst) addEnPassentMove(s, p, pawn, dst, c);
        myPawns ^= pawn;
    }
    u64 kn = p->knights & myPieces;
    while (kn) {
        u64 b = one(kn), m = KnightAtt[bsf(b)] & safe;
        while (m) { u64 d = one(m); addKnightMove(s, p, b, d, c); m ^= d; }
        kn ^= b;
    }
    u64 bs = p->bishopQueens & myPieces;
    while (bs) {
        u64 b = one(bs), m = bishopAttacks(b, ~x->allPieces) & safe;
        while (m) { u64 d = one(m); addSlidingMove(s, p, b, d, c); m ^= d; }
        bs ^= b;
    }
    u64 rs = p->rookQueens & myPieces;
    while (rs) {
        u64 b = one(rs), m = rookAttacks(b, ~x->allPieces) & safe;
        while (m) { u64 d = one(m); addSlidingMove(s, p, b, d, c); m ^= d; }
        rs ^= b;
    }
}

/* :1942-2261 */
u32 or_generate_boards(const OPos *p, OPos *out) {
    int c = side(p);
    Ctx x;
    Sink s = { out, 0 };
    prologue(p, c, &x);
    if (x.threatened & x.myKing) { generateBoardsOutOfCheck(p, c, &x, &s); return s.n; }
    u64 myPawns = x.allPawns & x.myPieces;
    u64 ept = epTargetOf(p, c);
    if (ept) {
        u64 epCap = c == WHITE ? southOne(ept) : northOne(ept);
        u64 src = (eastOne(epCap) | westOne(epCap)) & myPawns;
        while (src) {
            u64 pawn = one(src);
            if (pawn & x.pinned) {
                if (ept & Line[bsf(pawn)][x.k]) addEnPassentMove(&s, p, pawn, ept, c);
            } else {
                u64 prop = ~x.allPieces | epCap | pawn;
                u64 chk = (eastAttacks(x.enemyRooks, prop) | westAttacks(x.enemyRooks, prop)) & x.myKing;
                if (!chk) addEnPassentMove(&s, p, pawn, ept, c);
            }
            src ^= pawn;
        }
    }
    u64 dbl = RANK3 << (c * 24);
    u64 pp = myPawns & x.pinned;
    while (pp) {
        u64 pawn = one(pp);
        int pi = bsf(pawn);
        u64 line = Line[pi][x.k];
        u64 dst = (c == WHITE ? northOne(pawn) : southOne(pawn)) & line & ~x.allPieces;
        if (dst) {
            addSinglePawnMove(&s, p, pawn, dst, c, 0, pi);
            dst = (c == WHITE ? northOne(dst & dbl) : southOne(dst & dbl)) & ~x.allPieces;
            if (dst) addSinglePawnMove(&s, p, pawn, dst, c, 1, pi);
        }
        dst = (c == WHITE ? northWestOne(pawn) : southWestOne(pawn)) & line;
        dst |= (c == WHITE ? northEastOne(pawn) : southEastOne(pawn)) & line;
        if (dst & x.enemyPieces) addPawnMoves(&s, p, pawn, dst, c);
        pp ^= pawn;
    }
    myPawns &= ~x.pinned;
    while (myPawns) {
        u64 pawn = one(myPawns);
        u64 dst = (c == WHITE ? northOne(pawn) : southOne(pawn)) & ~x.allPieces;
        if (dst) {
            addPawnMoves(&s, p, pawn, dst, c);
            dst = (c == WHITE ? northOne(dst & dbl) : southOne(dst & dbl)) & ~x.allPieces;
            if (dst) addSinglePawnMove(&s, p, pawn, dst, c, 1, bsf(pawn));
        }
        dst = (c == WHITE ? northWestOne(pawn) : southWestOne(pawn)) & x.enemyPieces;
        if (dst) addPawnMoves(&s, p, pawn, dst, c);
        dst = (c == WHITE ? northEastOne(pawn) : southEastOne(pawn)) & x.enemyPieces;
        if (dst) addPawnMoves(&s, p, pawn, dst, c);
        myPawns ^= pawn;
    }
    if (c == WHITE) {
        if ((wc(p) & 1) && !(F1G1 & x.allPieces) && !(F1G1 & x.threatened)) addCastleMove(&s, p, BIT(E1), BIT(G1), BIT(H1), BIT(F1), c);
        if ((wc(p) & 2) && !(B1D1 & x.allPieces) && !(C1D1 & x.threatened)) addCastleMove(&s, p, BIT(E1), BIT(C1), BIT(A1), BIT(D1), c);
    } else {
        if ((bc(p) & 1) && !(F8G8 & x.allPieces) && !(F8G8 & x.threatened)) addCastleMove(&s, p, BIT(E8), BIT(G8), BIT(H8), BIT(F8), c);
        if ((bc(p) & 2) && !(B8D8 & x.allPieces) && !(C8D8 & x.threatened)) addCastleMove(&s, p, BIT(E8), BIT(C8), BIT(A8), BIT(D8), c);
    }
    u64 km = KingAtt[x.k] & ~(x.threatened | x.myPieces);
    while (km) { u64 d = one(km); addKingMove(&s, p, x.myKing, d, c); km ^= d; }
    u64 kn = p->knights & x.myPieces & ~x.pinned;
    while (kn) {
        u64 b = one(kn), m = KnightAtt[bsf(b)] & ~x.myPieces;
        while (m) { u64 d = one(m); addKnightMove(&s, p, b, d, c); m ^= d; }
        kn ^= b;
    }
    u64 myB = p->bishopQueens & x.myPieces;
    for (int pass = 0; pass < 2; pass++) {
        u64 bs = myB & (pass == 0 ? x.pinned : ~x.pinned);
        while (bs) {
            u64 b = one(bs), m = bishopAttacks(b, ~x.allPieces) & ~x.myPieces;
            if (pass == 0) m &= Line[bsf(b)][x.k];
            while (m) { u64 d = one(m); addSlidingMove(&s, p, b, d, c); m ^= d; }
            bs ^= b;
        }
    }
    u64 myR = p->rookQueens & x.myPieces;
    for (int pass = 0; pass < 2; pass++) {
        u64 rs = myR & (pass == 0 ? x.pinned : ~x.pinned);
        while (rs) {
            u64 b = one(rs), m = rookAttacks(b, ~x.allPieces) & ~x.myPieces;
            if (pass == 0) m &= Line[bsf(b)][x.k];
            while (m) { u64 d = one(m); addSlidingMove(&s, p, b, d, c); m ^= d; }
            rs ^= b;
        }
    }
    return s.n;
}

/* ------------------------------------------------------------- generateMoves */
/* CMove: from | to << 6 | flags << 12 (chess.h:234-293) */
enum { F_QUIET = 0, F_DPP = 1, F_KCASTLE = 2, F_QCASTLE = 3, F_CAP = 4, F_EP = 5,
       F_NPROMO = 8, F_BPROMO = 9, F_RPROMO = 10, F_QPROMO = 11 };
typedef struct { u16 *out; u32 n; } MSink;
static void addm(MSink *s, int from, int to, int fl) { s->out[s->n++] = (u16)((from & 63) | ((to & 63) << 6) | ((fl & 15) << 12)); }
/* addCompactPawnMoves :1723-1736 (N, B, Q, R) */
static void addpm(MSink *s, int from, u64 dst, int fl) {
    int to = bsf(dst);
    if (dst & (RANK1 | RANK8)) {
        addm(s, from, to, fl | F_NPROMO); addm(s, from, to, fl | F_BPROMO);
        addm(s, from, to, fl | F_QPROMO); addm(s, from, to, fl | F_RPROMO);
    } else addm(s, from, to, fl);
}
static int capf(const Ctx *x, u64 dst) { return (dst & x->enemyPieces) ? F_CAP : 0; }

/* :2266-2460: out of check the reference leaves king, knight and slider
 * captures unflagged ("TODO: set capture flag correctly"), pawn captures flagged */
static void generateMovesOutOfCheck(const OPos *p, int c, const Ctx *x, MSink *s) {
    u64 attackers = checkers(p, c, x);
    u64 km = KingAtt[x->k] & ~(x->threatened | x->myPieces);
    while (km) { u64 d = one(km); addm(s, x->k, bsf(d), 0); km ^= d; }
    if (multiple(attackers)) return;
    u64 safe = attackers | Between[x->k][bsf(attackers)];
    u64 myPieces = x->myPieces & ~x->pinned;
    u64 myPawns = x->allPawns & myPieces;
    u64 dbl = RANK3 << (c * 24);
    u64 ept = epTargetOf(p, c);
    u64 epCap = c == WHITE ? southOne(ept) : northOne(ept);
    if (epCap != attackers) ept = 0;
    while (myPawns) {
        u64 pawn = one(myPawns);
        int pf = bsf(pawn);
        u64 dst = (c == WHITE ? northOne(pawn) : southOne(pawn)) & ~x->allPieces;
        if (dst) {
            if (dst & safe) addpm(s, pf, dst, 0);
            else {
                dst = (c == WHITE ? northOne(dst & dbl) : southOne(dst & dbl)) & safe & ~x->allPieces;
                if (dst) addm(s, pf, bsf(dst), F_DPP);
            }
        }
        u64 wcap = c == WHITE ? northWestOne(pawn) : southWestOne(pawn);
        u64 ecap = c == WHITE ? northEastOne(pawn) : southEastOne(pawn);
        dst = (wcap | ecap) & x->enemyPieces & safe;
        if (dst) addpm(s, pf, dst, F_CAP);
        dst = (wcap | ecap) & ept;
        if (dst) addm(s, pf, bsf(dst), F_EP);
        myPawns ^= pawn;
    }
    u64 kn = p->knights & myPieces;
    while (kn) { u64 b = one(kn), m = KnightAtt[bsf(b)] & safe; while (m) { u64 d = one(m); addm(s, bsf(b), bsf(d), 0); m ^= d; } kn ^= b; }
    u64 bs = p->bishopQueens & myPieces;
    while (bs) { u64 b = one(bs), m = bishopAttacks(b, ~x->allPieces) & safe; while (m) { u64 d = one(m); addm(s, bsf(b), bsf(d), 0); m ^= d; } bs ^= b; }
    u64 rs = p->rookQueens & myPieces;
    while (rs) { u64 b = one(rs), m = rookAttacks(b, ~x->allPieces) & safe; while (m) { u64 d = one(m); addm(s, bsf(b), bsf(d), 0); m ^= d; } rs ^= b; }
}

/* :2471-2823; the host build flags every capture (:2525-2528) */
u32 or_generate_moves(const OPos *p, u16 *out) {
    int c = side(p);
    Ctx x;
    MSink s = { out, 0 };
    prologue(p, c, &x);
    if (x.threatened & x.myKing) { generateMovesOutOfCheck(p, c, &x, &s); return s.n; }
    u64 myPawns = x.allPawns & x.myPieces;
    u64 ept = epTargetOf(p, c);
    if (ept) {
        u64 epCap = c == WHITE ? southOne(ept) : northOne(ept);
        u64 src = (eastOne(epCap) | westOne(epCap)) & myPawns;
        while (src) {
            u64 pawn = one(src);
            if (pawn & x.pinned) {
                if (ept & Line[bsf(pawn)][x.k]) addm(&s, bsf(pawn), bsf(ept), F_EP);
            } else {
                u64 prop = ~x.allPieces | epCap | pawn;
                u64 chk = (eastAttacks(x.enemyRooks, prop) | westAttacks(x.enemyRooks, prop)) & x.myKing;
                if (!chk) addm(&s, bsf(pawn), bsf(ept), F_EP);
            }
            src ^= pawn;
        }
    }
    u64 dbl = RANK3 << (c * 24);
    u64 pp = myPawns & x.pinned;
    while (pp) {
        u64 pawn = one(pp);
        int pi = bsf(pawn);
        u64 line = Line[pi][x.k];
        u64 dst = (c == WHITE ? northOne(pawn) : southOne(pawn)) & line & ~x.allPieces;
        if (dst) {
            addm(&s, pi, bsf(dst), 0);
            dst = (c == WHITE ? northOne(dst & dbl) : southOne(dst & dbl)) & ~x.allPieces;
            if (dst) addm(&s, pi, bsf(dst), F_DPP);
        }
        dst = (c == WHITE ? northWestOne(pawn) : southWestOne(pawn)) & line;
        dst |= (c == WHITE ? northEastOne(pawn) : southEastOne(pawn)) & line;
        if (dst & x.enemyPieces) addpm(&s, pi, dst, F_CAP);
        pp ^= pawn;
    }
    myPawns &= ~x.pinned;
    while (myPawns) {
        u64 pawn = one(myPawns);
        int pf = bsf(pawn);
        u64 dst = (c == WHITE ? northOne(pawn) : southOne(pawn)) & ~x.allPieces;
        if (dst) {
            addpm(&s, pf, dst, 0);
            dst = (c == WHITE ? northOne(dst & dbl) : southOne(dst & dbl)) & ~x.allPieces;
            if (dst) addm(&s, pf, bsf(dst), F_DPP);
        }
        dst = (c == WHITE ? northWestOne(pawn) : southWestOne(pawn)) & x.enemyPieces;
        if (dst) addpm(&s, pf, dst, F_CAP);
        dst = (c == WHITE ? northEastOne(pawn) : southEastOne(pawn)) & x.enemyPieces;
        if (dst) addpm(&s, pf, dst, F_CAP);
        myPawns ^= pawn;
    }
    if (c == WHITE) {
        if ((wc(p) & 1) && !(F1G1 & x.allPieces) && !(F1G1 & x.threatened)) addm(&s, E1, G1, F_KCASTLE);
        if ((wc(p) & 2) && !(B1D1 & x.allPieces) && !(C1D1 & x.threatened)) addm(&s, E1, C1, F_QCASTLE);
    } else {
        if ((bc(p) & 1) && !(F8G8 & x.allPieces) && !(F8G8 & x.threatened)) addm(&s, E8, G8, F_KCASTLE);
        if ((bc(p) & 2) && !(B8D8 & x.allPieces) && !(C8D8 & x.threatened)) addm(&s, E8, C8, F_QCASTLE);
    }
    u64 km = KingAtt[x.k] & ~(x.threatened | x.myPieces);
    while (km) { u64 d = one(km); addm(&s, x.k, bsf(d), capf(&x, d)); km ^= d; }
    u64 kn = p->knights & x.myPieces & ~x.pinned;
    while (kn) { u64 b = one(kn), m = KnightAtt[bsf(b)] & ~x.myPieces; while (m) { u64 d = one(m); addm(&s, bsf(b), bsf(d), capf(&x, d)); m ^= d; } kn ^= b; }
    u64 myB = p->bishopQueens & x.myPieces;
    for (int pass = 0; pass < 2; pass++) {
        u64 bs = myB & (pass == 0 ? x.pinned : ~x.pinned);
        while (bs) {
            u64 b = one(bs), m = bishopAttacks(b, ~x.allPieces) & ~x.myPieces;
            if (pass == 0) m &= Line[bsf(b)][x.k];
            while (m) { u64 d = one(m); addm(&s, bsf(b), bsf(d), capf(&x, d)); m ^= d; }
            bs ^= b;
        }
    }
    u64 myR = p->rookQueens & x.myPieces;
    for (int pass = 0; pass < 2; pass++) {
        u64 rs = myR & (pass == 0 ? x.pinned : ~x.pinned);
        while (rs) {
            u64 b = one(rs), m = rookAttacks(b, ~x.allPieces) & ~x.myPieces;
            if (pass == 0) m &= Line[bsf(b)][x.k];
            while (m) { u64 d = one(m); addm(&s, bsf(b), bsf(d), capf(&x, d)); m ^= d; }
            rs ^= b;
        }
    }
    return s.n;
}

/* ------------------------------------------------------------------- Zobrist */
/* ZobristRandoms word offsets (chess.h:326-333) */
#define Z_PIECE(color, piece, sq) ((color) * 384 + ((piece) - 1) * 64 + (sq))
#define Z_CASTLE(color, qs) (768 + (color) * 2 + (qs))
#define Z_EP(file) (772 + (file))
#define Z_SIDE 780
static inline void zx(OKey *k, int w) { k->lo ^= ZLO[w]; k->hi ^= ZHI[w]; }

/* MoveGeneratorBitboard.h:3914-3980 */
void or_zobrist128(const OPos *p, OKey *out) {
    OKey k = { 0, 0 };
    if (side(p) == WHITE) zx(&k, Z_SIDE);
    if (wc(p) & 1) zx(&k, Z_CASTLE(WHITE, 0));
    if (wc(p) & 2) zx(&k, Z_CASTLE(WHITE, 1));
    if (bc(p) & 1) zx(&k, Z_CASTLE(BLACK, 0));
    if (bc(p) & 2) zx(&k, Z_CASTLE(BLACK, 1));
    if (ep(p)) zx(&k, Z_EP(ep(p) - 1));
    u64 allPawns = p->pawns & RANKS2TO7;
    u64 all = p->kings | allPawns | p->knights | p->bishopQueens | p->rookQueens;
    while (all) {
        u64 b = one(all);
        int sq = bsf(b), color = !(b & p->whitePieces);
        if (b & allPawns) zx(&k, Z_PIECE(color, PAWN, sq));
        else if (b & p->kings) zx(&k, Z_PIECE(color, KING, sq));
        else if (b & p->knights) zx(&k, Z_PIECE(color, KNIGHT, sq));
        else if (b & p->rookQueens & p->bishopQueens) zx(&k, Z_PIECE(color, QUEEN, sq));
        else if (b & p->rookQueens) zx(&k, Z_PIECE(color, ROOK, sq));
        else if (b & p->bishopQueens) zx(&k, Z_PIECE(color, BISHOP, sq));
        all ^= b;
    }
    *out = k;
}

static void castleKeys(const OPos *p, OKey *h) {
    if (wc(p) & 1) zx(h, Z_CASTLE(WHITE, 0));
    if (wc(p) & 2) zx(h, Z_CASTLE(WHITE, 1));
    if (bc(p) & 1) zx(h, Z_CASTLE(BLACK, 0));
    if (bc(p) & 2) zx(h, Z_CASTLE(BLACK, 1));
    if (ep(p)) zx(h, Z_EP(ep(p) - 1));
}

/* MoveGeneratorBitboard.h:3575-3836 with updateHash = true */
void or_make_move(OPos *p, OKey *h, u16 move) {
    int c = side(p);
    int from = move & 63, to = (move >> 6) & 63, fl = (move >> 12) & 15;
    u64 src = BIT(from), dst = BIT(to);
    u64 queens = p->bishopQueens & p->rookQueens;
    int piece;
    if (p->kings & src) piece = KING;
    else if (p->knights & src) piece = KNIGHT;
    else if ((p->pawns & RANKS2TO7) & src) piece = PAWN;
    else if (queens & src) piece = QUEEN;
    else if (p->bishopQueens & src) piece = BISHOP;
    else piece = ROOK;
    zx(h, Z_PIECE(c, piece, from));
    if (fl == 8 || fl == 12) piece = KNIGHT;
    else if (fl == 9 || fl == 13) piece = BISHOP;
    else if (fl == 10 || fl == 14) piece = ROOK;
    else if (fl == 11 || fl == 15) piece = QUEEN;
    {
        int dp = 0;
        if (p->kings & dst) dp = KING;
        else if (p->knights & dst) dp = KNIGHT;
        else if ((p->pawns & RANKS2TO7) & dst) dp = PAWN;
        else if (queens & dst) dp = QUEEN;
        else if (p->bishopQueens & dst) dp = BISHOP;
        else if (p->rookQueens & dst) dp = ROOK;
        if (dp) zx(h, Z_PIECE(!c, dp, to));
    }
    zx(h, Z_PIECE(c, piece, to));
    zx(h, Z_SIDE);
    castleKeys(p, h);
    p->bishopQueens &= ~src; p->rookQueens &= ~src; p->kings &= ~src; p->knights &= ~src; p->pawns &= ~(src & RANKS2TO7);
    p->bishopQueens &= ~dst; p->rookQueens &= ~dst; p->kings &= ~dst; p->knights &= ~dst; p->pawns &= ~(dst & RANKS2TO7);
    if (piece == KING) { p->kings |= dst; if (c == WHITE) set_wc(p, 0); else set_bc(p, 0); }
    if (piece == KNIGHT) p->knights |= dst;
    if (piece == PAWN) p->pawns |= dst;
    if (piece == BISHOP || piece == QUEEN) p->bishopQueens |= dst;
    if (piece == ROOK || piece == QUEEN) p->rookQueens |= dst;
    if (c == WHITE) p->whitePieces = (p->whitePieces ^ src) | dst;
    else p->whitePieces &= ~dst;
    if (fl == F_EP) {
        u64 cap = c == WHITE ? southOne(dst) : northOne(dst);
        p->pawns &= ~(cap & RANKS2TO7);
        zx(h, Z_PIECE(!c, PAWN, bsf(cap)));
        if (c == BLACK) p->whitePieces &= ~cap;
    }
    if (c == WHITE) {
        if (fl == F_KCASTLE) { p->rookQueens = (p->rookQueens ^ BIT(H1)) | BIT(F1); p->whitePieces = (p->whitePieces ^ BIT(H1)) | BIT(F1); zx(h, Z_PIECE(c, ROOK, H1)); zx(h, Z_PIECE(c, ROOK, F1)); }
        else if (fl == F_QCASTLE) { p->rookQueens = (p->rookQueens ^ BIT(A1)) | BIT(D1); p->whitePieces = (p->whitePieces ^ BIT(A1)) | BIT(D1); zx(h, Z_PIECE(c, ROOK, A1)); zx(h, Z_PIECE(c, ROOK, D1)); }
    } else {
        if (fl == F_KCASTLE) { p->rookQueens = (p->rookQueens ^ BIT(H8)) | BIT(F8); zx(h, Z_PIECE(c, ROOK, H8)); zx(h, Z_PIECE(c, ROOK, F8)); }
        else if (fl == F_QCASTLE) { p->rookQueens = (p->rookQueens ^ BIT(A8)) | BIT(D8); zx(h, Z_PIECE(c, ROOK, A8)); zx(h, Z_PIECE(c, ROOK, D8)); }
    }
    set_side(p, !c);
    set_ep(p, 0);
    updateCastleFlag(p, dst, c);
    if (piece == ROOK) updateCastleFlag(p, src, !c);
    if (fl == F_DPP) {
        u64 allPawns = p->pawns & RANKS2TO7;
        u64 all = p->kings | allPawns | p->knights | p->bishopQueens | p->rookQueens;
        u64 black = all & ~p->whitePieces;
        u64 enemy = c == WHITE ? black : p->whitePieces;
        if ((eastOne(dst) | westOne(dst)) & allPawns & enemy) set_ep(p, (from & 7) + 1);
    }
    castleKeys(p, h);
}

/* ----------------------------------------------------------------- perft_bb */
/* perft_bb.h:79-101 */
u64 or_perft(const OPos *p, u32 depth) {
    OPos child[256];
    if (depth == 0) return 1;
    if (depth == 1) return or_count_moves(p);
    u32 n = or_generate_boards(p, child);
    u64 sum = 0;
    for (u32 i = 0; i < n; i++) sum += or_perft(&child[i], depth - 1);
    return sum;
}

/* Root-move split over threads (the multi-threaded CPU baseline leg). */
typedef struct { const OPos *boards; u32 n; u32 depth; u64 *res; int next; pthread_mutex_t *mu; } Job;
static void *worker(void *arg) {
    Job *j = (Job *)arg;
    for (;;) {
        pthread_mutex_lock(j->mu);
        int i = j->next++;
        pthread_mutex_unlock(j->mu);
        if (i >= (int)j->n) break;
        j->res[i] = or_perft(&j->boards[i], j->depth);
    }
    return 0;
}
u64 or_perft_mt(const OPos *p, u32 depth, int threads) {
    if (depth <= 2 || threads <= 1) return or_perft(p, depth);
    OPos child[256];
    u32 n = or_generate_boards(p, child);
    /* split two plies deep so threads balance */
    OPos *grand = (OPos *)malloc(sizeof(OPos) * 256 * 256);
    u32 m = 0;
    for (u32 i = 0; i < n; i++) m += or_generate_boards(&child[i], grand + m);
    u64 *res = (u64 *)calloc(m ? m : 1, sizeof(u64));
    pthread_mutex_t mu = PTHREAD_MUTEX_INITIALIZER;
    Job j = { grand, m, depth - 2, res, 0, &mu };
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * threads);
    for (int t = 0; t < threads; t++) pthread_create(&th[t], 0, worker, &j);
    for (int t = 0; t < threads; t++) pthread_join(th[t], 0);
    u64 sum = 0;
    for (u32 i = 0; i < m; i++) sum += res[i];
    free(th); free(res); free(grand);
    return sum;
}

/* divide: per-root-move perft over the generateMoves list */
u32 or_divide(const OPos *p, u32 depth, u16 *moves, u64 *counts) {
    u32 n = or_generate_moves(p, moves);
    for (u32 i = 0; i < n; i++) {
        OPos c = *p;
        OKey h = { 0, 0 };
        or_make_move(&c, &h, moves[i]);
        counts[i] = depth <= 1 ? 1 : or_perft(&c, depth - 1);
    }
    return n;
}

/* ----------------------------------------------------------------------- FEN */
/* util.cpp:189-281 and board088ToHexBB util.cpp:92-140 (clocks are ignored,
 * as in the reference; the en-passant file is taken as written). */
int or_fen(const char *fen, OPos *out) {
    memset(out, 0, sizeof(*out));
    int i = 0, row = 0, col = 0;
    for (; fen[i]; i++) {
        char ch = fen[i];
        if (ch == '/' || ch == '\\') { row++; col = 0; }
        else if (ch >= '1' && ch <= '8') col += ch - '0';
        else {
            const char *pcs = "PNBRQKpnbrqk";
            const char *f = strchr(pcs, ch);
            if (f && ch) {
                int idx = (int)(f - pcs), color = idx >= 6, piece = idx % 6 + 1;
                int r = 7 - row;
                if (r < 0 || col > 7) return -1;
                u64 b = BIT(r * 8 + col);
                if (color == WHITE) out->whitePieces |= b;
                switch (piece) {
                case PAWN: out->pawns |= b; break;
                case KNIGHT: out->knights |= b; break;
                case BISHOP: out->bishopQueens |= b; break;
                case ROOK: out->rookQueens |= b; break;
                case QUEEN: out->bishopQueens |= b; out->rookQueens |= b; break;
                case KING: out->kings |= b; break;
                }
                col++;
            }
        }
        if (row >= 7 && col == 8) break;
    }
    if (!fen[i]) return -1;
    i++;
    while (fen[i] == ' ') i++;
    set_side(out, (fen[i] == 'b' || fen[i] == 'B') ? BLACK : WHITE);
    if (fen[i]) i++;
    while (fen[i] == ' ') i++;
    int w = 0, b = 0;
    while (fen[i] && fen[i] != ' ') {
        if (fen[i] == 'k') b |= 1;
        else if (fen[i] == 'q') b |= 2;
        else if (fen[i] == 'K') w |= 1;
        else if (fen[i] == 'Q') w |= 2;
        i++;
    }
    set_wc(out, w); set_bc(out, b);
    while (fen[i] == ' ') i++;
    if (fen[i] >= 'a' && fen[i] <= 'h') set_ep(out, fen[i] - 'a' + 1);
    return 0;
}
