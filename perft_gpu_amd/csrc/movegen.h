// movegen.h -- legal move counting, generation and make-move for gfx950.
//
// Written for 64-lane wavefronts: the leaf count is branch-light so that the
// lanes of a wave (sibling positions) stay converged.
//  * One "target" mask folds the reference's separate out-of-check path
//    (MoveGeneratorBitboard.h:2825-2994) into the normal one: not in check ->
//    ~own, single check -> checker | squares between, double check -> 0.
//    A pinned piece can never land on a check-resolving square, so the
//    result equals the reference's countMovesOutOfCheck.
//  * Pawns are counted in bulk (one shift + popcount per direction); pinned
//    pawns move only along the king line they sit on, which is applied as a
//    per-direction mask instead of a per-pawn loop.  Knights take one table
//    lookup + popcount each (at most two knights is cheaper than eight
//    direction shifts).
//  * With Kogge-Stone sliders, moves are also counted in bulk: the rays of
//    different sliders in one direction are disjoint (a ray stops at the
//    first occupied square), so popcount(fill_d(all movers) & target) sums
//    the per-piece counts.  Pinned sliders join the fills of the axis their
//    pin lies on.  The magic variants count per piece.
//  * Enemy attacks (with the moving king removed, so sliders see through
//    it), pins and slider checkers follow the reference's method
//    (findAttackedSquares :1254, findPinnedPieces :1213).
// The same source compiles for the host so the library can list moves and
// build workloads without a GPU; perft itself only runs on the device.
#pragma once
#include "pg_types.h"

namespace pg {

enum Slider { SL_KOGGE = 0, SL_PLAIN = 1, SL_FANCY = 2, SL_BYTE = 3, SL_COUNT = 4 };

// Fixed-shift fancy magic entry: factor, base offset into the shared
// 97264-entry table, and (byte-lookup variant) offset into the per-piece list
// of distinct attack sets.  Same fields as the reference's FancyMagicEntry
// (chess.h:343-359).
struct FancyEntry {
    u64 factor;
    int base;
    int offset;
};
static_assert(sizeof(FancyEntry) == 16, "fancy entry is 16 bytes");

constexpr int ROOK_BITS = 12, BISHOP_BITS = 9;
constexpr int FANCY_TABLE_SIZE = 97264, BYTE_ROOK_SIZE = 4900, BYTE_BISHOP_SIZE = 1428;
constexpr int ZOB_WORDS = 782;

// Table pointers (device or host memory; on the device the byte-lookup
// variant repoints some of them at LDS copies).
struct Tables {
    const u64 *knight, *king;            // [64] empty-board attacks
    const u64 *rookEmpty, *bishopEmpty;  // [64]
    const u64 *between;                  // [64*64] squares strictly between
    const u64 *line;                     // [64*64] full line through two squares
    const u64 *rookMask, *bishopMask;    // [64] relevant occupancy (edges off)
    const u64 *rookMagic, *bishopMagic;  // [64] plain magic factors
    const u64 *rookPlain, *bishopPlain;  // [64*4096], [64*512]
    const FancyEntry *rookFancy, *bishopFancy;  // [64]
    const u64 *fancy;                    // [97264]
    const u8 *byteIdx;                   // [97264]
    const u64 *byteRook, *byteBishop;    // [4900], [1428]
    const u64 *zobLo, *zobHi;            // [782] each
};

PG_HD int popc(u64 x) { return __builtin_popcountll(x); }
PG_HD int ctz(u64 x) { return __builtin_ctzll(x); }
PG_HD u64 bit(int s) { return 1ULL << s; }
PG_HD bool several(u64 x) { return (x & (x - 1)) != 0; }

// Three-input bitwise function given by its truth table F: bit i of F is the
// result for inputs (a, b, c) = (i >> 2 & 1, i >> 1 & 1, i & 1), so F for an
// expression is that expression evaluated on a = 0xF0, b = 0xCC, c = 0xAA.
// On gfx950 each 32-bit half is one v_bitop3_b32; the compiler does not form
// these from 64-bit and/or chains by itself.
template <u32 F>
PG_HD u64 bop3(u64 a, u64 b, u64 c) {
#if defined(__HIP_DEVICE_COMPILE__)
    const u32 lo = __builtin_amdgcn_bitop3_b32((u32)a, (u32)b, (u32)c, F);
    const u32 hi = __builtin_amdgcn_bitop3_b32((u32)(a >> 32), (u32)(b >> 32), (u32)(c >> 32), F);
    return ((u64)hi << 32) | lo;
#else
    u64 r = 0;
    for (int i = 0; i < 8; i++)
        if ((F >> i) & 1) r |= ((i & 4) ? a : ~a) & ((i & 2) ? b : ~b) & ((i & 1) ? c : ~c);
    return r;
#endif
}
PG_HD u64 or_and(u64 a, u64 b, u64 c) { return bop3<0xF8>(a, b, c); }    // a | (b & c)
PG_HD u64 and3(u64 a, u64 b, u64 c) { return bop3<0x80>(a, b, c); }      // a & b & c
PG_HD u64 or3(u64 a, u64 b, u64 c) { return bop3<0xFE>(a, b, c); }       // a | b | c
PG_HD u64 and_or(u64 a, u64 b, u64 c) { return bop3<0xE0>(a, b, c); }    // a & (b | c)
PG_HD u64 and_nor(u64 a, u64 b, u64 c) { return bop3<0x10>(a, b, c); }   // a & ~b & ~c
PG_HD u64 andn(u64 a, u64 b) { return bop3<0x30>(a, b, b); }             // a & ~b
PG_HD u64 or_xor(u64 a, u64 b, u64 c) { return bop3<0xF6>(a, b, c); }    // a | (b ^ c)
PG_HD u64 xor_and(u64 a, u64 b, u64 c) { return bop3<0x28>(a, b, c); }   // (a ^ b) & c

// A board the generator accepts: one king per side, no white bit on an empty
// square and an en-passant field of 0 (none) or a file 1..8.  (Inputs arriving over the C ABI are checked with this.)
PG_HD bool pos_sane(const Pos &p) {
    const u64 occ = pos_occ(p);
    return popc(p.kings & p.white) == 1 && popc(p.kings & ~p.white) == 1 && (p.white & ~occ) == 0 &&
           pos_ep(p) <= 8;
}

// ---------------------------------------------------------------- fills
// Kogge-Stone occluded fill, returns the squares attacked in one direction
// (first blocker included).  S = shift, M = wrap mask of the target file.
template <int S, u64 M>
PG_HD u64 fill_up(u64 g, u64 e) {
    e &= M;
    g |= (g << S) & e;
    e &= e << S;
    g |= (g << (2 * S)) & e;
    e &= e << (2 * S);
    g |= (g << (4 * S)) & e;
    return (g << S) & M;
}
template <int S, u64 M>
PG_HD u64 fill_dn(u64 g, u64 e) {
    e &= M;
    g |= (g >> S) & e;
    e &= e >> S;
    g |= (g >> (2 * S)) & e;
    e &= e >> (2 * S);
    g |= (g >> (4 * S)) & e;
    return (g >> S) & M;
}
PG_HD u64 att_n(u64 g, u64 e) { return fill_up<8, ~0ULL>(g, e); }
PG_HD u64 att_s(u64 g, u64 e) { return fill_dn<8, ~0ULL>(g, e); }
PG_HD u64 att_e(u64 g, u64 e) { return fill_up<1, ~FILEA>(g, e); }
PG_HD u64 att_w(u64 g, u64 e) { return fill_dn<1, ~FILEH>(g, e); }
PG_HD u64 att_ne(u64 g, u64 e) { return fill_up<9, ~FILEA>(g, e); }
PG_HD u64 att_nw(u64 g, u64 e) { return fill_up<7, ~FILEH>(g, e); }
PG_HD u64 att_se(u64 g, u64 e) { return fill_dn<7, ~FILEA>(g, e); }
PG_HD u64 att_sw(u64 g, u64 e) { return fill_dn<9, ~FILEH>(g, e); }
PG_HD u64 ks_rook(u64 g, u64 e) { return att_n(g, e) | att_s(g, e) | att_e(g, e) | att_w(g, e); }
PG_HD u64 ks_bishop(u64 g, u64 e) { return att_ne(g, e) | att_nw(g, e) | att_se(g, e) | att_sw(g, e); }

PG_HD u64 king_att(u64 k) {
    u64 a = ((k << 1) & ~FILEA) | ((k >> 1) & ~FILEH);
    k |= a;
    return a | (k << 8) | (k >> 8);
}
PG_HD u64 knight_att(u64 n) {
    u64 l1 = (n >> 1) & 0x7f7f7f7f7f7f7f7fULL, l2 = (n >> 2) & 0x3f3f3f3f3f3f3f3fULL;
    u64 r1 = (n << 1) & 0xfefefefefefefefeULL, r2 = (n << 2) & 0xfcfcfcfcfcfcfcfcULL;
    u64 h1 = l1 | r1, h2 = l2 | r2;
    return (h1 << 16) | (h1 >> 16) | (h2 << 8) | (h2 >> 8);
}

PG_HD u64 rev64(u64 x) {
#if defined(__clang__)
    return __builtin_bitreverse64(x);
#else
    x = ((x >> 1) & 0x5555555555555555ULL) | ((x & 0x5555555555555555ULL) << 1);
    x = ((x >> 2) & 0x3333333333333333ULL) | ((x & 0x3333333333333333ULL) << 2);
    x = ((x >> 4) & 0x0F0F0F0F0F0F0F0FULL) | ((x & 0x0F0F0F0F0F0F0F0FULL) << 4);
    return __builtin_bswap64(x);
#endif
}

PG_HD u64 nor2(u64 a, u64 b) { return bop3<0x03>(a, b, b); }  // ~(a | b)

// squares east of each slider in s up to and including its first blocker,
// given eo = ~(occ | FILEH); every slider must be in occ
PG_HD u64 east_sum(u64 s, u64 eo) { return (andn(s, FILEH) << 1) + eo; }

template <int S> PG_HD u64 shift(u64 x) {
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (S > 0 && S < 32) {
        u64 r;
        asm("v_lshlrev_b64 %0, %1, %2" : "=v"(r) : "i"(S), "v"(x));
        return r;
    } else if constexpr (S < 0 && S > -32) {
        u64 r;
        asm("v_lshrrev_b64 %0, %1, %2" : "=v"(r) : "i"(-S), "v"(x));
        return r;
    }
#endif
    return S > 0 ? x << S : x >> -S;
}
struct Prop3 {
    u64 a, b, c;
};
// propagators of a direction from the empty squares that may be entered
template <int S> PG_HD Prop3 props(u64 a) {
    Prop3 p;
    p.a = a;
    p.b = a & shift<S>(a);
    p.c = p.b & shift<2 * S>(p.b);
    return p;
}
// generators grown through the propagators (the ray squares shifted back
// by one step, still to be shifted by S and wrap-masked)
template <int S> PG_HD u64 grow(u64 g, const Prop3 &p) {
    g = or_and(g, shift<S>(g), p.a);
    g = or_and(g, shift<2 * S>(g), p.b);
    if constexpr (S == 8 || S == -8) {
        // a shift by 32 moves one half into the other: one 32-bit step
        u32 lo = (u32)g, hi = (u32)(g >> 32);
        if (S == 8) hi |= lo & (u32)(p.c >> 32);
        else lo |= hi & (u32)p.c;
        return ((u64)hi << 32) | lo;
    }
    return or_and(g, shift<4 * S>(g), p.c);
}
// the side's part of a piece word (every piece word lies inside occ)
template <int C> PG_HD u64 side_of(u64 b, u64 white) { return C == WHITE ? b & white : andn(b, white); }

// Attacks of the sliders in b (bishop-like) and r (rook-like) through the
// empty squares e, one pass over the eight directions with the propagators
// built once per direction; every slider must be off e.
PG_HD u64 ks_attacks(u64 b, u64 r, u64 e) {
    const Prop3 pn = props<8>(e), ps = props<-8>(e);
    u64 a = shift<8>(grow<8>(r, pn)) | shift<-8>(grow<-8>(r, ps));
    const u64 eA = andn(e, FILEA), eH = andn(e, FILEH);
    a = or_xor(a, eH, east_sum(r, eH));
    a = or_xor(a, eA, rev64(east_sum(rev64(r), rev64(eA))));
    const Prop3 p9 = props<9>(eA), m9 = props<-9>(eH), p7 = props<7>(eH), m7 = props<-7>(eA);
    a = or_and(a, shift<9>(grow<9>(b, p9)), ~FILEA);
    a = or_and(a, shift<-9>(grow<-9>(b, m9)), ~FILEH);
    a = or_and(a, shift<7>(grow<7>(b, p7)), ~FILEH);
    return or_and(a, shift<-7>(grow<-7>(b, m7)), ~FILEA);
}

// Attacks of one slider along one line through its square (the line may
// include the square): (o - s) ^ rev(rev(o) - rev(s)) with o the occupied
// squares of the line other than s (hyperbola quintessence, no tables).
PG_HD u64 line_att(u64 sb, u64 occ, u64 line) {
    const u64 o = bop3<0x40>(occ, line, sb);  // occ & line & ~sb
    return xor_and(o - sb, rev64(rev64(o) - rev64(sb)), line);
}

// pawn geometry per side
template <int C> PG_HD u64 fwd(u64 x) { return C == WHITE ? x << 8 : x >> 8; }
template <int C> PG_HD u64 fwd_west(u64 x) { return C == WHITE ? (x << 7) & ~FILEH : (x >> 9) & ~FILEH; }
template <int C> PG_HD u64 fwd_east(u64 x) { return C == WHITE ? (x << 9) & ~FILEA : (x >> 7) & ~FILEA; }
template <int C> PG_HD u64 pawn_att(u64 x) { return fwd_west<C>(x) | fwd_east<C>(x); }

// lines through a square
PG_HD u64 file_of(int s) { return FILEA << (s & 7); }
PG_HD u64 rank_of(int s) { return RANK1 << (s & 56); }
PG_HD u64 diag_of(int s) {  // a1-h8 direction
    int d = (s & 7) - (s >> 3);
    return d >= 0 ? DIAG_A1H8 >> (8 * d) : DIAG_A1H8 << (-8 * d);
}
PG_HD u64 anti_of(int s) {  // a8-h1 direction
    int d = (s & 7) + (s >> 3) - 7;
    return d >= 0 ? DIAG_A8H1 << (8 * d) : DIAG_A8H1 >> (-8 * d);
}

// ---------------------------------------------------------------- sliders
template <int V> struct Sl;

template <> struct Sl<SL_KOGGE> {
    static constexpr bool kBulk = true;
    PG_HD static u64 rook(const Tables &, int s, u64 occ) {
        const u64 sb = bit(s);
        return line_att(sb, occ, file_of(s)) | line_att(sb, occ, rank_of(s));
    }
    PG_HD static u64 bishop(const Tables &, int s, u64 occ) {
        const u64 sb = bit(s);
        return line_att(sb, occ, diag_of(s)) | line_att(sb, occ, anti_of(s));
    }
};

// high 32 bits of the low 64-bit word of a*b, i.e. (a*b) >> 32 mod 2^32:
// three 32-bit multiplies instead of a full 64x64 product.
PG_HD u32 mul_hi32_lo64(u64 a, u64 b) {
    u32 al = (u32)a, ah = (u32)(a >> 32), bl = (u32)b, bh = (u32)(b >> 32);
    u32 hi = (u32)(((u64)al * bl) >> 32);
    return hi + al * bh + ah * bl;
}

template <> struct Sl<SL_PLAIN> {
    static constexpr bool kBulk = false;
    PG_HD static u64 rook(const Tables &t, int s, u64 occ) {
        u32 i = mul_hi32_lo64(t.rookMagic[s], occ & t.rookMask[s]) >> (32 - ROOK_BITS);
        return t.rookPlain[(s << ROOK_BITS) + i];
    }
    PG_HD static u64 bishop(const Tables &t, int s, u64 occ) {
        u32 i = mul_hi32_lo64(t.bishopMagic[s], occ & t.bishopMask[s]) >> (32 - BISHOP_BITS);
        return t.bishopPlain[(s << BISHOP_BITS) + i];
    }
};

template <> struct Sl<SL_FANCY> {
    static constexpr bool kBulk = false;
    PG_HD static u64 rook(const Tables &t, int s, u64 occ) {
        FancyEntry f = t.rookFancy[s];
        return t.fancy[f.base + (mul_hi32_lo64(f.factor, occ & t.rookMask[s]) >> (32 - ROOK_BITS))];
    }
    PG_HD static u64 bishop(const Tables &t, int s, u64 occ) {
        FancyEntry f = t.bishopFancy[s];
        return t.fancy[f.base + (mul_hi32_lo64(f.factor, occ & t.bishopMask[s]) >> (32 - BISHOP_BITS))];
    }
};

template <> struct Sl<SL_BYTE> {
    static constexpr bool kBulk = false;
    PG_HD static u64 rook(const Tables &t, int s, u64 occ) {
        FancyEntry f = t.rookFancy[s];
        u32 i = mul_hi32_lo64(f.factor, occ & t.rookMask[s]) >> (32 - ROOK_BITS);
        return t.byteRook[t.byteIdx[f.base + i] + f.offset];
    }
    PG_HD static u64 bishop(const Tables &t, int s, u64 occ) {
        FancyEntry f = t.bishopFancy[s];
        u32 i = mul_hi32_lo64(f.factor, occ & t.bishopMask[s]) >> (32 - BISHOP_BITS);
        return t.byteBishop[t.byteIdx[f.base + i] + f.offset];
    }
};

// ---------------------------------------------------------------- context
struct MoveCtx {
    u64 occ, mine, enemy, myPawns, myKing, eRooks, pinned, checkers, threatened, target;
    int k;
};

template <int C, int V>
PG_HD void prologue(const Pos &p, const Tables &t, MoveCtx &x) {
    const u64 pawns = pos_pawns(p);
    x.occ = p.kings | pawns | p.knights | p.bishopQueens | p.rookQueens;
    const u64 black = x.occ & ~p.white;
    x.mine = C == WHITE ? p.white : black;
    x.enemy = C == WHITE ? black : p.white;
    x.myPawns = pawns & x.mine;
    x.myKing = p.kings & x.mine;
    x.k = ctz(x.myKing) & 63;  // in range even for a board without a king
    const u64 eP = pawns & x.enemy, eN = p.knights & x.enemy, eK = p.kings & x.enemy;
    const u64 eB = p.bishopQueens & x.enemy, eR = p.rookQueens & x.enemy;
    x.eRooks = eR;

    // squares the enemy attacks, sliders looking through our king
    u64 a;
    {
        const u64 pw = C == WHITE ? shift<-9>(eP) : shift<7>(eP), pe = C == WHITE ? shift<-7>(eP) : shift<9>(eP);
        a = or3(t.king[ctz(eK) & 63], andn(pw, FILEH), andn(pe, FILEA));
        const u64 h1 = or_and(andn(shift<-1>(eN), FILEH), shift<1>(eN), ~FILEA);
        const u64 h2 = or_and(andn(shift<-2>(eN), FILEG | FILEH), shift<2>(eN), ~(FILEA | FILEB));
        a = or3(a, shift<16>(h1), shift<-16>(h1));
        a = or3(a, shift<8>(h2), shift<-8>(h2));
    }
    if (Sl<V>::kBulk) {
        a |= ks_attacks(eB, eR, ~x.occ | x.myKing);
    } else {
        const u64 o = x.occ ^ x.myKing;
        for (u64 b = eB; b; b &= b - 1) a |= Sl<V>::bishop(t, ctz(b), o);
        for (u64 r = eR; r; r &= r - 1) a |= Sl<V>::rook(t, ctz(r), o);
    }
    x.threatened = a;

    // sliders on a king line: no blocker -> checker, one blocker -> pinned
    u64 cand = (t.rookEmpty[x.k] & eR) | (t.bishopEmpty[x.k] & eB);
    u64 pinned = 0, chk = 0;
    for (; cand; cand &= cand - 1) {
        const int s = ctz(cand);
        const u64 b = t.between[(x.k << 6) | s] & x.occ;
        if (b == 0) chk |= bit(s);
        else if (!several(b)) pinned |= b;
    }
    chk |= (t.knight[x.k] & eN) | (pawn_att<C>(x.myKing) & eP);
    x.pinned = pinned & x.mine;
    x.checkers = chk;
    u64 tgt = ~x.mine;
    if (chk) tgt &= several(chk) ? 0ULL : (chk | t.between[(x.k << 6) | ctz(chk)]);
    x.target = tgt;
}

enum { P_PAWN = 1, P_KNIGHT, P_BISHOP, P_ROOK, P_QUEEN, P_KING };

// Bounded move writer: never stores past `cap`, but keeps counting so a
// caller can detect a mismatch with the counted size.  With `tag` set it also
// stores owner | moving piece << 8 per move (the fused leaf's window), so the
// child can be made without looking the piece up.
struct MoveSink {
    u16 *out;
    u32 cap, n;
    u16 *tag;
    u32 owner;
    PG_HD void add(u16 m, int piece) {
        if (n < cap) {
            out[n] = m;
            if (tag) tag[n] = (u16)(owner | ((u32)piece << 8));
        }
        n++;
    }
};

// en-passant: number of legal captures (MoveGeneratorBitboard.h:3050-3085,
// and :2891-2902 when in check); writes them when `out` is given.
template <int C>
PG_HD int ep_moves(const Pos &p, const Tables &t, const MoveCtx &x, MoveSink *out) {
    const int f = pos_ep(p);
    if (!f) return 0;
    const u64 dst = C == WHITE ? bit(39 + f) : bit(15 + f);
    const u64 cap = C == WHITE ? dst >> 8 : dst << 8;
    u64 src = (((cap << 1) & ~FILEA) | ((cap >> 1) & ~FILEH)) & x.myPawns;
    int n = 0;
    const int to = ctz(dst);
    if (x.checkers) {
        if (x.checkers != cap) return 0;
        src &= ~x.pinned;
        for (; src; src &= src - 1) {
            if (out) out->add((u16)(ctz(src) | (to << 6) | (MF_EP_CAPTURE << 12)), P_PAWN);
            n++;
        }
        return n;
    }
    for (; src; src &= src - 1) {
        const int s = ctz(src);
        bool ok;
        if (x.pinned & bit(s)) {
            ok = (dst & t.line[(s << 6) | x.k]) != 0;
        } else {
            // both pawns leave the rank: no rook or queen may then see the king along it
            const u64 e = ~x.occ | cap | bit(s), eH = andn(e, FILEH), eA = andn(e, FILEA);
            const u64 ray = (eH ^ east_sum(x.eRooks, eH)) | (eA ^ rev64(east_sum(rev64(x.eRooks), rev64(eA))));
            ok = (ray & x.myKing) == 0;
        }
        if (ok) {
            if (out) out->add((u16)(s | (to << 6) | (MF_EP_CAPTURE << 12)), P_PAWN);
            n++;
        }
    }
    return n;
}

template <int C>
PG_HD int castle_count(const Pos &p, const MoveCtx &x) {
    if (x.checkers) return 0;
    const int r = C == WHITE ? pos_wcastle(p) : pos_bcastle(p);
    const int sh = C == WHITE ? 0 : 56;
    int n = 0;
    if ((r & 1) && !((0x60ULL << sh) & (x.occ | x.threatened))) n++;
    if ((r & 2) && !((0x0EULL << sh) & x.occ) && !((0x0CULL << sh) & x.threatened)) n++;
    return n;
}

// ---------------------------------------------------------------- count
template <int C> PG_HD u32 count_moves_ks(const Pos &p, const Tables &t);

template <int C, int V>
PG_HD u32 count_moves(const Pos &p, const Tables &t) {
    if constexpr (V == SL_KOGGE) return count_moves_ks<C>(p, t);
    MoveCtx x;
    prologue<C, V>(p, t, x);
    const u64 tgt = x.target, empty = ~x.occ;
    int n = popc(and_nor(t.king[x.k], x.mine, x.threatened));
    if (tgt == 0) return n;  // double check: king moves only

    // movers; a pinned piece moves only along its pin line
    u64 pf = x.myPawns, pw = x.myPawns, pe = x.myPawns, myN = p.knights & x.mine;
    if (x.pinned) {
        const u64 np = ~x.pinned, dk = diag_of(x.k), ak = anti_of(x.k);
        pf = and_or(x.myPawns, np, file_of(x.k));
        pw = and_or(x.myPawns, np, C == WHITE ? ak : dk);
        pe = and_or(x.myPawns, np, C == WHITE ? dk : ak);
        myN = andn(myN, x.pinned);
    }

    // pawns
    const u64 PROMO = C == WHITE ? RANK8 : RANK1;
    const u64 p1 = fwd<C>(pf) & empty;
    const u64 p2 = fwd<C>(p1 & (C == WHITE ? RANK3 : RANK6)) & empty & tgt;
    const u64 t1 = p1 & tgt;
    n += popc(t1) + 3 * popc(t1 & PROMO) + popc(p2);
    const u64 et = x.enemy & tgt;
    const u64 cw = and3(C == WHITE ? shift<7>(pw) : shift<-9>(pw), ~FILEH, et);
    const u64 ce = and3(C == WHITE ? shift<9>(pe) : shift<-7>(pe), ~FILEA, et);
    n += popc(cw) + 3 * popc(cw & PROMO) + popc(ce) + 3 * popc(ce & PROMO);

    // knights (pinned knights never move)
    for (u64 b = myN; b; b &= b - 1) n += popc(t.knight[ctz(b)] & tgt);

    // sliders, per piece
    for (u64 b = p.bishopQueens & x.mine; b; b &= b - 1) {
        const int s = ctz(b);
        u64 m = Sl<V>::bishop(t, s, x.occ) & tgt;
        if (x.pinned & bit(s)) m &= t.line[(s << 6) | x.k];
        n += popc(m);
    }
    for (u64 r = p.rookQueens & x.mine; r; r &= r - 1) {
        const int s = ctz(r);
        u64 m = Sl<V>::rook(t, s, x.occ) & tgt;
        if (x.pinned & bit(s)) m &= t.line[(s << 6) | x.k];
        n += popc(m);
    }
    n += castle_count<C>(p, x);
    n += ep_moves<C>(p, t, x, nullptr);
    return (u32)n;
}

// ---------------------------------------------------------------- count, Kogge-Stone
// The Kogge-Stone count in one pass over the eight directions, written to
// keep the VALU instruction count down:
//  * and/or chains are three-input bitop3 instructions (two per 64-bit word);
//  * each direction's propagator chain (empty, empty & empty<<s, ...) is
//    built once and shared by the enemy-attack fill and the own-move fill;
//  * enemy sliders fill through empty squares only; the one square the
//    moving king must not step to because a slider checks it through the
//    king ("behind" it on the check line) is added explicitly;
//  * east rays of every slider at once by one add: with the h-file as
//    sentinel, o = occ | FILEH and e = ~o, the squares from each slider up
//    to its first blocker are o ^ (o - 2s) = e ^ (e + 2s); west rays the
//    same on the bit-reversed board;
//  * pinned pieces need the king-line masks only when something is pinned.
template <int C>
PG_HD u32 count_moves_ks(const Pos &p, const Tables &t) {
    MoveCtx x;
    const u64 W = p.white;
    const u64 pawns = pos_pawns(p);
    x.occ = or3(p.kings, pawns, or3(p.knights, p.bishopQueens, p.rookQueens));
    x.mine = C == WHITE ? W : andn(x.occ, W);
    x.enemy = C == WHITE ? andn(x.occ, W) : W;
    x.myPawns = side_of<C>(pawns, W);
    x.myKing = side_of<C>(p.kings, W);
    x.k = ctz(x.myKing) & 63;
    const int k = x.k;
    const u64 eP = side_of<1 - C>(pawns, W), eN = side_of<1 - C>(p.knights, W);
    const u64 eB = side_of<1 - C>(p.bishopQueens, W), eR = side_of<1 - C>(p.rookQueens, W);
    x.eRooks = eR;

    // sliders on a king line: no blocker -> checker, one blocker -> pinned
    u64 cand = or_and(t.rookEmpty[k] & eR, t.bishopEmpty[k], eB);
    u64 pinned = 0, chk = 0, behind = 0;
    for (; cand; cand &= cand - 1) {
        const int s = ctz(cand);
        const u64 btw = t.between[(k << 6) | s];
        const u64 b = btw & x.occ;
        if (b == 0) {
            chk |= bit(s);
            behind |= and_nor(t.line[(s << 6) | k] & t.king[k], btw, bit(s));
        } else if (!several(b)) {
            pinned |= b;
        }
    }
    {   // pawn and knight checks
        const u64 kw = C == WHITE ? shift<7>(x.myKing) : shift<-9>(x.myKing);
        const u64 ke = C == WHITE ? shift<9>(x.myKing) : shift<-7>(x.myKing);
        chk = or_and(or_and(chk, or_and(andn(kw, FILEH), ke, ~FILEA), eP), t.knight[k], eN);
    }
    x.pinned = pinned & x.mine;
    x.checkers = chk;
    u64 tgt = ~x.mine;
    if (chk) tgt &= several(chk) ? 0ULL : (chk | t.between[(k << 6) | ctz(chk)]);
    x.target = tgt;

    // movers; a pinned piece moves only along its pin line
    const u64 rq = side_of<C>(p.rookQueens, W), bq = side_of<C>(p.bishopQueens, W);
    u64 rf = rq, rr = rq, bd = bq, ba = bq, pf = x.myPawns, pw = x.myPawns, pe = x.myPawns;
    u64 myN = side_of<C>(p.knights, W);
    if (x.pinned) {
        const u64 np = ~x.pinned;
        const u64 fk = file_of(k), dk = diag_of(k), ak = anti_of(k);
        rf = and_or(rq, np, fk);
        rr = and_or(rq, np, rank_of(k));
        bd = and_or(bq, np, dk);
        ba = and_or(bq, np, ak);
        pf = and_or(x.myPawns, np, fk);
        pw = and_or(x.myPawns, np, C == WHITE ? ak : dk);
        pe = and_or(x.myPawns, np, C == WHITE ? dk : ak);
        myN = andn(myN, x.pinned);
    }

    // enemy attacks without sliders: king, pawns, knights
    const u64 eK = side_of<1 - C>(p.kings, W);
    u64 att;
    {
        const u64 pw_ = C == WHITE ? shift<-9>(eP) : shift<7>(eP), pe_ = C == WHITE ? shift<-7>(eP) : shift<9>(eP);
        att = or3(t.king[ctz(eK) & 63], or_and(andn(pw_, FILEH), pe_, ~FILEA), behind);
        const u64 h1 = or_and(andn(shift<-1>(eN), FILEH), shift<1>(eN), ~FILEA);
        const u64 h2 = or_and(andn(shift<-2>(eN), FILEG | FILEH), shift<2>(eN), ~(FILEA | FILEB));
        att = or3(att, shift<16>(h1), shift<-16>(h1));
        att = or3(att, shift<8>(h2), shift<-8>(h2));
    }

    const u64 empty = ~x.occ;
    int n = 0;
    {   // north and south
        const Prop3 pn = props<8>(empty), ps = props<-8>(empty);
        const u64 an = grow<8>(eR, pn), as = grow<-8>(eR, ps);
        att = or3(att, shift<8>(an), shift<-8>(as));
        n += popc((grow<8>(rf, pn) << 8) & tgt);
        n += popc((grow<-8>(rf, ps) >> 8) & tgt);
    }
    const u64 eA = nor2(x.occ, FILEA), eH = nor2(x.occ, FILEH);
    {   // east: one add per side
        att = or_xor(att, eH, east_sum(eR, eH));
        n += popc(xor_and(eH, east_sum(rr, eH), tgt));
    }
    {   // west: east on the bit-reversed board
        const u64 eoR = rev64(eA);
        att = or_xor(att, eA, rev64(east_sum(rev64(eR), eoR)));
        n += popc(xor_and(eA, rev64(east_sum(rev64(rr), eoR)), tgt));
    }
    {   // north-east and south-west
        const Prop3 pa = props<9>(eA), pb = props<-9>(eH);
        att = or_and(att, shift<9>(grow<9>(eB, pa)), ~FILEA);
        att = or_and(att, shift<-9>(grow<-9>(eB, pb)), ~FILEH);
        n += popc(and3(shift<9>(grow<9>(bd, pa)), ~FILEA, tgt));
        n += popc(and3(shift<-9>(grow<-9>(bd, pb)), ~FILEH, tgt));
    }
    {   // north-west and south-east
        const Prop3 pa = props<7>(eH), pb = props<-7>(eA);
        att = or_and(att, shift<7>(grow<7>(eB, pa)), ~FILEH);
        att = or_and(att, shift<-7>(grow<-7>(eB, pb)), ~FILEA);
        n += popc(and3(shift<7>(grow<7>(ba, pa)), ~FILEH, tgt));
        n += popc(and3(shift<-7>(grow<-7>(ba, pb)), ~FILEA, tgt));
    }
    x.threatened = att;
    n += popc(and_nor(t.king[k], x.mine, att));

    // pawns: pushes on the low or high half only where the rank allows
    const u64 PROMO = C == WHITE ? RANK8 : RANK1;
    const u64 p1 = fwd<C>(pf) & empty;
    const u64 p2 = fwd<C>(p1 & (C == WHITE ? RANK3 : RANK6)) & empty & tgt;
    const u64 t1 = p1 & tgt;
    n += popc(t1) + 3 * popc(t1 & PROMO) + popc(p2);
    const u64 et = x.enemy & tgt;
    const u64 cw = and3(C == WHITE ? shift<7>(pw) : shift<-9>(pw), ~FILEH, et);
    const u64 ce = and3(C == WHITE ? shift<9>(pe) : shift<-7>(pe), ~FILEA, et);
    n += popc(cw) + 3 * popc(cw & PROMO) + popc(ce) + 3 * popc(ce & PROMO);
    // knights
    for (u64 b = myN; b; b &= b - 1) n += popc(t.knight[ctz(b)] & tgt);
    n += castle_count<C>(p, x);
    n += ep_moves<C>(p, t, x, nullptr);
    return (u32)n;
}

// ---------------------------------------------------------------- generate
PG_HD u16 mk_move(int from, int to, int flags) { return (u16)(from | (to << 6) | (flags << 12)); }

// calls f(square) for every set square of m, low half then high half (a
// 32-bit find-first-set and clear per square instead of the 64-bit ones)
template <class F>
PG_HD void each_square(u64 m, F &&f) {
    for (u32 lo = (u32)m; lo; lo &= lo - 1) f(__builtin_ctz(lo));
    for (u32 hi = (u32)(m >> 32); hi; hi &= hi - 1) f(32 + __builtin_ctz(hi));
}

// Writes the legal moves of p to out (at most cap of them); returns the
// number of legal moves (at most 218).
// Pawn captures carry the capture flag; other captures do not, as on the
// reference's device path (MoveGeneratorBitboard.h:2525-2529).
template <int C, int V>
PG_HD u32 generate_moves(const Pos &p, const Tables &t, u16 *out, u32 cap, u16 *tag = nullptr, u32 owner = 0) {
    MoveCtx x;
    prologue<C, V>(p, t, x);
    const u64 tgt = x.target, empty = ~x.occ;
    MoveSink o{out, cap, 0, tag, owner};
    each_square(and_nor(t.king[x.k], x.mine, x.threatened),
                [&](int to) { o.add(mk_move(x.k, to, MF_QUIET), P_KING); });
    if (tgt == 0) return o.n;

    // a pinned pawn moves only along its pin line
    u64 pf = x.myPawns, pw = x.myPawns, pe = x.myPawns;
    if (x.pinned) {
        const u64 np = ~x.pinned, dk = diag_of(x.k), ak = anti_of(x.k);
        pf = and_or(x.myPawns, np, file_of(x.k));
        pw = and_or(x.myPawns, np, C == WHITE ? ak : dk);
        pe = and_or(x.myPawns, np, C == WHITE ? dk : ak);
    }
    const int up = C == WHITE ? 8 : -8;
    const u64 PROMO = C == WHITE ? RANK8 : RANK1;
    auto promos = [&](int from, int to, int cap) {
        o.add(mk_move(from, to, MF_PROMO | cap | 0), P_PAWN);
        o.add(mk_move(from, to, MF_PROMO | cap | 1), P_PAWN);
        o.add(mk_move(from, to, MF_PROMO | cap | 3), P_PAWN);
        o.add(mk_move(from, to, MF_PROMO | cap | 2), P_PAWN);
    };
    // pushes
    const u64 p1 = fwd<C>(pf) & empty;
    const u64 p2 = fwd<C>(p1 & (C == WHITE ? RANK3 : RANK6)) & empty & tgt;
    each_square(andn(p1 & tgt, PROMO), [&](int to) { o.add(mk_move(to - up, to, MF_QUIET), P_PAWN); });
    each_square(p1 & tgt & PROMO, [&](int to) { promos(to - up, to, 0); });
    each_square(p2, [&](int to) { o.add(mk_move(to - 2 * up, to, MF_DOUBLE_PUSH), P_PAWN); });
    // captures
    const u64 et = x.enemy & tgt;
    const u64 cw = and3(C == WHITE ? shift<7>(pw) : shift<-9>(pw), ~FILEH, et);
    const u64 ce = and3(C == WHITE ? shift<9>(pe) : shift<-7>(pe), ~FILEA, et);
    constexpr int dw = C == WHITE ? 7 : -9, de = C == WHITE ? 9 : -7;
    each_square(andn(cw, PROMO), [&](int to) { o.add(mk_move(to - dw, to, MF_CAPTURE), P_PAWN); });
    each_square(andn(ce, PROMO), [&](int to) { o.add(mk_move(to - de, to, MF_CAPTURE), P_PAWN); });
    each_square(cw & PROMO, [&](int to) { promos(to - dw, to, MF_CAPTURE); });
    each_square(ce & PROMO, [&](int to) { promos(to - de, to, MF_CAPTURE); });
    ep_moves<C>(p, t, x, &o);
    // castling
    if (!x.checkers) {
        const int r = C == WHITE ? pos_wcastle(p) : pos_bcastle(p);
        const int sh = C == WHITE ? 0 : 56;
        if ((r & 1) && !((0x60ULL << sh) & (x.occ | x.threatened))) o.add(mk_move(4 + sh, 6 + sh, MF_KING_CASTLE), P_KING);
        if ((r & 2) && !((0x0EULL << sh) & x.occ) && !((0x0CULL << sh) & x.threatened)) o.add(mk_move(4 + sh, 2 + sh, MF_QUEEN_CASTLE), P_KING);
    }
    // knights (pinned knights never move)
    for (u64 b = andn(p.knights & x.mine, x.pinned); b; b &= b - 1) {
        const int s = ctz(b);
        each_square(t.knight[s] & tgt, [&](int to) { o.add(mk_move(s, to, MF_QUIET), P_KNIGHT); });
    }
    // sliders
    for (u64 b = p.bishopQueens & x.mine; b; b &= b - 1) {
        const int s = ctz(b);
        u64 m = Sl<V>::bishop(t, s, x.occ) & tgt;
        if (x.pinned & bit(s)) m &= t.line[(s << 6) | x.k];
        const int pc = (p.rookQueens & bit(s)) ? P_QUEEN : P_BISHOP;
        each_square(m, [&](int to) { o.add(mk_move(s, to, MF_QUIET), pc); });
    }
    for (u64 r = p.rookQueens & x.mine; r; r &= r - 1) {
        const int s = ctz(r);
        u64 m = Sl<V>::rook(t, s, x.occ) & tgt;
        if (x.pinned & bit(s)) m &= t.line[(s << 6) | x.k];
        const int pc = (p.bishopQueens & bit(s)) ? P_QUEEN : P_ROOK;
        each_square(m, [&](int to) { o.add(mk_move(s, to, MF_QUIET), pc); });
    }
    return o.n;
}

// ---------------------------------------------------------------- make move
// Zobrist word offsets in the 782-word key set (chess.h:326-333)
PG_HD int zob_piece(int color, int piece /*1..6*/, int sq) { return color * 384 + (piece - 1) * 64 + sq; }
constexpr int ZOB_CASTLE = 768, ZOB_EP = 772, ZOB_SIDE = 780;

PG_HD void zob_xor(Key128 &h, const Tables &t, int w) {
    h.lo ^= t.zobLo[w];
    h.hi ^= t.zobHi[w];
}
PG_HD void zob_state(Key128 &h, const Tables &t, const Pos &p) {
    const int w = pos_wcastle(p), b = pos_bcastle(p), e = pos_ep(p);
    if (w & 1) zob_xor(h, t, ZOB_CASTLE + 0);
    if (w & 2) zob_xor(h, t, ZOB_CASTLE + 1);
    if (b & 1) zob_xor(h, t, ZOB_CASTLE + 2);
    if (b & 2) zob_xor(h, t, ZOB_CASTLE + 3);
    if (e) zob_xor(h, t, ZOB_EP + e - 1);
}

PG_HD int piece_on(const Pos &p, u64 b) {
    if (p.kings & b) return P_KING;
    if (p.knights & b) return P_KNIGHT;
    if (pos_pawns(p) & b) return P_PAWN;
    if (p.bishopQueens & p.rookQueens & b) return P_QUEEN;
    if (p.bishopQueens & b) return P_BISHOP;
    if (p.rookQueens & b) return P_ROOK;
    return 0;
}

// Applies a legal move (MoveGeneratorBitboard.h:3575-3836): the en-passant
// file is recorded only when an enemy pawn can take, and castling rights are
// cleared by king moves, rook moves from a corner and captures on a corner.
// With HASH the 128-bit key is updated incrementally.
// make_move_as takes the moving piece from the caller (the generator records
// it); make_move looks it up.
template <int C, bool HASH>
PG_HD void make_move_as(Pos &p, u16 mv, int piece, Key128 *h, const Tables &t) {
    const int from = mv & 63, to = (mv >> 6) & 63, fl = mv >> 12;
    const u64 src = bit(from), dst = bit(to);
    if (HASH) {
        zob_xor(*h, t, zob_piece(C, piece, from));
        const int cap = piece_on(p, dst);
        if (cap) zob_xor(*h, t, zob_piece(1 - C, cap, to));
        zob_xor(*h, t, ZOB_SIDE);
        zob_state(*h, t, p);
    }
    const bool movedRook = piece == P_ROOK;
    if (fl & MF_PROMO) {
        const int pr = fl & 3;  // N, B, R, Q
        piece = pr == 0 ? P_KNIGHT : pr == 1 ? P_BISHOP : pr == 2 ? P_ROOK : P_QUEEN;
    }
    if (HASH) zob_xor(*h, t, zob_piece(C, piece, to));
    const u64 clr = ~(src | dst);
    p.bishopQueens &= clr;
    p.rookQueens &= clr;
    p.kings &= clr;
    p.knights &= clr;
    p.pawns &= ~((src | dst) & RANKS2TO7);
    if (piece == P_KING) p.kings |= dst;
    if (piece == P_KNIGHT) p.knights |= dst;
    if (piece == P_PAWN) p.pawns |= dst;
    if (piece == P_BISHOP || piece == P_QUEEN) p.bishopQueens |= dst;
    if (piece == P_ROOK || piece == P_QUEEN) p.rookQueens |= dst;
    if (C == WHITE) p.white = (p.white ^ src) | dst;
    else p.white &= ~dst;

    if (fl == MF_EP_CAPTURE) {
        const u64 cap = C == WHITE ? dst >> 8 : dst << 8;
        p.pawns &= ~cap;
        if (C == BLACK) p.white &= ~cap;
        if (HASH) zob_xor(*h, t, zob_piece(1 - C, P_PAWN, ctz(cap)));
    }
    if (fl == MF_KING_CASTLE || fl == MF_QUEEN_CASTLE) {
        const int sh = C == WHITE ? 0 : 56;
        const int rf = (fl == MF_KING_CASTLE ? 7 : 0) + sh, rt = (fl == MF_KING_CASTLE ? 5 : 3) + sh;
        p.rookQueens = (p.rookQueens ^ bit(rf)) | bit(rt);
        if (C == WHITE) p.white = (p.white ^ bit(rf)) | bit(rt);
        if (HASH) {
            zob_xor(*h, t, zob_piece(C, P_ROOK, rf));
            zob_xor(*h, t, zob_piece(C, P_ROOK, rt));
        }
    }

    // game state: side, en passant, castling
    u64 st = p.pawns & ~(0xFFULL | (1ULL << 63));
    int wc = pos_wcastle(p), bc = pos_bcastle(p);
    if (piece == P_KING) { if (C == WHITE) wc = 0; else bc = 0; }
    // a capture on the enemy's corner (updateCastleFlag, :1311)
    if (C == WHITE) {
        if (dst & bit(63)) bc &= ~1;
        else if (dst & bit(56)) bc &= ~2;
    } else {
        if (dst & bit(7)) wc &= ~1;
        else if (dst & bit(0)) wc &= ~2;
    }
    if (movedRook && piece == P_ROOK) {
        if (C == WHITE) {
            if (src & bit(7)) wc &= ~1;
            else if (src & bit(0)) wc &= ~2;
        } else {
            if (src & bit(63)) bc &= ~1;
            else if (src & bit(56)) bc &= ~2;
        }
    }
    int epf = 0;
    if (fl == MF_DOUBLE_PUSH) {
        const u64 ePawns = (C == WHITE) ? (p.pawns & RANKS2TO7 & ~p.white) : (p.pawns & RANKS2TO7 & p.white);
        if ((((dst << 1) & ~FILEA) | ((dst >> 1) & ~FILEH)) & ePawns) epf = (from & 7) + 1;
    }
    p.pawns = st | (u64)wc | ((u64)bc << 2) | ((u64)epf << 4) | ((u64)(1 - C) << 63);
    if (HASH) zob_state(*h, t, p);
}

template <int C, bool HASH>
PG_HD void make_move(Pos &p, u16 mv, Key128 *h, const Tables &t) {
    int piece = piece_on(p, bit(mv & 63));
    if (!piece) piece = P_PAWN;  // only for a move that is not legal here; keeps key indices in range
    make_move_as<C, HASH>(p, mv, piece, h, t);
}

// Full 128-bit key (MoveGeneratorBitboard.h:3914-3980); the side key is
// present when white is to move.
PG_HD Key128 zobrist128(const Pos &p, const Tables &t) {
    Key128 h = {0, 0};
    if (pos_side(p) == WHITE) zob_xor(h, t, ZOB_SIDE);
    zob_state(h, t, p);
    for (u64 all = pos_occ(p); all; all &= all - 1) {
        const int s = ctz(all);
        const int color = (p.white >> s) & 1 ? WHITE : BLACK;
        zob_xor(h, t, zob_piece(color, piece_on(p, bit(s)), s));
    }
    return h;
}

// side dispatch helpers
template <int V>
PG_HD u32 count_any(const Pos &p, const Tables &t) {
    return pos_side(p) ? count_moves<BLACK, V>(p, t) : count_moves<WHITE, V>(p, t);
}
template <int V>
PG_HD u32 generate_any(const Pos &p, const Tables &t, u16 *out, u32 cap = 256, u16 *tag = nullptr, u32 owner = 0) {
    return pos_side(p) ? generate_moves<BLACK, V>(p, t, out, cap, tag, owner)
                       : generate_moves<WHITE, V>(p, t, out, cap, tag, owner);
}
PG_HD void make_any(Pos &p, u16 mv, const Tables &t) {
    if (pos_side(p)) make_move<BLACK, false>(p, mv, nullptr, t);
    else make_move<WHITE, false>(p, mv, nullptr, t);
}
PG_HD void make_any_hash(Pos &p, u16 mv, Key128 &h, const Tables &t) {
    if (pos_side(p)) make_move<BLACK, true>(p, mv, &h, t);
    else make_move<WHITE, true>(p, mv, &h, t);
}

}  // namespace pg
