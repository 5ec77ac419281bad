// tables.cpp -- builds the move generator's lookup tables on the host.
//
// Same tables as MoveGeneratorBitboard::init (MoveGeneratorBitboard.h:982-
// 1080) and Magics.cpp, constructed by this code:
//  * empty-board attacks, squares-between and full-line tables;
//  * relevant-occupancy masks (edges removed, :1026-1047);
//  * plain magics: 12-bit rook / 9-bit bishop tables found by a seeded search
//    (the reference seeds from the clock, :1025; a fixed seed keeps builds
//    reproducible);
//  * fixed-shift fancy magics from the published factors and bases, with the
//    collision check of Magics.cpp:62-106;
//  * byte-lookup variant: per-square list of distinct attack sets and a byte
//    index per fancy slot (Magics.cpp:69-92 layout; offsets accumulate in
//    square order, :1059-1076).
#include "tables.h"
#include <cstring>
#include <cstdio>
#include <mutex>
#include "zobrist_keys.inc"
#include "fancy_magics.inc"

namespace pg {

static u64 between_sq(int a, int b) {
    int fa = a & 7, ra = a >> 3, fb = b & 7, rb = b >> 3;
    int df = fb - fa, dr = rb - ra;
    if (a == b) return 0;
    if (!(df == 0 || dr == 0 || df == dr || df == -dr)) return 0;
    int sf = (df > 0) - (df < 0), sr = (dr > 0) - (dr < 0);
    u64 m = 0;
    for (int f = fa + sf, r = ra + sr; f != fb || r != rb; f += sf, r += sr) m |= bit(r * 8 + f);
    return m;
}

static u64 line_sq(int a, int b) {
    if (file_of(a) & bit(b)) return file_of(a);
    if (rank_of(a) & bit(b)) return rank_of(a);
    if (diag_of(a) & bit(b)) return diag_of(a);
    if (anti_of(a) & bit(b)) return anti_of(a);
    return 0;
}

// splitmix64
static u64 next_rand(u64 &s) {
    u64 z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// enumerate all subsets of mask (carry-rippler); returns count
static int subsets(u64 mask, u64 *occ, u64 *att, int sq, bool rook) {
    int n = 0;
    u64 s = 0;
    do {
        occ[n] = s;
        att[n] = rook ? ks_rook(bit(sq), ~s) : ks_bishop(bit(sq), ~s);
        n++;
        s = (s - mask) & mask;
    } while (s);
    return n;
}

static u64 find_plain_magic(u64 mask, int bits, const u64 *occ, const u64 *att, int n, u64 *table, u64 &rng) {
    std::vector<u32> epoch(1u << bits, 0);
    for (u32 tries = 1;; tries++) {
        u64 m = next_rand(rng) & next_rand(rng) & next_rand(rng);
        if (popc((mask * m) & 0xFF00000000000000ULL) < 6) continue;
        bool ok = true;
        for (int i = 0; i < n && ok; i++) {
            u32 idx = (u32)((occ[i] * m) >> (64 - bits));
            if (epoch[idx] != tries) {
                epoch[idx] = tries;
                table[idx] = att[i];
            } else if (table[idx] != att[i]) {
                ok = false;
            }
        }
        if (ok) {
            for (u32 i = 0; i < (1u << bits); i++)
                if (epoch[i] != tries) table[i] = 0;
            return m;
        }
    }
}

int HostTables::build(u64 seed, char *err, int errlen) {
    magicSeed = seed;
    for (int s = 0; s < 64; s++) {
        knight[s] = knight_att(bit(s));
        king[s] = king_att(bit(s));
        rookEmpty[s] = ks_rook(bit(s), ~0ULL);
        bishopEmpty[s] = ks_bishop(bit(s), ~0ULL);
    }
    for (int a = 0; a < 64; a++)
        for (int b = 0; b < 64; b++) {
            between[a * 64 + b] = between_sq(a, b);
            line[a * 64 + b] = a == b ? file_of(a) : line_sq(a, b);
        }
    for (int s = 0; s < 64; s++) {
        u64 sq = bit(s), m = rookEmpty[s] & ~sq;
        if (!(sq & RANK1)) m &= ~RANK1;
        if (!(sq & RANK8)) m &= ~RANK8;
        if (!(sq & FILEA)) m &= ~FILEA;
        if (!(sq & FILEH)) m &= ~FILEH;
        rookMask[s] = m;
        bishopMask[s] = bishopEmpty[s] & ~sq & 0x007E7E7E7E7E7E00ULL;
    }

    std::vector<u64> occ(4096), att(4096);
    rookPlain.assign(64u << ROOK_BITS, 0);
    bishopPlain.assign(64u << BISHOP_BITS, 0);
    u64 rng = seed;
    for (int s = 0; s < 64; s++) {
        int n = subsets(rookMask[s], occ.data(), att.data(), s, true);
        rookMagic[s] = find_plain_magic(rookMask[s], ROOK_BITS, occ.data(), att.data(), n, &rookPlain[(size_t)s << ROOK_BITS], rng);
        n = subsets(bishopMask[s], occ.data(), att.data(), s, false);
        bishopMagic[s] = find_plain_magic(bishopMask[s], BISHOP_BITS, occ.data(), att.data(), n, &bishopPlain[(size_t)s << BISHOP_BITS], rng);
    }

    fancy.assign(FANCY_TABLE_SIZE, 0);
    byteIdx.assign(FANCY_TABLE_SIZE, 0);
    byteRook.assign(BYTE_ROOK_SIZE, 0);
    byteBishop.assign(BYTE_BISHOP_SIZE, 0);
    std::vector<short> owner(FANCY_TABLE_SIZE, -1);
    int offR = 0, offB = 0;
    for (int s = 0; s < 64; s++) {
        for (int pass = 0; pass < 2; pass++) {
            const bool rook = pass == 0;
            const PgFancyConst &c = rook ? PG_FANCY_ROOK[s] : PG_FANCY_BISHOP[s];
            const int bits = rook ? ROOK_BITS : BISHOP_BITS;
            const short who = (short)(s + (rook ? 64 : 0));
            int n = subsets(rook ? rookMask[s] : bishopMask[s], occ.data(), att.data(), s, rook);
            std::vector<u64> uniq;
            for (int i = 0; i < n; i++) {
                const int slot = c.base + (int)((occ[i] * c.factor) >> (64 - bits));
                if (slot < 0 || slot >= FANCY_TABLE_SIZE) {
                    snprintf(err, errlen, "fancy slot out of range (sq %d)", s);
                    return -1;
                }
                if (fancy[slot] == 0) {
                    fancy[slot] = att[i];
                    owner[slot] = who;
                    size_t j = 0;
                    while (j < uniq.size() && uniq[j] != att[i]) j++;
                    if (j == uniq.size()) uniq.push_back(att[i]);
                    if (j > 255) {
                        snprintf(err, errlen, "more than 256 attack sets (sq %d)", s);
                        return -1;
                    }
                    byteIdx[slot] = (u8)j;
                } else if (fancy[slot] != att[i] || owner[slot] != who) {
                    snprintf(err, errlen, "fancy magic collision (sq %d, %s)", s, rook ? "rook" : "bishop");
                    return -1;
                }
            }
            int &off = rook ? offR : offB;
            std::vector<u64> &dst = rook ? byteRook : byteBishop;
            if (off + (int)uniq.size() > (int)dst.size()) {
                snprintf(err, errlen, "byte-lookup list overflow (sq %d)", s);
                return -1;
            }
            std::copy(uniq.begin(), uniq.end(), dst.begin() + off);
            FancyEntry &e = rook ? rookFancy[s] : bishopFancy[s];
            e.factor = c.factor;
            e.base = c.base;
            e.offset = off;
            off += (int)uniq.size();
        }
    }
    if (offR != BYTE_ROOK_SIZE || offB != BYTE_BISHOP_SIZE) {
        snprintf(err, errlen, "byte-lookup sizes %d/%d", offR, offB);
        return -1;
    }
    memcpy(zobLo, PG_ZOBRIST_LO, sizeof(zobLo));
    memcpy(zobHi, PG_ZOBRIST_HI, sizeof(zobHi));
    return 0;
}

Tables HostTables::view() const {
    Tables t;
    t.knight = knight; t.king = king; t.rookEmpty = rookEmpty; t.bishopEmpty = bishopEmpty;
    t.between = between; t.line = line; t.rookMask = rookMask; t.bishopMask = bishopMask;
    t.rookMagic = rookMagic; t.bishopMagic = bishopMagic;
    t.rookPlain = rookPlain.data(); t.bishopPlain = bishopPlain.data();
    t.rookFancy = rookFancy; t.bishopFancy = bishopFancy; t.fancy = fancy.data();
    t.byteIdx = byteIdx.data(); t.byteRook = byteRook.data(); t.byteBishop = byteBishop.data();
    t.zobLo = zobLo; t.zobHi = zobHi;
    return t;
}

static const size_t kSizes[19] = {
    64 * 8, 64 * 8, 64 * 8, 64 * 8,              // knight king rookEmpty bishopEmpty
    4096 * 8, 4096 * 8,                          // between line
    64 * 8, 64 * 8, 64 * 8, 64 * 8,              // masks, plain factors
    (64u << ROOK_BITS) * 8, (64u << BISHOP_BITS) * 8,
    64 * 16, 64 * 16,                            // fancy entries
    FANCY_TABLE_SIZE * 8, FANCY_TABLE_SIZE,      // fancy, byte index
    BYTE_ROOK_SIZE * 8, BYTE_BISHOP_SIZE * 8,
    2 * ZOB_WORDS * 8,                           // zobrist lo + hi
};

TableLayout table_layout() {
    TableLayout L;
    size_t o = 0;
    for (int i = 0; i < 19; i++) {
        L.off[i] = o;
        o += (kSizes[i] + 255) & ~(size_t)255;
    }
    L.total = o;
    return L;
}

void pack_tables(const HostTables &h, std::vector<u8> &blob) {
    TableLayout L = table_layout();
    blob.assign(L.total, 0);
    const void *src[19] = {h.knight, h.king, h.rookEmpty, h.bishopEmpty, h.between, h.line, h.rookMask,
                           h.bishopMask, h.rookMagic, h.bishopMagic, h.rookPlain.data(), h.bishopPlain.data(),
                           h.rookFancy, h.bishopFancy, h.fancy.data(), h.byteIdx.data(), h.byteRook.data(),
                           h.byteBishop.data(), nullptr};
    for (int i = 0; i < 18; i++) memcpy(&blob[L.off[i]], src[i], kSizes[i]);
    memcpy(&blob[L.off[18]], h.zobLo, ZOB_WORDS * 8);
    memcpy(&blob[L.off[18] + ZOB_WORDS * 8], h.zobHi, ZOB_WORDS * 8);
}

Tables tables_at(const u8 *b) {
    TableLayout L = table_layout();
    Tables t;
    t.knight = (const u64 *)(b + L.off[0]);
    t.king = (const u64 *)(b + L.off[1]);
    t.rookEmpty = (const u64 *)(b + L.off[2]);
    t.bishopEmpty = (const u64 *)(b + L.off[3]);
    t.between = (const u64 *)(b + L.off[4]);
    t.line = (const u64 *)(b + L.off[5]);
    t.rookMask = (const u64 *)(b + L.off[6]);
    t.bishopMask = (const u64 *)(b + L.off[7]);
    t.rookMagic = (const u64 *)(b + L.off[8]);
    t.bishopMagic = (const u64 *)(b + L.off[9]);
    t.rookPlain = (const u64 *)(b + L.off[10]);
    t.bishopPlain = (const u64 *)(b + L.off[11]);
    t.rookFancy = (const FancyEntry *)(b + L.off[12]);
    t.bishopFancy = (const FancyEntry *)(b + L.off[13]);
    t.fancy = (const u64 *)(b + L.off[14]);
    t.byteIdx = (const u8 *)(b + L.off[15]);
    t.byteRook = (const u64 *)(b + L.off[16]);
    t.byteBishop = (const u64 *)(b + L.off[17]);
    t.zobLo = (const u64 *)(b + L.off[18]);
    t.zobHi = (const u64 *)(b + L.off[18] + ZOB_WORDS * 8);
    return t;
}

static char g_table_err[256];

const HostTables *host_tables(const char **err) {
    static HostTables *t = nullptr;
    static std::once_flag once;
    std::call_once(once, [] {
        HostTables *h = new HostTables();
        if (h->build(0x5EED0F0EC0DEULL, g_table_err, sizeof(g_table_err)) == 0) t = h;
        else delete h;
    });
    if (!t && err) *err = g_table_err;
    return t;
}

}  // namespace pg
