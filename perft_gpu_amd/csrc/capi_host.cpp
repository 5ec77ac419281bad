// capi_host.cpp -- host-side entry points of the C ABI: FEN conversion,
// Zobrist keys, host move listing, the seeded position generator, and
// pg_cpu_perft, the reference's single-thread CPU perft as its own entry
// point (no device entry point calls it: device perft never falls back to
// the host).
#include <cstdio>
#include <cstring>
#include <algorithm>
#include "../../include/perft_gpu.h"
#include "tables.h"

using namespace pg;

namespace pg {
void set_error(const char *fmt, ...);  // engine.hip
}

namespace {

// FEN piece placement, side, castling and en-passant file (util.cpp:189-281;
// the reference also stops after the en-passant field).
int parse_fen(const char *fen, Pos &p) {
    memset(&p, 0, sizeof(p));
    int i = 0, row = 0, col = 0;
    static const char *kPieces = "PNBRQKpnbrqk";
    for (; fen[i]; i++) {
        const char ch = fen[i];
        if (ch == '/' || ch == '\\') {
            row++;
            col = 0;
        } else if (ch >= '1' && ch <= '8') {
            col += ch - '0';
        } else if (const char *f = strchr(kPieces, ch)) {
            const int idx = (int)(f - kPieces), color = idx >= 6, piece = idx % 6;
            const int r = 7 - row;
            if (r < 0 || col > 7) return PG_EINVAL;
            const u64 b = bit(r * 8 + col);
            if (color == WHITE) p.white |= b;
            switch (piece) {
            case 0: p.pawns |= b; break;
            case 1: p.knights |= b; break;
            case 2: p.bishopQueens |= b; break;
            case 3: p.rookQueens |= b; break;
            case 4: p.bishopQueens |= b; p.rookQueens |= b; break;
            default: p.kings |= b; break;
            }
            col++;
        }
        if (row >= 7 && col == 8) break;
    }
    if (!fen[i]) return PG_EINVAL;
    i++;
    while (fen[i] == ' ') i++;
    const int side = (fen[i] == 'b' || fen[i] == 'B') ? BLACK : WHITE;
    if (fen[i]) i++;
    while (fen[i] == ' ') i++;
    int wc = 0, bc = 0;
    for (; fen[i] && fen[i] != ' '; i++) {
        if (fen[i] == 'K') wc |= 1;
        else if (fen[i] == 'Q') wc |= 2;
        else if (fen[i] == 'k') bc |= 1;
        else if (fen[i] == 'q') bc |= 2;
    }
    while (fen[i] == ' ') i++;
    const int ep = (fen[i] >= 'a' && fen[i] <= 'h') ? fen[i] - 'a' + 1 : 0;
    p.pawns |= (u64)wc | ((u64)bc << 2) | ((u64)ep << 4) | ((u64)side << 63);
    // one king each, or the generator's king-square lookups are undefined
    if (!pos_sane(p)) return PG_EINVAL;
    return PG_OK;
}

const HostTables *tables_or_null() { return host_tables(nullptr); }

u64 next_rand(u64 &s) {  // splitmix64
    u64 z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// order key independent of capture flags: from, to, promotion piece
u32 move_key(u16 m) {
    const u32 fl = m >> 12;
    const u32 promo = (fl & 8) ? (fl & 3) + 1 : 0;
    return (m & 0xFFF) | (promo << 12);
}

}  // namespace

extern "C" {

int pg_fen_to_pos(const char *fen, pg_pos48 *out) {
    if (!fen || !out) return PG_EINVAL;
    Pos p;
    const int rc = parse_fen(fen, p);
    if (rc) return rc;
    memcpy(out, &p, sizeof(p));
    return PG_OK;
}

int pg_pos_to_fen(const pg_pos48 *pos, char *buf, int buflen) {
    if (!pos || !buf || buflen < 100) return PG_EINVAL;
    Pos p;
    memcpy(&p, pos, sizeof(p));
    char *o = buf;
    const u64 pawns = pos_pawns(p);
    for (int r = 7; r >= 0; r--) {
        int empty = 0;
        for (int f = 0; f < 8; f++) {
            const u64 b = bit(r * 8 + f);
            char ch = 0;
            if (p.kings & b) ch = 'k';
            else if (pawns & b) ch = 'p';
            else if (p.knights & b) ch = 'n';
            else if (p.bishopQueens & p.rookQueens & b) ch = 'q';
            else if (p.bishopQueens & b) ch = 'b';
            else if (p.rookQueens & b) ch = 'r';
            if (!ch) { empty++; continue; }
            if (empty) { *o++ = (char)('0' + empty); empty = 0; }
            *o++ = (p.white & b) ? (char)(ch - 32) : ch;
        }
        if (empty) *o++ = (char)('0' + empty);
        if (r) *o++ = '/';
    }
    *o++ = ' ';
    *o++ = pos_side(p) ? 'b' : 'w';
    *o++ = ' ';
    const int wc = pos_wcastle(p), bc = pos_bcastle(p);
    if (!wc && !bc) *o++ = '-';
    if (wc & 1) *o++ = 'K';
    if (wc & 2) *o++ = 'Q';
    if (bc & 1) *o++ = 'k';
    if (bc & 2) *o++ = 'q';
    *o++ = ' ';
    const int ep = pos_ep(p);
    if (ep) {
        *o++ = (char)('a' + ep - 1);
        *o++ = pos_side(p) ? '3' : '6';
    } else {
        *o++ = '-';
    }
    snprintf(o, (size_t)(buflen - (o - buf)), " 0 1");
    return PG_OK;
}

int pg_zobrist128(const pg_pos48 *pos, pg_key128 *out) {
    const HostTables *h = tables_or_null();
    if (!h) return PG_ETABLE;
    if (!pos || !out) return PG_EINVAL;
    Pos p;
    memcpy(&p, pos, sizeof(p));
    if (!pos_sane(p)) return PG_EINVAL;
    const Key128 k = zobrist128(p, h->view());
    out->lo = k.lo;
    out->hi = k.hi;
    return PG_OK;
}

int pg_host_generate_moves(const pg_pos48 *pos, pg_move *out, uint32_t *n) {
    const HostTables *h = tables_or_null();
    if (!h) return PG_ETABLE;
    if (!pos || !out || !n) return PG_EINVAL;
    Pos p;
    memcpy(&p, pos, sizeof(p));
    if (!pos_sane(p)) return PG_EINVAL;
    *n = generate_any<SL_FANCY>(p, h->view(), out, 256);
    return PG_OK;
}

int pg_host_count_moves(const pg_pos48 *pos, uint32_t n, int slider, uint32_t *out) {
    const HostTables *h = tables_or_null();
    if (!h) return PG_ETABLE;
    if (n && (!pos || !out)) return PG_EINVAL;
    if (slider < 0 || slider >= SL_COUNT) return PG_EINVAL;
    const Tables t = h->view();
    for (uint32_t i = 0; i < n; i++) {
        Pos p;
        memcpy(&p, &pos[i], sizeof(p));
        if (!pos_sane(p)) return PG_EINVAL;
        switch (slider) {
        case SL_KOGGE: out[i] = count_any<SL_KOGGE>(p, t); break;
        case SL_PLAIN: out[i] = count_any<SL_PLAIN>(p, t); break;
        case SL_FANCY: out[i] = count_any<SL_FANCY>(p, t); break;
        default: out[i] = count_any<SL_BYTE>(p, t); break;
        }
    }
    return PG_OK;
}

}  // extern "C"

namespace {
// perft_bb (perft_bb.h:79-101): children by generation and make, the
// legal-move count at depth 1
u64 host_perft(const Pos &p, u32 depth, const Tables &t, bool *ovf) {
    if (depth == 0) return 1;
    if (depth == 1) return count_any<SL_FANCY>(p, t);
    u16 mv[256];
    const u32 n = generate_any<SL_FANCY>(p, t, mv, 256);
    u64 s = 0;
    for (u32 i = 0; i < n && !*ovf; i++) {
        Pos q = p;
        make_any(q, mv[i], t);
        const u64 c = host_perft(q, depth - 1, t, ovf);
        if (s + c < s) *ovf = true;
        s += c;
    }
    return s;
}
}  // namespace

extern "C" {

int pg_cpu_perft(const pg_pos48 *pos, uint32_t depth, uint64_t *out) {
    const HostTables *h = tables_or_null();
    if (!h) return PG_ETABLE;
    if (!pos || !out || depth > 15) return PG_EINVAL;
    Pos p;
    memcpy(&p, pos, sizeof(p));
    if (!pos_sane(p)) return PG_EINVAL;
    bool ovf = false;
    const u64 v = host_perft(p, depth, h->view(), &ovf);
    if (ovf) {
        set_error("perft(%u) does not fit 64 bits", depth);
        return PG_EOVERFLOW;
    }
    *out = v;
    return PG_OK;
}

int pg_host_children(const pg_pos48 *pos, uint32_t n, pg_pos48 *out, uint64_t cap, uint32_t *counts) {
    const HostTables *h = tables_or_null();
    if (!h) return PG_ETABLE;
    if (n && (!pos || !counts)) return PG_EINVAL;
    const Tables t = h->view();
    u64 w = 0;
    u16 mv[256];
    for (uint32_t i = 0; i < n; i++) {
        Pos p;
        memcpy(&p, &pos[i], sizeof(p));
        if (!pos_sane(p)) return PG_EINVAL;
        const u32 c = generate_any<SL_FANCY>(p, t, mv, 256);
        counts[i] = c;
        if (w + c > cap) return PG_ENOMEM;
        for (u32 j = 0; j < c; j++) {
            Pos q = p;
            make_any(q, mv[j], t);
            memcpy(&out[w + j], &q, sizeof(q));
        }
        w += c;
    }
    return PG_OK;
}

int pg_host_make_move(pg_pos48 *pos, pg_move move, pg_key128 *key) {
    const HostTables *h = tables_or_null();
    if (!h) return PG_ETABLE;
    if (!pos) return PG_EINVAL;
    Pos p;
    memcpy(&p, pos, sizeof(p));
    if (!pos_sane(p)) return PG_EINVAL;
    if (key) {
        Key128 k = {key->lo, key->hi};
        make_any_hash(p, move, k, h->view());
        key->lo = k.lo;
        key->hi = k.hi;
    } else {
        make_any(p, move, h->view());
    }
    memcpy(pos, &p, sizeof(p));
    return PG_OK;
}

int pg_host_slider_attacks(int slider, int rook, const uint8_t *sq, const uint64_t *occ, uint32_t n, uint64_t *out) {
    const HostTables *h = tables_or_null();
    if (!h) return PG_ETABLE;
    if (n && (!sq || !occ || !out)) return PG_EINVAL;
    if (slider < 0 || slider >= SL_COUNT) return PG_EINVAL;
    const Tables t = h->view();
    for (uint32_t i = 0; i < n; i++) {
        if (sq[i] > 63) return PG_EINVAL;
        const int s = sq[i];
        const u64 o = occ[i];
        u64 a;
        switch (slider) {
        case SL_KOGGE: a = rook ? Sl<SL_KOGGE>::rook(t, s, o) : Sl<SL_KOGGE>::bishop(t, s, o); break;
        case SL_PLAIN: a = rook ? Sl<SL_PLAIN>::rook(t, s, o) : Sl<SL_PLAIN>::bishop(t, s, o); break;
        case SL_FANCY: a = rook ? Sl<SL_FANCY>::rook(t, s, o) : Sl<SL_FANCY>::bishop(t, s, o); break;
        default: a = rook ? Sl<SL_BYTE>::rook(t, s, o) : Sl<SL_BYTE>::bishop(t, s, o); break;
        }
        out[i] = a;
    }
    return PG_OK;
}

int pg_random_positions(uint64_t seed, uint32_t n, uint32_t min_ply, uint32_t max_ply, pg_pos48 *out, uint32_t *plies) {
    const HostTables *h = tables_or_null();
    if (!h) return PG_ETABLE;
    if ((!out && n) || min_ply > max_ply || max_ply > 1000) return PG_EINVAL;
    const Tables t = h->view();
    Pos start;
    parse_fen("rnbqkbnr/pppppppp/8/8/8/8/PPPPPPPP/RNBQKBNR w KQkq - 0 1", start);
    u64 rng = seed;
    u16 mv[256];
    for (u32 i = 0; i < n; i++) {
        for (;;) {
            const u32 ply = min_ply + (u32)(next_rand(rng) % (u64)(max_ply - min_ply + 1));
            Pos p = start;
            bool ok = true;
            for (u32 j = 0; j < ply && ok; j++) {
                const u32 k = generate_any<SL_FANCY>(p, t, mv, 256);
                if (!k) { ok = false; break; }
                std::sort(mv, mv + k, [](u16 a, u16 b) { return move_key(a) < move_key(b); });
                make_any(p, mv[next_rand(rng) % k], t);
            }
            if (ok && count_any<SL_FANCY>(p, t) == 0) ok = false;
            if (!ok) continue;
            memcpy(&out[i], &p, sizeof(p));
            if (plies) plies[i] = ply;
            break;
        }
    }
    return PG_OK;
}

}  // extern "C"
