// ref_driver.cpp -- C entry points over the reference's own CPU move
// generator, compiled from the reference checkout into oracle/_ref/.
//
// TEST INFRASTRUCTURE ONLY (the checker, never the product).  It links the
// reference's translation units randoms.cpp, GlobalVars.cpp, Magics.cpp and
// util.cpp unmodified and includes MoveGeneratorBitboard.h with
// SKIP_CUDA_CODE (the host half of that header).  The recursion of
// perft_bb (perft_bb.h:79-101) lives in a file that also holds __global__
// code, so its twenty lines are restated here over the reference's own
// countMoves / generateBoards templates.
#define SKIP_CUDA_CODE
#include "MoveGeneratorBitboard.h"
#include <thread>
#include <vector>
#include <atomic>

static bool g_ref_inited = false;

extern "C" {

void ref_init(void) {
    if (g_ref_inited) return;
    MoveGeneratorBitboard::init();
    g_ref_inited = true;
}

static uint32 ref_count(HexaBitBoardPosition *p) {
    return p->chance == BLACK ? MoveGeneratorBitboard::countMoves<BLACK>(p)
                              : MoveGeneratorBitboard::countMoves<WHITE>(p);
}

static uint32 ref_gen_boards(HexaBitBoardPosition *p, HexaBitBoardPosition *out) {
    return p->chance == BLACK ? MoveGeneratorBitboard::generateBoards<BLACK>(p, out)
                              : MoveGeneratorBitboard::generateBoards<WHITE>(p, out);
}

// perft_bb.h:79-101
unsigned long long ref_perft(const void *pos48, unsigned depth) {
    HexaBitBoardPosition p;
    memcpy(&p, pos48, sizeof(p));
    if (depth == 0) return 1;
    if (depth == 1) return ref_count(&p);
    HexaBitBoardPosition child[MAX_MOVES];
    uint32 n = ref_gen_boards(&p, child);
    unsigned long long sum = 0;
    for (uint32 i = 0; i < n; i++) sum += ref_perft(&child[i], depth - 1);
    return sum;
}

// perft_bb over a split two plies below the root, on `threads` threads.
unsigned long long ref_perft_mt(const void *pos48, unsigned depth, int threads) {
    if (depth <= 2 || threads <= 1) return ref_perft(pos48, depth);
    HexaBitBoardPosition p;
    memcpy(&p, pos48, sizeof(p));
    HexaBitBoardPosition child[MAX_MOVES];
    uint32 n = ref_gen_boards(&p, child);
    std::vector<HexaBitBoardPosition> grand;
    for (uint32 i = 0; i < n; i++) {
        HexaBitBoardPosition g[MAX_MOVES];
        uint32 m = ref_gen_boards(&child[i], g);
        grand.insert(grand.end(), g, g + m);
    }
    std::vector<unsigned long long> res(grand.size(), 0);
    std::atomic<size_t> next(0);
    std::vector<std::thread> th;
    for (int t = 0; t < threads; t++)
        th.emplace_back([&]() {
            for (;;) {
                size_t i = next.fetch_add(1);
                if (i >= grand.size()) break;
                res[i] = ref_perft(&grand[i], depth - 2);
            }
        });
    for (auto &t : th) t.join();
    unsigned long long sum = 0;
    for (auto v : res) sum += v;
    return sum;
}

unsigned ref_count_moves(const void *pos48) {
    HexaBitBoardPosition p;
    memcpy(&p, pos48, sizeof(p));
    return ref_count(&p);
}

unsigned ref_generate_boards(const void *pos48, void *out48) {
    HexaBitBoardPosition p;
    memcpy(&p, pos48, sizeof(p));
    return ref_gen_boards(&p, (HexaBitBoardPosition *)out48);
}

unsigned ref_generate_moves(const void *pos48, unsigned short *out) {
    HexaBitBoardPosition p;
    memcpy(&p, pos48, sizeof(p));
    CMove moves[MAX_MOVES];
    uint32 n = p.chance == BLACK ? MoveGeneratorBitboard::generateMoves<BLACK>(&p, moves)
                                 : MoveGeneratorBitboard::generateMoves<WHITE>(&p, moves);
    memcpy(out, moves, n * sizeof(CMove));
    return n;
}

// makeMove<chance, updateHash=true> with the 128-bit key (MoveGeneratorBitboard.h:3575)
void ref_make_move(void *pos48, unsigned long long *key2, unsigned short move) {
    HexaBitBoardPosition *p = (HexaBitBoardPosition *)pos48;
    HashKey128b h(key2[0], key2[1]);
    CMove m;
    memcpy(&m, &move, sizeof(m));
    if (p->chance == BLACK) MoveGeneratorBitboard::makeMove<BLACK, true>(p, h, m);
    else MoveGeneratorBitboard::makeMove<WHITE, true>(p, h, m);
    key2[0] = h.lowPart;
    key2[1] = h.highPart;
}

void ref_zobrist128(const void *pos48, unsigned long long *key2) {
    HexaBitBoardPosition p;
    memcpy(&p, pos48, sizeof(p));
    HashKey128b h = MoveGeneratorBitboard::computeZobristKey128b(&p);
    key2[0] = h.lowPart;
    key2[1] = h.highPart;
}

// Utils::readFENString + Utils::board088ToHexBB (util.cpp:189, :92).  The
// reference parser walks past the end of a FEN that stops after the castling
// field, so callers pass all four leading fields.
int ref_fen(const char *fen, void *out48) {
    char buf[1024];
    strncpy(buf, fen, sizeof(buf) - 1);
    buf[sizeof(buf) - 1] = 0;
    BoardPosition b;
    Utils::readFENString(buf, &b);
    Utils::board088ToHexBB((HexaBitBoardPosition *)out48, &b);
    return 0;
}

// Table contents the magic variants must reproduce.
unsigned long long ref_fancy_table(int i) { return fancy_magic_lookup_table[i]; }
unsigned ref_byte_table(int i) { return fancy_byte_magic_lookup_table[i]; }
unsigned long long ref_byte_rook(int i) { return fancy_byte_RookLookup[i]; }
unsigned long long ref_byte_bishop(int i) { return fancy_byte_BishopLookup[i]; }
int ref_fancy_offset(int rook, int sq) { return rook ? rook_magics_fancy[sq].offset : bishop_magics_fancy[sq].offset; }

}  // extern "C"
