// engine.h -- per-device perft context (the object behind pg_ctx).
#pragma once
#include <hip/hip_runtime.h>
#include <vector>
#include "../../include/perft_gpu.h"
#include "tables.h"
#include "tt.h"

namespace pg {

void set_error(const char *fmt, ...);
int scratch_setup(int device, u8 **out);  // device address of the static scratch

#define HIP_TRY(expr)                                                   \
    do {                                                                \
        hipError_t e_ = (expr);                                         \
        if (e_ != hipSuccess) {                                         \
            set_error("HIP error %s at line %d", hipGetErrorString(e_), __LINE__); \
            return PG_EHIP;                                             \
        }                                                               \
    } while (0)

#define PG_VARIANT_CALL(v, fn, ...)                  \
    switch (v) {                                     \
    case SL_KOGGE: fn<SL_KOGGE>(__VA_ARGS__); break; \
    case SL_PLAIN: fn<SL_PLAIN>(__VA_ARGS__); break; \
    case SL_FANCY: fn<SL_FANCY>(__VA_ARGS__); break; \
    default: fn<SL_BYTE>(__VA_ARGS__); break;        \
    }


struct TTable;
struct Slot;
struct Server;

struct Engine {
    int device = 0, cus = 0, bpc = 8, slider = SL_KOGGE;
    size_t ldsPerBlock = 0;
    int clockKHz = 0;
    bool timing = false;
    pg_config cfg{};
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    u8 *dTables = nullptr;
    Tables dt{};
    u8 *arena = nullptr;     // tree arena, used as a stack
    size_t arenaBytes = 0, top = 0, peak = 0;
    u8 *dSmall = nullptr;    // accumulator and error flag
    u64 *dAcc = nullptr;
    u64 *dItems = nullptr;   // items consumed by fused leaf launches
    int leafMode = 0;        // 0 fused last two plies, 1 materialized item list
    bool graphPerft = true;  // pg_perft may run as one graph launch
    int batchGrid = -1;      // work-groups per kernel of pg_perft_batch: -1 sized to the batch, 0 whole chip
    u32 *dErr = nullptr;
    u32 *dOvf = nullptr;     // set when a 64-bit total wraps
    u64 bias = 0;            // test hook: starting value of every root total
    u64 *hostBuf = nullptr;  // pinned, 256 bytes: [0] results, [1] flags, [8..13] a staged root
    u8 *scratch = nullptr;   // device address of the static scratch (pg_perft_noalloc)
    pg_stats stats{};

    void *alloc(size_t bytes);
    u32 grid(u64 n, int wg) const;
    u32 leaf_grid(u64 n) const;  // work-groups for the fused leaf
    int read_u64(const u64 *d, u64 *h);
    int scan(const u32 *cnt, u64 n, u64 *off, u64 *total);
    int leaf(int v, const Pos *b, u64 n, const u32 *par, const u16 *mv, u64 m);
    int leaf_fused(int v, const Pos *b, const u32 *cnt, u64 n);
    int leaf_moves(int v, const Pos *gp, const u32 *par, const u16 *mv, u64 n);
    int bfs(int v, const Pos *boards, u64 n, const u32 *cnt, int r, int ply);
    int split(int v, const Pos *boards, u64 n, const u32 *cnt, int r, int ply);
    int perft(const Pos &root, u32 depth, int v, u64 *out);
    int check_flags(const char *what);  // consistency and overflow flags of the last call

    // batched searches (graph.hip)
    Slot *batchSlot = nullptr;
    // tables and slot servers created on this context; pg_shutdown releases
    // those still alive
    std::vector<TTable *> tts;
    std::vector<Server *> servers;

    // transposition-table perft
    bool dedup = true;
    int tt_bfs(int v, TTable &tt, Pos *B, Key128 *K, u64 n, u32 *cnt, int r, u64 *P, int ply);
    int tt_split(int v, TTable &tt, Pos *B, Key128 *K, u64 n, u32 *cnt, int r, u64 *P, int ply);
    int tt_last(int v, TTable &tt, Pos *B, Key128 *K, u64 n, u32 *cnt, u64 *P, int ply);
    int perft_tt(TTable &tt, const Pos &root, u32 depth, int v, u64 *out);
};

int slot_init(Engine &e, Slot &sl, u8 *region, size_t bytes, hipStream_t stream);
void slot_free(Slot &sl);
void slot_delete(Slot *sl);
void slot_forget_tt(Slot &sl, const TTable *tt);  // drop graphs that bake this table in
int batch_perft(Engine &e, Slot &sl, TTable *tt, const Pos *roots, u64 n, u32 depth, int v, int dedup, int bpc,
                u64 *out, pg_batch_stats *bs);
int perft_graph(Engine &e, const Pos &root, u32 depth, int v, u64 *out, bool *fits);
void tt_free(TTable *t);      // release a table (no users left)
void server_free(Server *s);  // release a slot server
int server_submit(Server &s, uint32_t slot, const pg_pos48 *pos, uint32_t n, uint32_t depth, int slider, int dedup,
                  int wgCap);

// Device transposition table (the object behind pg_tt).
struct TTable {
    Engine *eng = nullptr;
    int users = 0;  // slot servers sharing this table
    u8 *mem = nullptr;
    size_t bytes = 0;
    TTView view{};
    int bits[TT_MAX_DEPTH] = {0};
    TTCounters *dCtr = nullptr;
};

}  // namespace pg
