// deep.cpp -- perft of depths whose totals pass 2^64, with divided output.
//
// The reference walks the top of the tree on the host with InfInt totals and
// a CPU "complete" transposition table, prints one line per move of every
// node deeper than DIVIDED_PERFT_DEPTH, and hands each node at its launch
// depth to the GPU (perft_bb_cpu_launcher, launcher.h:1181-1389; dividedPerft,
// :1392-1450; complete TT, :598-697).  This is the same walk written over the
// C ABI: host nodes sum 128-bit totals and are memoised in a host table keyed
// by (128-bit Zobrist key, remaining depth); every node at the launch depth is
// one device call (pg_perft_tt with a table, pg_perft without), whose 64-bit
// result is overflow-checked on the device.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>
#include "../../include/perft_gpu.h"

namespace {

typedef unsigned __int128 u128;

struct NodeKey {
    uint64_t lo, hi;
    uint32_t depth;
    bool operator==(const NodeKey &o) const { return lo == o.lo && hi == o.hi && depth == o.depth; }
};
struct NodeHash {
    size_t operator()(const NodeKey &k) const { return (size_t)(k.lo ^ (k.hi >> 7) ^ ((uint64_t)k.depth * 0x9E3779B97F4A7C15ULL)); }
};

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

uint64_t occupancy(const pg_pos48 &p) {
    return p.kings | (p.pawns & 0x00FFFFFFFFFFFF00ULL) | p.knights | p.bishopQueens | p.rookQueens;
}

// Utils::getCompactMoveString (utils.h:251-278): "e2-e4 ", '*' for a capture
void move_string(const pg_pos48 &p, pg_move m, char out[8]) {
    const int from = m & 63, to = (m >> 6) & 63, flags = m >> 12;
    const bool cap = flags == 5 || (occupancy(p) >> to & 1);
    snprintf(out, 8, "%c%d%c%c%d ", 'a' + (from & 7), (from >> 3) + 1, cap ? '*' : '-', 'a' + (to & 7), (to >> 3) + 1);
}

struct DeepWalk {
    pg_ctx *ctx;
    pg_tt *tt;
    int slider;
    uint32_t launch, divided;
    pg_divide_cb cb;
    void *user;
    std::unordered_map<NodeKey, u128, NodeHash> table;
    size_t maxEntries = (size_t)1 << 22;
    pg_deep_stats st{};
    int rc = PG_OK;

    u128 device(const pg_pos48 &p, uint32_t depth) {
        uint64_t v = 0;
        const double t0 = now_ms();
        const int r = tt ? pg_perft_tt(ctx, tt, &p, depth, slider, &v) : pg_perft(ctx, &p, depth, slider, &v);
        st.device_ms += now_ms() - t0;
        st.device_calls++;
        if (r != PG_OK) rc = r;
        return v;
    }

    u128 walk(const pg_pos48 &p, const pg_key128 &key, uint32_t depth, std::string &path) {
        if (rc != PG_OK) return 0;
        if (depth <= launch) return device(p, depth);
        st.host_nodes++;
        const NodeKey nk{key.lo, key.hi, depth};
        auto it = table.find(nk);
        if (it != table.end()) {
            st.host_tt_hits++;
            return it->second;
        }
        pg_move moves[256];
        uint32_t n = 0;
        if ((rc = pg_host_generate_moves(&p, moves, &n)) != PG_OK) return 0;
        // quiet moves first, as sortMoves orders them (launcher.h:367-391)
        std::stable_partition(moves, moves + n, [](pg_move m) { return (m >> 12) == 0; });
        u128 total = 0;
        for (uint32_t i = 0; i < n && rc == PG_OK; i++) {
            pg_pos48 c = p;
            pg_key128 ck = key;
            if ((rc = pg_host_make_move(&c, moves[i], &ck)) != PG_OK) break;
            char ms[8];
            move_string(p, moves[i], ms);
            const size_t mark = path.size();
            path += ms;
            const u128 child = walk(c, ck, depth - 1, path);
            if (rc == PG_OK && cb && depth > divided) cb(user, path.c_str(), (uint64_t)child, (uint64_t)(child >> 64));
            path.resize(mark);
            total += child;
        }
        if (rc == PG_OK && table.size() < maxEntries) table.emplace(nk, total);
        return total;
    }
};

}  // namespace

extern "C" int pg_perft_deep(pg_ctx *ctx, pg_tt *tt, const pg_pos48 *pos, uint32_t depth, uint32_t launch_depth,
                             uint32_t divided_depth, int slider, pg_divide_cb cb, void *user, uint64_t out[2],
                             pg_deep_stats *stats) {
    if (!ctx || !pos || !out || depth > 64 || launch_depth < 1 || launch_depth > 15) return PG_EINVAL;
    const double t0 = now_ms();
    DeepWalk w{ctx, tt, slider, launch_depth, divided_depth, cb, user, {}, (size_t)1 << 22, {}, PG_OK};
    pg_key128 key;
    int rc = pg_zobrist128(pos, &key);
    if (rc != PG_OK) return rc;
    std::string path = "..";
    const u128 total = depth == 0 ? 1 : w.walk(*pos, key, depth, path);
    w.st.total_ms = now_ms() - t0;
    if (stats) *stats = w.st;
    if (w.rc != PG_OK) return w.rc;
    out[0] = (uint64_t)total;
    out[1] = (uint64_t)(total >> 64);
    return PG_OK;
}
