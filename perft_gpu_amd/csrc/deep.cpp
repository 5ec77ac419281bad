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
//
// With a disk table (pg_perft_deep_disk) the nodes at one remaining depth are
// also looked up in, and written to, a file laid out as the reference's
// perfthash.dat (checkAndCreateDiskHash, launcher.h:1698-1722; DiskHashEntry,
// chess.h:518-525): 2^12 index entries of 32 bytes, then chained entries
// appended at the end; an entry is {key lo, key hi, perft, depth, next}, an
// all-zero key is a free slot, next = ~0 ends a chain.  Every finished node
// at that depth is written (and synced) before the walk goes on, so a run
// that is interrupted resumes from the subtrees it has finished.  In the
// reference the probe and store are compiled out (diskTTProbe/diskTTStore,
// launcher.h:700-789, used under ENABLE_DISK_HASH == 7 at :1220-1231 and
// :1381-1386); the file is created and left blank.
#include <fcntl.h>
#include <unistd.h>
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

// perfthash.dat: 32-byte entries, 2^12 of them indexed by the low key bits
struct DiskEntry {
    uint64_t lo, hi, perft;
    uint32_t depth, next;
};
static_assert(sizeof(DiskEntry) == 32, "DiskHashEntry is 32 bytes");
constexpr uint64_t DISK_INDEX = 1u << 12;
constexpr uint32_t DISK_END = ~0u;

struct DiskTable {
    int fd = -1;
    uint64_t entries = 0;  // file length in entries

    int open_file(const char *path) {
        fd = open(path, O_RDWR | O_CREAT, 0644);
        if (fd < 0) return PG_EINVAL;
        const off_t len = lseek(fd, 0, SEEK_END);
        if (len < 0 || len % (off_t)sizeof(DiskEntry)) return PG_EINVAL;
        entries = (uint64_t)len / sizeof(DiskEntry);
        if (entries == 0) {  // a new table: the blank index part
            std::vector<DiskEntry> blank(DISK_INDEX);
            memset(blank.data(), 0, blank.size() * sizeof(DiskEntry));
            if (pwrite(fd, blank.data(), blank.size() * sizeof(DiskEntry), 0) != (ssize_t)(blank.size() * sizeof(DiskEntry)))
                return PG_EINVAL;
            entries = DISK_INDEX;
        }
        return entries >= DISK_INDEX ? PG_OK : PG_EINVAL;
    }
    ~DiskTable() {
        if (fd >= 0) close(fd);
    }
    bool read_at(uint64_t i, DiskEntry &e) const {
        return i < entries && pread(fd, &e, sizeof(e), (off_t)(i * sizeof(e))) == (ssize_t)sizeof(e);
    }
    bool write_at(uint64_t i, const DiskEntry &e) const {
        return pwrite(fd, &e, sizeof(e), (off_t)(i * sizeof(e))) == (ssize_t)sizeof(e);
    }
    // 1: found (*v), 0: absent, < 0: error
    int probe(uint64_t lo, uint64_t hi, uint32_t depth, uint64_t *v) const {
        uint64_t i = lo & (DISK_INDEX - 1);
        for (uint64_t steps = 0; steps <= entries; steps++) {
            DiskEntry e;
            if (!read_at(i, e)) return PG_EINVAL;
            if (e.lo == 0 && e.hi == 0) return 0;
            if (e.lo == lo && e.hi == hi && e.depth == depth) {
                *v = e.perft;
                return 1;
            }
            if (e.next == DISK_END) return 0;
            i = e.next;
        }
        return PG_EINVAL;  // a chain that loops: not a table this code wrote
    }
    // into the first free slot of the key's chain, or a new entry at the end
    int store(uint64_t lo, uint64_t hi, uint32_t depth, uint64_t v) {
        uint64_t i = lo & (DISK_INDEX - 1);
        const DiskEntry mine{lo, hi, v, depth, DISK_END};
        for (uint64_t steps = 0; steps <= entries; steps++) {
            DiskEntry e;
            if (!read_at(i, e)) return PG_EINVAL;
            if (e.lo == 0 && e.hi == 0) return write_at(i, mine) && fdatasync(fd) == 0 ? PG_OK : PG_EINVAL;
            if (e.next == DISK_END) {
                const uint64_t j = entries;
                if (j >= DISK_END) return PG_ENOMEM;
                // the new entry first, then the link to it, so an interrupted
                // store leaves at worst an unreachable entry
                if (!write_at(j, mine)) return PG_EINVAL;
                entries++;
                e.next = (uint32_t)j;
                return write_at(i, e) && fdatasync(fd) == 0 ? PG_OK : PG_EINVAL;
            }
            i = e.next;
        }
        return PG_EINVAL;
    }
};

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
    DiskTable *disk = nullptr;
    uint32_t diskDepth = 0;

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
        if (disk && depth == diskDepth) {
            uint64_t v = 0;
            const int r = disk->probe(key.lo, key.hi, depth, &v);
            if (r < 0) {
                rc = r;
                return 0;
            }
            if (r == 1) {
                st.disk_hits++;
                return v;
            }
            const u128 total = walk_node(p, key, depth, path);
            // totals of 2^64 and above are not kept (the entry holds 64 bits)
            if (rc == PG_OK && (total >> 64) == 0) {
                if ((rc = disk->store(key.lo, key.hi, depth, (uint64_t)total)) == PG_OK) st.disk_stores++;
            }
            return total;
        }
        return walk_node(p, key, depth, path);
    }

    u128 walk_node(const pg_pos48 &p, const pg_key128 &key, uint32_t depth, std::string &path) {
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

extern "C" int pg_perft_deep_disk(pg_ctx *ctx, pg_tt *tt, const pg_pos48 *pos, uint32_t depth, uint32_t launch_depth,
                                  uint32_t divided_depth, int slider, pg_divide_cb cb, void *user,
                                  const char *hash_path, uint32_t disk_depth, uint64_t out[2], pg_deep_stats *stats) {
    if (!ctx || !pos || !out || depth > 64 || launch_depth < 1 || launch_depth > 15) return PG_EINVAL;
    const double t0 = now_ms();
    DeepWalk w{ctx, tt, slider, launch_depth, divided_depth, cb, user, {}, (size_t)1 << 22, {}, PG_OK};
    DiskTable disk;
    if (hash_path) {
        // DISK_HASH_LEVEL (launcher.h:40, :1463): two plies below the root
        if (disk_depth == 0) disk_depth = depth > 2 ? depth - 2 : 1;
        if (disk_depth > depth || disk.open_file(hash_path) != PG_OK) return PG_EINVAL;
        w.disk = &disk;
        w.diskDepth = disk_depth;
    }
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

extern "C" int pg_perft_deep(pg_ctx *ctx, pg_tt *tt, const pg_pos48 *pos, uint32_t depth, uint32_t launch_depth,
                             uint32_t divided_depth, int slider, pg_divide_cb cb, void *user, uint64_t out[2],
                             pg_deep_stats *stats) {
    return pg_perft_deep_disk(ctx, tt, pos, depth, launch_depth, divided_depth, slider, cb, user, nullptr, 0, out,
                              stats);
}
