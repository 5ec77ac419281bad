// tt_engine.hip -- breadth-first perft with the device transposition table.
//
// Replaces perft_bb_gpu_simple_hash (perft_bb.h:1242-1589): per level, the
// children are made with the incremental 128-bit key, the table of their
// remaining depth is probed (a hit resolves the child), identical children
// are collapsed to one owner (FIND_DUPLICATES_IN_BFS, perft_bb.h:788-846;
// here on the full 128-bit key, the same test a table hit passes), the live
// children are expanded recursively, and on the way up every computed perft
// is stored and summed into its parent (calcPerftNFromPerftNminus1,
// perft_bb.h:849-961).
// A level that does not fit the arena is split in halves; the first half's
// stores then serve the second half.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include "engine.h"

using namespace pg;

namespace pg {

template <int V> static void launch_expand_tt(Engine &c, TTable &tt, const Pos *B, const Key128 *K, const u32 *par,
                                              const u16 *mv, u64 m, Pos *C, Key128 *CK, u64 *CP, u32 *Ccnt, u8 *S,
                                              u32 *D, int rc) {
    constexpr int W = wg_size<V>();
    k_expand_tt<V><<<c.grid(m, W), W, 0, c.stream>>>(B, K, par, mv, m, C, CK, CP, Ccnt, S, D, tt.view, rc, tt.dCtr, c.dt);
}
template <int V>
static void launch_leaf_per(Engine &c, const Pos *B, const u32 *cnt, u64 n, u64 *P, const u32 *idx) {
    if constexpr (V != SL_BYTE)
        k_leaf_fused<V, true><<<c.leaf_grid(n), LEAF_WG, 0, c.stream>>>(B, cnt, n, c.dAcc, c.dItems, P, c.dErr, c.dOvf,
                                                                        c.dt, idx);
}
template <int V> static void launch_expand_keys(Engine &c, TTable &tt, const Pos *B, const Key128 *K, const u32 *par,
                                                const u16 *mv, u64 m, Key128 *CK, u64 *CP, u8 *S, int rc) {
    if constexpr (V != SL_BYTE) {
        constexpr int W = wg_size<V>();
        k_expand_keys<V><<<c.grid(m, W), W, 0, c.stream>>>(B, K, par, mv, m, CK, CP, S, tt.view, rc, tt.dCtr, c.dt);
    }
}
template <int V>
static void launch_leaf_moves_per(Engine &c, const Pos *B, const u32 *par, const u16 *mv, const u32 *idx, u64 n, u64 *P) {
    if constexpr (V != SL_BYTE)
        k_leaf_moves_per<V><<<c.leaf_grid(n), LEAF_WG, 0, c.stream>>>(B, par, mv, idx, n, c.dItems, P, c.dErr, c.dOvf,
                                                                      c.dt);
}
template <int V> static void launch_gen(Engine &c, const Pos *b, u64 n, const u64 *off, u16 *mv, u32 *par) {
    constexpr int W = wg_size<V>();
    k_generate<V><<<c.grid(n, W), W, 0, c.stream>>>(b, n, off, mv, par, c.dErr, c.dt);
}
template <int V> static void launch_cnt(Engine &c, const Pos *b, u64 n, u32 *cnt) {
    constexpr int W = wg_size<V>();
    k_count<V><<<c.grid(n, W), W, 0, c.stream>>>(b, n, cnt, c.dt);
}

static u64 pow2_at_least(u64 x) {
    u64 p = 1;
    while (p < x) p <<= 1;
    return p;
}

int Engine::tt_split(int v, TTable &tt, Pos *B, Key128 *K, u64 n, u32 *cnt, int r, u64 *P, int ply) {
    if (n <= 1) {
        set_error("arena too small for one board's children at ply %d", ply);
        return PG_ENOMEM;
    }
    stats.splits++;
    const u64 h = n / 2;
    int rc = tt_bfs(v, tt, B, K, h, cnt, r, P, ply);
    if (rc) return rc;
    return tt_bfs(v, tt, B + h, K + h, n - h, cnt + h, r, P + h, ply);
}

// P[i] = perft(r) of every valid board i (boards marked invalid keep P[i]).
int Engine::tt_bfs(int v, TTable &tt, Pos *B, Key128 *K, u64 n, u32 *cnt, int r, u64 *P, int ply) {
    if (n == 0) return PG_OK;
    if (r == 1) {
        k_perft1<<<grid(n, 256), 256, 0, stream>>>(B, n, cnt, P, ply == 0 ? bias : 0, dOvf);
        HIP_TRY(hipGetLastError());
        return PG_OK;
    }
    if (r == 2 && v != SL_BYTE) {
        // Boards resolved from the table or collapsed into an equal board
        // have count 0 and already hold their perft (a live board without
        // moves holds 0), so the leaf runs over the list of boards with
        // moves: its tiles are then full of work.  Without room for the list
        // the leaf takes every board and skips the others itself.
        const size_t mark = top;
        const u64 chunks = (n + LIVE_CHUNK - 1) / LIVE_CHUNK;
        u32 *chunkCount = n <= 0xFFFFFFFFULL ? (u32 *)alloc(chunks * sizeof(u32)) : nullptr;
        u64 *chunkOff = chunkCount ? (u64 *)alloc((chunks + 1) * sizeof(u64)) : nullptr;
        u32 *idx = nullptr;
        u64 nLive = n;
        if (chunkOff) {
            k_live_count<<<(u32)chunks, LIVE_WG, 0, stream>>>(cnt, nullptr, n, chunkCount);
            HIP_TRY(hipGetLastError());
            int rc = scan(chunkCount, chunks, chunkOff, &nLive);
            if (rc && rc != PG_ENOMEM) {
                top = mark;
                return rc;
            }
            idx = rc == PG_OK ? (u32 *)alloc(nLive * sizeof(u32) + 4) : nullptr;
            if (idx) {
                k_live_write<<<(u32)chunks, LIVE_WG, 0, stream>>>(cnt, nullptr, n, chunkOff, idx);
                HIP_TRY(hipGetLastError());
            } else {
                nLive = n;
            }
        }
        if (timing) HIP_TRY(hipEventRecord(ev0, stream));
        if (nLive) PG_VARIANT_CALL(v, launch_leaf_per, *this, B, cnt, nLive, P, idx);
        top = mark;
        if (ply == 0 && bias) k_add_base<<<1, 64, 0, stream>>>(P, n, bias, dOvf);
        HIP_TRY(hipGetLastError());
        if (timing) {
            HIP_TRY(hipEventRecord(ev1, stream));
            HIP_TRY(hipEventSynchronize(ev1));
            float ms = 0;
            HIP_TRY(hipEventElapsedTime(&ms, ev0, ev1));
            stats.leaf_kernel_ms += ms;
        }
        stats.leaf_launches++;
        stats.leaf_parents += n;
        return PG_OK;
    }
    if (r == 3 && v != SL_BYTE) return tt_last(v, tt, B, K, n, cnt, P, ply);
    if (n > 0xFFFFFFFFULL) return tt_split(v, tt, B, K, n, cnt, r, P, ply);
    const size_t mark = top;
    u64 *off = (u64 *)alloc((n + 1) * sizeof(u64));
    if (!off) { top = mark; return tt_split(v, tt, B, K, n, cnt, r, P, ply); }
    u64 m = 0;
    int rc = scan(cnt, n, off, &m);
    if (rc == PG_ENOMEM) { top = mark; return tt_split(v, tt, B, K, n, cnt, r, P, ply); }
    if (rc) return rc;
    if (m > 0xFFFFFFFFULL) { top = mark; return tt_split(v, tt, B, K, n, cnt, r, P, ply); }
    const int rcd = r - 1;  // children's remaining depth
    Pos *C = nullptr;
    Key128 *CK = nullptr;
    u64 *CP = nullptr;
    u32 *Ccnt = nullptr, *D = nullptr;
    u8 *S = nullptr;
    if (m) {
        C = (Pos *)alloc(m * sizeof(Pos));
        CK = C ? (Key128 *)alloc(m * sizeof(Key128)) : nullptr;
        CP = CK ? (u64 *)alloc(m * sizeof(u64)) : nullptr;
        Ccnt = CP ? (u32 *)alloc(m * sizeof(u32)) : nullptr;
        D = Ccnt ? (u32 *)alloc(m * sizeof(u32)) : nullptr;
        S = D ? (u8 *)alloc(m) : nullptr;
    }
    const size_t keep = top;
    u16 *mv = (m && S) ? (u16 *)alloc(m * sizeof(u16)) : nullptr;
    u32 *par = mv ? (u32 *)alloc(m * sizeof(u32)) : nullptr;
    const bool wantSet = dedup && rcd >= 2;
    const u64 setSlots = wantSet ? pow2_at_least(2 * m) : 0;
    u32 *set = (par && wantSet) ? (u32 *)alloc(setSlots * sizeof(u32)) : nullptr;
    if (m && (!par || (wantSet && !set))) {
        top = mark;
        return tt_split(v, tt, B, K, n, cnt, r, P, ply);
    }
    if (m) {
        if (ply + 1 < 16) stats.nodes[ply + 1] += m;
        PG_VARIANT_CALL(v, launch_gen, *this, B, n, off, mv, par);
        PG_VARIANT_CALL(v, launch_expand_tt, *this, tt, B, K, par, mv, m, C, CK, CP, Ccnt, S, D, rcd);
        if (set) {
            HIP_TRY(hipMemsetAsync(set, 0, setSlots * sizeof(u32), stream));
            k_dedup<<<grid(m, 256), 256, 0, stream>>>(C, CK, m, Ccnt, S, D, set, setSlots - 1, dErr, tt.dCtr, rcd);
        }
        HIP_TRY(hipGetLastError());
        top = keep;  // moves, parents and the set are no longer needed
        rc = tt_bfs(v, tt, C, CK, m, Ccnt, rcd, CP, ply + 1);
        if (rc) return rc;
        k_resolve_store<<<grid(m, 256), 256, 0, stream>>>(CK, m, CP, S, D, tt.view, rcd, tt.dCtr);
    }
    k_segsum<<<grid(n, 256), 256, 0, stream>>>(B, n, off, CP, P, ply == 0 ? bias : 0, dOvf);
    HIP_TRY(hipGetLastError());
    top = mark;
    return PG_OK;
}

// The boards of remaining depth 3 (B, n) and the children they expand to,
// which are never written as boards: per child its key, perft, owner, state,
// move and parent index (35 bytes, and 8 for the collapse set) instead of
// the 81 of a written board with its count.  k_expand_keys makes each child
// in registers for its key (and probes a table of remaining depth 2 if one
// exists), the collapse runs on the keys, the live children are listed by
// state, and the fused leaf makes each listed child from its parent and move,
// generates its moves in LDS and counts their moves: perft(2) per child.
// Collapsed children cost no move count, and the smaller level needs fewer
// splits, so more children are collapsed.
int Engine::tt_last(int v, TTable &tt, Pos *B, Key128 *K, u64 n, u32 *cnt, u64 *P, int ply) {
    const int r = 3, rcd = 2;
    if (n > 0xFFFFFFFFULL) return tt_split(v, tt, B, K, n, cnt, r, P, ply);
    const size_t mark = top;
    u64 *off = (u64 *)alloc((n + 1) * sizeof(u64));
    if (!off) { top = mark; return tt_split(v, tt, B, K, n, cnt, r, P, ply); }
    u64 m = 0;
    int rc = scan(cnt, n, off, &m);
    if (rc == PG_ENOMEM) { top = mark; return tt_split(v, tt, B, K, n, cnt, r, P, ply); }
    if (rc) return rc;
    if (m > 0xFFFFFFFFULL) { top = mark; return tt_split(v, tt, B, K, n, cnt, r, P, ply); }
    u64 *CP = nullptr;
    if (m) {
        const u64 chunks = (m + LIVE_CHUNK - 1) / LIVE_CHUNK;
        Key128 *CK = (Key128 *)alloc(m * sizeof(Key128));
        CP = CK ? (u64 *)alloc(m * sizeof(u64)) : nullptr;
        u32 *D = CP ? (u32 *)alloc(m * sizeof(u32)) : nullptr;
        u8 *S = D ? (u8 *)alloc(m) : nullptr;
        u16 *mv = S ? (u16 *)alloc(m * sizeof(u16)) : nullptr;
        u32 *par = mv ? (u32 *)alloc(m * sizeof(u32)) : nullptr;
        u32 *chunkCount = par ? (u32 *)alloc(chunks * sizeof(u32)) : nullptr;
        u64 *chunkOff = chunkCount ? (u64 *)alloc((chunks + 1) * sizeof(u64)) : nullptr;
        u32 *idx = chunkOff ? (u32 *)alloc(m * sizeof(u32) + 4) : nullptr;
        const size_t keep = top;
        const u64 setSlots = dedup ? pow2_at_least(2 * m) : 0;
        u32 *set = (idx && dedup) ? (u32 *)alloc(setSlots * sizeof(u32)) : nullptr;
        if (!idx || (dedup && !set)) {
            top = mark;
            return tt_split(v, tt, B, K, n, cnt, r, P, ply);
        }
        if (ply + 1 < 16) stats.nodes[ply + 1] += m;
        PG_VARIANT_CALL(v, launch_gen, *this, B, n, off, mv, par);
        PG_VARIANT_CALL(v, launch_expand_keys, *this, tt, B, K, par, mv, m, CK, CP, S, rcd);
        if (set) {
            HIP_TRY(hipMemsetAsync(set, 0, setSlots * sizeof(u32), stream));
            k_dedup<<<grid(m, 256), 256, 0, stream>>>(nullptr, CK, m, nullptr, S, D, set, setSlots - 1, dErr, tt.dCtr,
                                                       rcd);
        }
        HIP_TRY(hipGetLastError());
        top = keep;  // the set is no longer needed
        k_live_count<<<(u32)chunks, LIVE_WG, 0, stream>>>(nullptr, S, m, chunkCount);
        HIP_TRY(hipGetLastError());
        u64 nLive = 0;
        rc = scan(chunkCount, chunks, chunkOff, &nLive);
        if (rc) return rc;
        k_live_write<<<(u32)chunks, LIVE_WG, 0, stream>>>(nullptr, S, m, chunkOff, idx);
        HIP_TRY(hipGetLastError());
        if (timing) HIP_TRY(hipEventRecord(ev0, stream));
        if (nLive) PG_VARIANT_CALL(v, launch_leaf_moves_per, *this, B, par, mv, idx, nLive, CP);
        HIP_TRY(hipGetLastError());
        if (timing) {
            HIP_TRY(hipEventRecord(ev1, stream));
            HIP_TRY(hipEventSynchronize(ev1));
            float ms = 0;
            HIP_TRY(hipEventElapsedTime(&ms, ev0, ev1));
            stats.leaf_kernel_ms += ms;
        }
        stats.leaf_launches++;
        stats.leaf_parents += nLive;
        k_resolve_store<<<grid(m, 256), 256, 0, stream>>>(CK, m, CP, S, D, tt.view, rcd, tt.dCtr);
    }
    k_segsum<<<grid(n, 256), 256, 0, stream>>>(B, n, off, CP, P, ply == 0 ? bias : 0, dOvf);
    HIP_TRY(hipGetLastError());
    top = mark;
    return PG_OK;
}

int Engine::perft_tt(TTable &tt, const Pos &root, u32 depth, int v, u64 *out) {
    memset(&stats, 0, sizeof(stats));
    const auto t0 = std::chrono::steady_clock::now();
    top = 0;
    peak = 0;
    if (depth == 0) { *out = 1; return PG_OK; }
    const HostTables *ht = host_tables(nullptr);
    if (!ht) return PG_ETABLE;
    const Key128 key = zobrist128(root, ht->view());
    // device scratch: board, key, count, perft, probe result
    u8 *d = (u8 *)alloc(256);
    if (!d) return PG_ENOMEM;
    Pos *dB = (Pos *)d;
    Key128 *dK = (Key128 *)(d + 64);
    u32 *dCnt = (u32 *)(d + 96);
    u32 *dHit = (u32 *)(d + 100);
    u64 *dP = (u64 *)(d + 128);
    HIP_TRY(hipMemcpyAsync(dB, &root, sizeof(Pos), hipMemcpyHostToDevice, stream));
    HIP_TRY(hipMemcpyAsync(dK, &key, sizeof(Key128), hipMemcpyHostToDevice, stream));
    k_init_call<<<1, 64, 0, stream>>>(dP, 0, dItems, dErr, dOvf);
    k_init_call<<<1, 64, 0, stream>>>(dAcc, 0, nullptr, dErr, dOvf);
    u64 res = 0;
    bool rootHit = false;
    if (depth >= 2 && tt.view.base[depth] && bias == 0) {
        k_tt_single<<<1, 64, 0, stream>>>(tt.view, (int)depth, key, dP, dHit, 0);
        u32 hit = 0;
        HIP_TRY(hipMemcpyAsync(hostBuf, dHit, 4, hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipStreamSynchronize(stream));
        memcpy(&hit, hostBuf, 4);
        rootHit = hit != 0;
    }
    if (!rootHit) {
        PG_VARIANT_CALL(v, launch_cnt, *this, dB, 1, dCnt);
        int rc = tt_bfs(v, tt, dB, dK, 1, dCnt, (int)depth, dP, 0);
        if (rc) return rc;
        // (a biased test total is never stored: it is not the root's perft)
        if (depth >= 2 && bias == 0) k_tt_single<<<1, 64, 0, stream>>>(tt.view, (int)depth, key, dP, dHit, 1);
        HIP_TRY(hipGetLastError());
    }
    int rc = read_u64(dP, &res);
    if (rc) return rc;
    rc = check_flags("TT perft");
    if (rc) return rc;
    u64 items = 0;
    rc = read_u64(dItems, &items);
    if (rc) return rc;
    stats.leaf_items += items;
    stats.leaves = res;
    stats.peak_arena = peak;
    stats.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    *out = res;
    return PG_OK;
}

}  // namespace pg

// ====================================================================== C ABI
struct pg_tt : public pg::TTable {};
struct pg_ctx;  // == pg::Engine (engine.hip)

// Splits `bytes` over remaining depths 3..15: a quarter to depth 3, a quarter
// to depth 4 (16-byte entries), the other half shared by depths 5..15 (32-byte
// entries); each table a power of two of slots.  Remaining depth 2 gets no
// table: the in-level duplicate collapse already removes the transpositions
// within a level, and a hit there saves only one perft(2), less than its
// probe and store cost (start-position perft(11) with 64 GB: 24.1 s without
// it, 36.0 s with half the budget on it).
static void tt_plan(u64 bytes, int bits[TT_MAX_DEPTH], u64 sizes[TT_MAX_DEPTH]) {
    for (int r = 0; r < TT_MAX_DEPTH; r++) { bits[r] = 0; sizes[r] = 0; }
    auto fit = [](u64 budget, u64 entry) {
        int b = 0;
        while (b < 40 && (entry << (b + 1)) <= budget) b++;
        return b;
    };
    const u64 share[5] = {0, 0, 0, bytes / 4, bytes / 4};
    for (int r = 3; r <= TT_SHALLOW_MAX; r++) {
        bits[r] = fit(share[r], 16);
        sizes[r] = 16ull << bits[r];
    }
    const u64 deep = (bytes / 2) / (TT_MAX_DEPTH - 1 - TT_SHALLOW_MAX);
    for (int r = TT_SHALLOW_MAX + 1; r < TT_MAX_DEPTH; r++) {
        bits[r] = fit(deep, 32);
        sizes[r] = 32ull << bits[r];
    }
}

void pg::tt_free(TTable *t) {
    Engine *e = t->eng;
    (void)hipSetDevice(e->device);
    (void)hipStreamSynchronize(e->stream);
    if (e->batchSlot) slot_forget_tt(*e->batchSlot, t);
    for (size_t i = 0; i < e->tts.size(); i++)
        if (e->tts[i] == t) {
            e->tts.erase(e->tts.begin() + i);
            break;
        }
    (void)hipFree(t->mem);
    (void)hipFree(t->dCtr);
    delete static_cast<pg_tt *>(t);
}

extern "C" {

int pg_tt_create(pg_ctx *ctx, uint64_t bytes, pg_tt **out) {
    if (!ctx || !out || bytes < (1ull << 20)) return PG_EINVAL;
    *out = nullptr;
    Engine *e = reinterpret_cast<Engine *>(ctx);
    (void)hipSetDevice(e->device);
    pg_tt *t = new pg_tt();
    t->eng = e;
    u64 sizes[TT_MAX_DEPTH];
    tt_plan(bytes, t->bits, sizes);
    u64 total = 0;
    for (int r = 0; r < TT_MAX_DEPTH; r++) total += sizes[r];
    if (hipMalloc(&t->mem, total) != hipSuccess) {
        delete t;
        set_error("transposition table of %llu bytes does not fit", (unsigned long long)total);
        return PG_ENOMEM;
    }
    if (hipMalloc(&t->dCtr, sizeof(TTCounters)) != hipSuccess) {
        (void)hipFree(t->mem);
        delete t;
        return PG_ENOMEM;
    }
    t->bytes = total;
    u64 o = 0;
    for (int r = 0; r < TT_MAX_DEPTH; r++) {
        t->view.base[r] = sizes[r] ? t->mem + o : nullptr;
        t->view.mask[r] = sizes[r] ? (1ull << t->bits[r]) - 1 : 0;
        o += sizes[r];
    }
    // the reference never clears its tables (launcher.h:237-245); these are
    if (hipMemset(t->mem, 0, total) != hipSuccess || hipMemset(t->dCtr, 0, sizeof(TTCounters)) != hipSuccess) {
        (void)hipFree(t->mem);
        (void)hipFree(t->dCtr);
        delete t;
        return PG_EHIP;
    }
    e->tts.push_back(t);
    *out = t;
    return PG_OK;
}

int pg_tt_clear(pg_tt *t) {
    if (!t) return PG_EINVAL;
    (void)hipSetDevice(t->eng->device);
    if (hipMemset(t->mem, 0, t->bytes) != hipSuccess) return PG_EHIP;
    if (hipMemset(t->dCtr, 0, sizeof(TTCounters)) != hipSuccess) return PG_EHIP;
    return PG_OK;
}

int pg_tt_destroy(pg_tt *t) {
    if (!t) return PG_EINVAL;
    if (t->users > 0) {
        set_error("the table is shared by %d slot server(s): destroy them first", t->users);
        return PG_EINVAL;
    }
    tt_free(t);
    return PG_OK;
}

int pg_tt_info(pg_tt *t, uint64_t *bytes, int32_t *bits_per_depth /* [16] */) {
    if (!t) return PG_EINVAL;
    if (bytes) *bytes = t->bytes;
    if (bits_per_depth)
        for (int r = 0; r < TT_MAX_DEPTH; r++) bits_per_depth[r] = t->view.base[r] ? t->bits[r] : -1;
    return PG_OK;
}

int pg_tt_counters(pg_tt *t, uint64_t *out /* [16][4] */) {
    if (!t || !out) return PG_EINVAL;
    (void)hipSetDevice(t->eng->device);
    std::vector<TTCounters> h(1);
    if (hipMemcpy(h.data(), t->dCtr, sizeof(TTCounters), hipMemcpyDeviceToHost) != hipSuccess) return PG_EHIP;
    for (int r = 0; r < TT_MAX_DEPTH; r++)
        for (int k = 0; k < 4; k++) {
            uint64_t sum = 0;
            for (int c = 0; c < TT_CTR_COPIES; c++) sum += h[0].c[c][r][k];
            out[r * 4 + k] = sum;
        }
    return PG_OK;
}

int pg_perft_tt(pg_ctx *ctx, pg_tt *t, const pg_pos48 *pos, uint32_t depth, int slider, uint64_t *out) {
    if (!ctx || !t || !pos || !out || depth >= TT_MAX_DEPTH) return PG_EINVAL;
    Engine *e = reinterpret_cast<Engine *>(ctx);
    if (t->eng != e) return PG_EINVAL;
    const int v = slider == PG_SLIDER_AUTO ? e->slider : slider;
    if (v < 0 || v >= SL_COUNT) return PG_EINVAL;
    Pos p;
    memcpy(&p, pos, sizeof(p));
    if (!pos_sane(p)) {
        set_error("position needs exactly one king per side");
        return PG_EINVAL;
    }
    (void)hipSetDevice(e->device);
    return e->perft_tt(*t, p, depth, v, reinterpret_cast<u64 *>(out));
}

// Checkpoint of a table's contents: a 128-byte header (magic, bytes, the
// per-depth layout, a wrapping 64-bit sum of all words) followed by the raw
// entries, copied through one pinned staging buffer.  Entries verify their own
// key on every probe, so a table restored from any file can only miss, never
// return a wrong count; the sum still rejects a damaged or foreign file.
struct TTFileHeader {
    char magic[8];
    u64 bytes;
    int32_t bits[TT_MAX_DEPTH];
    u64 sum;
    u8 pad[128 - 8 - 8 - 4 * TT_MAX_DEPTH - 8];
};
static_assert(sizeof(TTFileHeader) == 128, "checkpoint header is 128 bytes");
static const char TT_MAGIC[8] = {'P', 'G', 'T', 'T', '0', '0', '0', '1'};
constexpr size_t TT_STAGE = 64ull << 20;

static int tt_copy_file(pg_tt *t, FILE *f, bool save, u64 *sum) {
    u8 *stage = nullptr;
    if (hipHostMalloc(&stage, TT_STAGE) != hipSuccess) return PG_ENOMEM;
    int rc = PG_OK;
    u64 s = 0;
    for (u64 o = 0; o < t->bytes && rc == PG_OK; o += TT_STAGE) {
        const size_t n = (size_t)std::min<u64>(TT_STAGE, t->bytes - o);
        if (save) {
            if (hipMemcpy(stage, t->mem + o, n, hipMemcpyDeviceToHost) != hipSuccess) rc = PG_EHIP;
            else if (fwrite(stage, 1, n, f) != n) rc = PG_EINVAL;
        } else {
            if (fread(stage, 1, n, f) != n) rc = PG_EINVAL;
            else if (hipMemcpy(t->mem + o, stage, n, hipMemcpyHostToDevice) != hipSuccess) rc = PG_EHIP;
        }
        const u64 *w = reinterpret_cast<const u64 *>(stage);
        for (size_t i = 0; i < n / 8; i++) s += w[i];
    }
    (void)hipHostFree(stage);
    *sum = s;
    return rc;
}

int pg_tt_save(pg_tt *t, const char *path) {
    if (!t || !path) return PG_EINVAL;
    (void)hipSetDevice(t->eng->device);
    if (hipStreamSynchronize(t->eng->stream) != hipSuccess) return PG_EHIP;
    FILE *f = fopen(path, "wb");
    if (!f) {
        set_error("cannot write %s", path);
        return PG_EINVAL;
    }
    TTFileHeader h{};
    memcpy(h.magic, TT_MAGIC, 8);
    h.bytes = t->bytes;
    for (int r = 0; r < TT_MAX_DEPTH; r++) h.bits[r] = t->view.base[r] ? t->bits[r] : -1;
    int rc = fwrite(&h, sizeof(h), 1, f) == 1 ? PG_OK : PG_EINVAL;
    u64 sum = 0;
    if (rc == PG_OK) rc = tt_copy_file(t, f, true, &sum);
    if (rc == PG_OK) {
        h.sum = sum;
        if (fseek(f, 0, SEEK_SET) != 0 || fwrite(&h, sizeof(h), 1, f) != 1) rc = PG_EINVAL;
    }
    if (fclose(f) != 0 && rc == PG_OK) rc = PG_EINVAL;
    if (rc) set_error("writing %s failed", path);
    return rc;
}

int pg_tt_load(pg_tt *t, const char *path) {
    if (!t || !path) return PG_EINVAL;
    (void)hipSetDevice(t->eng->device);
    if (hipStreamSynchronize(t->eng->stream) != hipSuccess) return PG_EHIP;
    FILE *f = fopen(path, "rb");
    if (!f) {
        set_error("cannot read %s", path);
        return PG_EINVAL;
    }
    TTFileHeader h{};
    int rc = fread(&h, sizeof(h), 1, f) == 1 ? PG_OK : PG_EINVAL;
    if (rc == PG_OK && (memcmp(h.magic, TT_MAGIC, 8) != 0 || h.bytes != t->bytes)) rc = PG_EINVAL;
    for (int r = 0; rc == PG_OK && r < TT_MAX_DEPTH; r++)
        if (h.bits[r] != (t->view.base[r] ? t->bits[r] : -1)) rc = PG_EINVAL;
    if (rc) {
        fclose(f);
        set_error("%s is not a checkpoint of a table with this layout", path);
        return rc;
    }
    u64 sum = 0;
    rc = tt_copy_file(t, f, false, &sum);
    fclose(f);
    if (rc == PG_OK && sum != h.sum) rc = PG_EINVAL;
    if (rc) {
        set_error("reading %s failed or its checksum does not match; the table is cleared", path);
        (void)hipMemset(t->mem, 0, t->bytes);
    }
    return rc;
}

int pg_tt_poison(pg_tt *t, uint64_t seed) {
    if (!t) return PG_EINVAL;
    Engine *e = t->eng;
    (void)hipSetDevice(e->device);
    const u64 n = t->bytes / 16;
    k_fill_random<<<e->grid(n, 256), 256, 0, e->stream>>>((ulonglong2 *)t->mem, n, seed);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(e->stream) != hipSuccess) return PG_EHIP;
    return PG_OK;
}

int pg_set_dedup(pg_ctx *ctx, int on) {
    if (!ctx) return PG_EINVAL;
    reinterpret_cast<Engine *>(ctx)->dedup = on != 0;
    return PG_OK;
}

}  // extern "C"
