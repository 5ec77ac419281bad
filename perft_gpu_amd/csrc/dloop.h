// dloop.h -- device-driven level loop: the breadth-first perft of a batch of
// roots with no host round trip between levels, so the whole search can be
// captured once into a HIP graph and replayed.
//
// Sizes and array pointers of every level live in a Plan in device memory.
// Single-thread planning kernels read the scan total, carve the next level's
// arrays out of a bump region (stack discipline: transient move/parent/set
// arrays are released after expansion) and publish the level's size; the
// work kernels are persistent grid-stride loops that read their size from
// the plan.  If a level does not fit, the planner raises `overflow` and
// publishes size 0, every later kernel becomes a no-op, and the host reruns
// the batch in halves.  Per-root results come back up the levels by
// segmented sums (children of a board are contiguous), as in the TT path.
#pragma once
#include "tt.h"

namespace pg {

struct DLevel {
    u64 n;            // boards at this level
    Pos *b;
    Key128 *k;
    u32 *cnt;
    u64 *off;         // n + 1 exclusive offsets of the children
    u64 *perft;
    u8 *state;
    u32 *dup;
    u64 *tiles;       // scan tile sums
    u32 *live;        // leaf level: the boards with moves (kd_live), when there was room
    u16 *mv;          // moves / parents of the children: transient, or kept (keep) when
    u32 *par;         //   the children's boards are not stored (the leaf makes them)
    u32 *set;         // transient: duplicate set
    u64 setMask;
    u64 release;      // bump top to return to after expansion
    u32 keep;         // mv / par outlive the expansion
};

struct Plan {
    u8 *base;
    u64 cap, top;
    u32 overflow, err;  // overflow: a level did not fit the region
    u64 leafItems;
    u64 acc;
    u64 bias;           // test hook: starting value of every root's total
    u32 ovf;            // a 64-bit total wrapped
    u64 peak;           // highest bump top reached (bytes of the region used)
    u64 liveN;          // entries of the leaf level's live list
    DLevel lv[TT_MAX_DEPTH];
};

__device__ __forceinline__ void *plan_alloc(Plan *P, u64 bytes) {
    const u64 a = (P->top + 255) & ~255ULL;
    if (a + bytes > P->cap) {
        P->overflow = 1;
        return nullptr;
    }
    P->top = a + bytes;
    if (P->top > P->peak) P->peak = P->top;
    return P->base + a;
}

// as plan_alloc, for an optional array: no room leaves the plan as it was
__device__ __forceinline__ void *plan_try_alloc(Plan *P, u64 bytes) {
    const u64 a = (P->top + 255) & ~255ULL;
    if (a + bytes > P->cap) return nullptr;
    P->top = a + bytes;
    if (P->top > P->peak) P->peak = P->top;
    return P->base + a;
}

#define PG_GRID_LOOP_N(i, n) PG_GRID_LOOP(i, n)

// level-0 move counts (deeper levels get them from expansion)
template <int V>
__global__ __launch_bounds__(wg_size<V>()) void kd_count(Plan *P, Tables t) {
    stage_tables<V>(t);
    const DLevel L = P->lv[0];
    PG_GRID_LOOP(i, L.n) L.cnt[i] = count_any<V>(L.b[i], t);
}

// roots: probe the table of depth D; hits are resolved before expansion
static __global__ __launch_bounds__(256) void kd_root_probe(Plan *P, TTView tt, int D, TTCounters *ctr) {
    const DLevel L = P->lv[0];
    PG_GRID_LOOP(i, L.n) {
        u64 v = 0;
        const bool probe = D >= 2 && tt.base[D] && L.k;
        const bool hit = probe && tt_probe(tt, D, L.k[i], &v);
        if (ctr) {
            ctr_count(ctr, D, 0, probe);
            ctr_count(ctr, D, 1, hit);
        }
        // the roots are left untouched so that a rerun after an overflow sees
        // the same input; a hit is recorded in state/perft and expands to nothing
        L.state[i] = hit ? ST_HIT : ST_LIVE;
        L.perft[i] = v;
        if (hit) L.cnt[i] = 0;
    }
}

static __global__ __launch_bounds__(256) void kd_root_store(Plan *P, TTView tt, int D, TTCounters *ctr) {
    const DLevel L = P->lv[0];
    if (P->overflow || P->bias) return;  // a biased test total is not a perft
    PG_GRID_LOOP(i, L.n) {
        const bool st = D >= 2 && L.k && L.state[i] == ST_LIVE && tt_store(tt, D, L.k[i], L.perft[i]);
        if (ctr) ctr_count(ctr, D, 2, st);
    }
}

// release the transient arrays of level l (after its children are made)
__device__ __forceinline__ void release_level(Plan *P, int l) {
    DLevel &L = P->lv[l];
    if (P->lv[l + 1].n) P->top = L.release;
    if (!L.keep) {
        L.mv = nullptr;
        L.par = nullptr;
    }
    L.set = nullptr;
}

// release level l-1's transient arrays (l > 0), then allocate the offsets
// and scan tiles of level l
static __global__ void kd_plan_scan(Plan *P, int l) {
    if (threadIdx.x || blockIdx.x) return;
    if (l > 0) release_level(P, l - 1);
    DLevel &L = P->lv[l];
    L.off = nullptr;
    L.tiles = nullptr;
    if (L.n == 0) return;
    L.off = (u64 *)plan_alloc(P, (L.n + 1) * 8);
    L.tiles = L.off ? (u64 *)plan_alloc(P, ((L.n + SCAN_TILE - 1) / SCAN_TILE) * 8) : nullptr;
    if (!L.tiles) L.n = 0;
}

static __global__ __launch_bounds__(SCAN_WG) void kd_scan_reduce(Plan *P, int l) {
    const DLevel L = P->lv[l];
    const u64 tiles = (L.n + SCAN_TILE - 1) / SCAN_TILE;
    for (u64 tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        const u64 base = tile * SCAN_TILE + (u64)threadIdx.x * SCAN_IPT;
        u32 v[SCAN_IPT];
        load_tile(L.cnt, L.n, base, v);
        u64 s = 0;
#pragma unroll
        for (int j = 0; j < SCAN_IPT; j++) s += v[j];
        u64 tot;
        block_excl_scan<SCAN_WG>(s, &tot);
        if (threadIdx.x == 0) L.tiles[tile] = tot;
    }
}

__device__ void plan_children(Plan *P, int l, int wantKeys, int wantSet, int total);

// scan of the tile sums; then one thread plans level l+1 from the total
static __global__ __launch_bounds__(1024) void kd_scan_spine(Plan *P, int l, int wantKeys, int wantSet, int total) {
    const DLevel L = P->lv[l];
    if (L.n) {
        const u64 tiles = (L.n + SCAN_TILE - 1) / SCAN_TILE;
        u64 carry = 0;
        for (u64 s = 0; s < tiles; s += 1024) {
            const u64 i = s + threadIdx.x;
            const u64 v = i < tiles ? L.tiles[i] : 0;
            u64 tot;
            const u64 ex = block_excl_scan<1024>(v, &tot);
            if (i < tiles) L.tiles[i] = carry + ex;
            carry += tot;
        }
        if (threadIdx.x == 0) L.off[L.n] = carry;
    }
    if (threadIdx.x == 0) plan_children(P, l, wantKeys, wantSet, total);
}

static __global__ __launch_bounds__(SCAN_WG) void kd_scan_down(Plan *P, int l) {
    const DLevel L = P->lv[l];
    const u64 tiles = (L.n + SCAN_TILE - 1) / SCAN_TILE;
    for (u64 tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        const u64 base = tile * SCAN_TILE + (u64)threadIdx.x * SCAN_IPT;
        u32 v[SCAN_IPT];
        load_tile(L.cnt, L.n, base, v);
        u64 s = 0;
#pragma unroll
        for (int j = 0; j < SCAN_IPT; j++) s += v[j];
        u64 tot;
        u64 run = L.tiles[tile] + block_excl_scan<SCAN_WG>(s, &tot);
#pragma unroll
        for (int j = 0; j < SCAN_IPT; j++) {
            if (base + j < L.n) L.off[base + j] = run;
            run += v[j];
        }
    }
}

// allocate level l+1 (children of l) and the transient arrays; a total-only
// search keeps just boards and counts per level, and for its last expansion
// (total == 2) only the moves and parents: the leaf makes those children.
// total == 3: the last level of a search with per-board results and keys
// keeps no boards or counts, only keys, perfts, states and owners; the moves
// and parents stay until the leaf, which makes the children from them.
__device__ void plan_children(Plan *P, int l, int wantKeys, int wantSet, int total) {
    DLevel &L = P->lv[l];
    DLevel &C = P->lv[l + 1];
    memset(&C, 0, sizeof(C));
    L.keep = 0;
    const u64 m = L.n ? L.off[L.n] : 0;
    if (m == 0 || m > 0xFFFFFFFFULL) {
        if (m) P->overflow = 1;
        return;
    }
    if (total == 3) {
        C.perft = (u64 *)plan_alloc(P, m * 8);
        C.state = (u8 *)plan_alloc(P, m);
        C.dup = (u32 *)plan_alloc(P, m * 4);
        C.k = (Key128 *)plan_alloc(P, m * sizeof(Key128));
        L.mv = (u16 *)plan_alloc(P, m * 2);
        L.par = (u32 *)plan_alloc(P, m * 4);
        L.keep = 1;
        L.release = P->top;  // only the set goes after the expansion
        if (wantSet) {
            u64 s = 1;
            while (s < 2 * m) s <<= 1;
            L.set = (u32 *)plan_alloc(P, s * 4);
            L.setMask = s - 1;
        }
        if (P->overflow) {
            memset(&C, 0, sizeof(C));
            return;
        }
        C.n = m;
        return;
    }
    if (total == 2) {
        L.release = P->top;
        L.mv = (u16 *)plan_alloc(P, m * 2);
        L.par = (u32 *)plan_alloc(P, m * 4);
        if (!P->overflow) C.n = m;
        return;
    }
    C.b = (Pos *)plan_alloc(P, m * sizeof(Pos));
    C.cnt = (u32 *)plan_alloc(P, m * 4);
    if (!total) {
        C.perft = (u64 *)plan_alloc(P, m * 8);
        C.state = (u8 *)plan_alloc(P, m);
        C.dup = (u32 *)plan_alloc(P, m * 4);
    }
    if (wantKeys) C.k = (Key128 *)plan_alloc(P, m * sizeof(Key128));
    L.release = P->top;
    L.mv = (u16 *)plan_alloc(P, m * 2);
    L.par = (u32 *)plan_alloc(P, m * 4);
    if (wantSet) {
        u64 s = 1;
        while (s < 2 * m) s <<= 1;
        L.set = (u32 *)plan_alloc(P, s * 4);
        L.setMask = s - 1;
    }
    if (P->overflow) {
        memset(&C, 0, sizeof(C));
        return;
    }
    C.n = m;
}

template <int V>
__global__ __launch_bounds__(wg_size<V>()) void kd_generate(Plan *P, int l, Tables t) {
    stage_tables<V>(t);
    const DLevel L = P->lv[l];
    if (P->lv[l + 1].n == 0) return;
    if constexpr (V == SL_BYTE) {
        PG_GRID_LOOP(i, L.n) {
            const u64 o = L.off[i];
            const u32 room = (u32)(L.off[i + 1] - o);
            if (room == 0) continue;
            const u32 c = generate_any<V>(L.b[i], t, L.mv + o, room);
            if (c != room) P->err = 1;
            for (u32 j = 0; j < room; j++) L.par[o + j] = (u32)i;
        }
    } else {
        generate_tiles<V>(L.b, L.n, L.off, L.mv, L.par, &P->err, t);
    }
}

// children: make (with key), probe the child depth's table, count moves
template <int V>
__global__ __launch_bounds__(wg_size<V>()) void kd_expand(Plan *P, int l, TTView tt, int rc, TTCounters *ctr, Tables t) {
    stage_tables<V>(t);
    const DLevel L = P->lv[l];
    const DLevel C = P->lv[l + 1];
    PG_GRID_LOOP(i, C.n) {
        const u32 pi = L.par[i];
        Pos p = L.b[pi];
        u64 v = 0;
        bool hit = false;
        if (C.k) {
            Key128 h = L.k[pi];
            make_any_hash(p, L.mv[i], h, t);
            C.k[i] = h;
            const bool probe = rc >= 2 && tt.base[rc];
            hit = probe && tt_probe(tt, rc, h, &v);
            if (ctr) {
                ctr_count(ctr, rc, 0, probe);
                ctr_count(ctr, rc, 1, hit);
            }
        } else {
            make_any(p, L.mv[i], t);
        }
        C.cnt[i] = hit ? 0 : count_any<V>(p, t);
        C.b[i] = p;
        if (C.state) {  // absent in a total-only search
            C.state[i] = hit ? ST_HIT : ST_LIVE;
            C.perft[i] = v;
            C.dup[i] = (u32)i;
        }
    }
}

// the last level when it keeps keys only (plan_children, total == 3): each
// child is made in registers for its key and its table probe
template <int V>
__global__ __launch_bounds__(wg_size<V>()) void kd_expand_keys(Plan *P, int l, TTView tt, int rc, TTCounters *ctr,
                                                      Tables t) {
    stage_tables<V>(t);
    const DLevel L = P->lv[l];
    const DLevel C = P->lv[l + 1];
    PG_GRID_LOOP(i, C.n) {
        const u32 pi = L.par[i];
        Pos p = L.b[pi];
        Key128 h = L.k[pi];
        make_any_hash(p, L.mv[i], h, t);
        u64 v = 0;
        const bool probe = rc >= 2 && tt.base[rc];
        const bool hit = probe && tt_probe(tt, rc, h, &v);
        if (ctr) {
            ctr_count(ctr, rc, 0, probe);
            ctr_count(ctr, rc, 1, hit);
        }
        C.k[i] = h;
        C.state[i] = hit ? ST_HIT : ST_LIVE;
        if (hit) C.perft[i] = v;  // the leaf writes the live children's, the resolve the duplicates'
    }
}

// equal children of level l+1 when their boards are not stored: both made
// again from parent and move
__device__ __forceinline__ bool same_child(const DLevel &L, u64 a, u64 b, const Tables &t) {
    Pos pa = L.b[L.par[a]], pb = L.b[L.par[b]];
    make_any(pa, L.mv[a], t);
    make_any(pb, L.mv[b], t);
    return same_board(pa, pb);
}

// fill kernel for the duplicate set (graphs get no memset nodes)
static __global__ __launch_bounds__(256) void kd_clear_set(Plan *P, int l) {
    const DLevel L = P->lv[l];
    if (!L.set || P->lv[l + 1].n == 0) return;
    PG_GRID_LOOP(i, L.setMask + 1) L.set[i] = 0;
}

// one child after its first probe (kd_dedup): follow the probe sequence
// until the child owns a slot or finds its equal
__device__ __forceinline__ void kd_dedup_finish(Plan *P, const DLevel &L, const DLevel &C, u64 i, bool live, u64 s,
                                                u32 cur, TTCounters *ctr, int rc, int keyOnly, const Tables &t) {
    bool isDup = false;
    if (live) {
        for (u64 step = 0;; step++) {
            if (cur == 0) break;
            const u32 o = cur - 1;
            if (C.k[o].lo == C.k[i].lo && C.k[o].hi == C.k[i].hi &&
                (keyOnly || (C.b ? same_board(C.b[o], C.b[i]) : same_child(L, o, i, t)))) {
                C.dup[i] = o;
                isDup = true;
                break;
            }
            if (step >= L.setMask) { P->err = 3; break; }
            s = (s + 1) & L.setMask;
            cur = atomicCAS(&L.set[s], 0u, (u32)i + 1);
        }
    }
    if (ctr) ctr_count(ctr, rc, 3, isDup);
    if (isDup) {
        C.state[i] = ST_DUP;
        if (C.cnt) C.cnt[i] = 0;  // a level that keeps keys only has no counts
    }
}

// in-level duplicates of level l+1 (as k_dedup: two children per thread,
// both first probes in flight).  keyOnly (a search with a table, which
// accepts hits on the 128-bit key): equal keys are equal positions; without
// a table the boards are compared as well.
static __global__ __launch_bounds__(256) void kd_dedup(Plan *P, int l, TTCounters *ctr, int rc, int keyOnly, Tables t) {
    const DLevel L = P->lv[l];
    const DLevel C = P->lv[l + 1];
    if (!L.set) return;
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i0 = (u64)blockIdx.x * blockDim.x + threadIdx.x; i0 < C.n; i0 += 2 * stride) {
        const u64 i1 = i0 + stride;
        const bool has1 = i1 < C.n;
        const bool l0 = C.state[i0] == ST_LIVE, l1 = has1 && C.state[i1] == ST_LIVE;
        const u64 s0 = (C.k[i0].lo >> 7) & L.setMask, s1 = l1 ? (C.k[i1].lo >> 7) & L.setMask : 0;
        const u32 c0 = l0 ? atomicCAS(&L.set[s0], 0u, (u32)i0 + 1) : 0;
        const u32 c1 = l1 ? atomicCAS(&L.set[s1], 0u, (u32)i1 + 1) : 0;
        kd_dedup_finish(P, L, C, i0, l0, s0, c0, ctr, rc, keyOnly, t);
        if (has1) kd_dedup_finish(P, L, C, i1, l1, s1, c1, ctr, rc, keyOnly, t);
    }
}

// The live list of the leaf level l (as k_live_count / k_live_write, in one
// pass): boards resolved from the table or collapsed into an equal board
// have count 0, so the fused leaf runs over the boards with moves only.  A
// live board without moves gets its perft(2) = base here.  kd_plan_live
// makes room for the list (none: the leaf takes every board); each
// work-group lists a chunk of boards in order at an offset it reserves with
// one atomic, so chunks land in any order.
static __global__ void kd_plan_live(Plan *P, int l) {
    if (threadIdx.x || blockIdx.x) return;
    if (l > 0) release_level(P, l - 1);  // the moves, parents and set that made level l
    DLevel &L = P->lv[l];
    P->liveN = 0;
    L.live = L.n && L.n <= 0xFFFFFFFFULL ? (u32 *)plan_try_alloc(P, L.n * 4) : nullptr;
}

static __global__ __launch_bounds__(LIVE_WG) void kd_live(Plan *P, int l) {
    const DLevel L = P->lv[l];
    if (!L.live) return;
    __shared__ u64 sBase;
    const u64 base = l == 0 ? P->bias : 0;
    const u64 chunks = (L.n + LIVE_CHUNK - 1) / LIVE_CHUNK;
    for (u64 ch = blockIdx.x; ch < chunks; ch += gridDim.x) {
        // each thread takes LIVE_PER_THREAD consecutive boards, so the list
        // keeps the level's order (siblings stay neighbours in the leaf)
        const u64 first = ch * LIVE_CHUNK + (u64)threadIdx.x * LIVE_PER_THREAD;
        u32 f = 0, c = 0;
#pragma unroll
        for (int k = 0; k < LIVE_PER_THREAD; k++) {
            const u64 i = first + k;
            if (!L.cnt) {  // a level that keeps keys only: its live children, counted by the leaf
                const bool lv = i < L.n && L.state[i] == ST_LIVE;
                if (lv) f |= 1u << k;
                c += lv;
                continue;
            }
            const u32 n = i < L.n ? L.cnt[i] : 0;
            if (n) f |= 1u << k;
            c += n != 0;
            if (i < L.n && n == 0 && L.state[i] == ST_LIVE) L.perft[i] = base;
        }
        u64 total;
        const u64 pos = block_excl_scan<LIVE_WG>(c, &total);
        if (threadIdx.x == 0) sBase = total ? atomicAdd((unsigned long long *)&P->liveN, total) : 0;
        __syncthreads();
        u64 at = sBase + pos;
        for (; f; f &= f - 1) L.live[at++] = (u32)(first + __builtin_ctz(f));
        __syncthreads();
    }
}

// perft(2) of the boards of level l, fused (see leaf_tiles): per board into
// the level's perft array, or (TOTAL, a single-root search that only needs
// the sum) added to the plan's total
template <int V, bool TOTAL>
__global__ __launch_bounds__(LEAF_WG) __attribute__((amdgpu_waves_per_eu(V == SL_KOGGE ? 4 : 5, 8))) void kd_leaf(Plan *P, int l, Tables t) {
    const DLevel L = P->lv[l];
    const bool listed = !TOTAL && L.live;
    leaf_tiles<V, !TOTAL>(L.b, L.cnt, listed ? P->liveN : L.n, &P->acc, &P->leafItems, L.perft, L.state,
                          l == 0 ? P->bias : 0, &P->err, &P->ovf, t, listed ? L.live : nullptr);
}

// perft(2) of the children of level l-1 that form level l when level l keeps
// keys only: the leaf makes each listed child from its parent and move (or,
// without room for the list, every child, writing the live ones)
template <int V>
__global__ __launch_bounds__(LEAF_WG) __attribute__((amdgpu_waves_per_eu(V == SL_KOGGE ? 4 : 5, 8))) void kd_leaf_keys(
    Plan *P, int l, Tables t) {
    const DLevel L = P->lv[l];
    const DLevel A = P->lv[l - 1];
    const bool listed = L.live != nullptr;
    leaf_tiles<V, true, true>(A.b, nullptr, listed ? P->liveN : L.n, &P->acc, &P->leafItems, L.perft,
                              listed ? nullptr : L.state, 0, &P->err, &P->ovf, t, listed ? L.live : nullptr, A.par,
                              A.mv);
}

// perft(3) of the boards of level l summed into the plan's total: the
// children of level l (its move list) are made, counted and expanded inside
// the fused leaf (leaf_tiles, MOVES)
template <int V>
__global__ __launch_bounds__(LEAF_WG) __attribute__((amdgpu_waves_per_eu(V == SL_KOGGE ? 4 : 5, 8))) void kd_leaf_moves(
    Plan *P, int l, Tables t) {
    const DLevel L = P->lv[l];
    leaf_tiles<V, false, true>(L.b, nullptr, P->lv[l + 1].n, &P->acc, &P->leafItems, nullptr, nullptr, 0, &P->err,
                               &P->ovf, t, nullptr, L.par, L.mv);
}

// the plan's total += the move counts of level l (perft(1) of a total-only search)
static __global__ __launch_bounds__(256) void kd_sum_counts(Plan *P, int l) {
    const DLevel L = P->lv[l];
    u64 s = 0;
    PG_GRID_LOOP(i, L.n) s += L.cnt[i];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0 && s) add_checked(&P->acc, s, &P->ovf);
}

static __global__ __launch_bounds__(256) void kd_perft1(Plan *P, int l) {
    const DLevel L = P->lv[l];
    const u64 base = l == 0 ? P->bias : 0;
    PG_GRID_LOOP(i, L.n) {
        if (L.state[i] != ST_LIVE) continue;
        const u64 v = base + L.cnt[i];
        if (v < base) P->ovf = 1;
        L.perft[i] = v;
    }
}

// level l+1 after its perfts are known: duplicates take their owner's value,
// live children are stored (remaining depth rc)
static __global__ __launch_bounds__(256) void kd_resolve(Plan *P, int l, TTView tt, int rc, TTCounters *ctr) {
    const DLevel C = P->lv[l + 1];
    if (P->overflow) return;
    PG_GRID_LOOP(i, C.n) {
        const u8 st = C.state[i];
        bool stored = false;
        if (st == ST_DUP) C.perft[i] = C.perft[C.dup[i]];
        else if (st == ST_LIVE && C.k && rc >= 2 && tt.base[rc]) stored = tt_store(tt, rc, C.k[i], C.perft[i]);
        if (ctr) ctr_count(ctr, rc, 2, stored);
    }
}

static __global__ __launch_bounds__(256) void kd_segsum(Plan *P, int l) {
    const DLevel L = P->lv[l];
    const DLevel C = P->lv[l + 1];
    if (P->overflow) return;
    const u64 base = l == 0 ? P->bias : 0;
    PG_GRID_LOOP(i, L.n) {
        if (L.state[i] != ST_LIVE) continue;
        L.perft[i] = sum_checked(C.perft, L.off[i], L.off[i + 1], base, &P->ovf);
    }
}

}  // namespace pg
