// tt.h -- device transposition table for perft, and the kernels of the
// breadth-first perft that uses it.
//
// One HBM allocation, zeroed at creation, split into one table per remaining
// depth r (the reference keeps one table per depth too, launcher.h:69-123).
// Index = key.lo & mask; every entry verifies the full 128-bit key:
//  * shallow entries (r <= 4), 16 bytes, the reference's HashKey128b scheme
//    (perft_bb.h:890-892): w0 = (key.lo & ~mask) | perft, w1 = key.hi ^ w0.
//    The index bits hold the value, so only perft < 2^bits is stored.
//  * deep entries (r >= 5), 32 bytes (the reference packs 24, chess.h:474):
//    {key.lo ^ v, key.hi ^ v, v, ~(key.lo ^ key.hi ^ v)}.
// Writers use plain 16-byte stores without locks; a torn or concurrently
// overwritten entry fails the XOR check and reads as a miss, so a hit is
// always the value stored for that key.
#pragma once
#include "kernels.h"

namespace pg {

constexpr int TT_MAX_DEPTH = 16;
constexpr int TT_SHALLOW_MAX = 4;

struct TTView {
    u8 *base[TT_MAX_DEPTH];  // null = no table for that remaining depth
    u64 mask[TT_MAX_DEPTH];  // slots - 1
};

__device__ __forceinline__ bool tt_probe(const TTView &tt, int r, const Key128 &k, u64 *v) {
    u8 *b = tt.base[r];
    if (!b) return false;
    const u64 m = tt.mask[r];
    const u64 slot = k.lo & m;
    if (r <= TT_SHALLOW_MAX) {
        const ulonglong2 e = *reinterpret_cast<const ulonglong2 *>(b + slot * 16);
        if ((e.y ^ e.x) == k.hi && (e.x & ~m) == (k.lo & ~m)) {
            *v = e.x & m;
            return true;
        }
        return false;
    }
    const ulonglong2 *p = reinterpret_cast<const ulonglong2 *>(b + slot * 32);
    const ulonglong2 e0 = *reinterpret_cast<const ulonglong2 *>(p);
    const ulonglong2 e1 = *reinterpret_cast<const ulonglong2 *>(p + 1);
    const u64 val = e1.x;
    if ((e0.x ^ val) == k.lo && (e0.y ^ val) == k.hi && e1.y == ~(k.lo ^ k.hi ^ val)) {
        *v = val;
        return true;
    }
    return false;
}

__device__ __forceinline__ bool tt_store(const TTView &tt, int r, const Key128 &k, u64 v) {
    u8 *b = tt.base[r];
    if (!b) return false;
    const u64 m = tt.mask[r];
    const u64 slot = k.lo & m;
    if (r <= TT_SHALLOW_MAX) {
        if (v > m) return false;  // value does not fit the index bits
        ulonglong2 e;
        e.x = (k.lo & ~m) | v;
        e.y = k.hi ^ e.x;
        *reinterpret_cast<ulonglong2 *>(b + slot * 16) = e;
        return true;
    }
    ulonglong2 *p = reinterpret_cast<ulonglong2 *>(b + slot * 32);
    ulonglong2 e0, e1;
    e0.x = k.lo ^ v;
    e0.y = k.hi ^ v;
    e1.x = v;
    e1.y = ~(k.lo ^ k.hi ^ v);
    p[0] = e0;
    p[1] = e1;
    return true;
}


// counters: [copy][r][0] probes, [r][1] hits, [r][2] stores, [r][3]
// duplicates.  Work-groups add into copy blockIdx % TT_CTR_COPIES (512 bytes
// apart), so the waves of a large kernel do not all update the same few
// addresses -- atomics to one address serialize at its L2 channel and delay
// every other access that channel serves, including those of kernels on
// other streams.  Readers sum the copies.
constexpr int TT_CTR_COPIES = 64;
struct TTCounters {
    unsigned long long c[TT_CTR_COPIES][TT_MAX_DEPTH][4];
};

__device__ __forceinline__ void ctr_count(TTCounters *ctr, int r, int k, bool pred) {
    const u64 m = __ballot(pred);
    if ((threadIdx.x & 63) == __ffsll((long long)__ballot(1)) - 1 && m)
        atomicAdd(&ctr->c[blockIdx.x % TT_CTR_COPIES][r][k], (unsigned long long)__popcll(m));
}

// child = make(parent, move) with the incremental key; probe the table of the
// child's remaining depth rc; a hit resolves the child (its board is marked
// invalid so it expands to nothing), a miss counts the child's moves.
template <int V>
__global__ __launch_bounds__(wg_size<V>()) void k_expand_tt(const Pos *__restrict__ parents, const Key128 *__restrict__ pkeys,
                                                   const u32 *__restrict__ pidx, const u16 *__restrict__ moves, u64 m,
                                                   Pos *__restrict__ children, Key128 *__restrict__ ckeys,
                                                   u64 *__restrict__ cperft, u32 *__restrict__ cnt, u8 *__restrict__ state,
                                                   u32 *__restrict__ dup, TTView tt, int rc, TTCounters *ctr, Tables t) {
    stage_tables<V>(t);
    PG_GRID_LOOP(i, m) {
        const u32 pi = pidx[i];
        Pos p = parents[pi];
        Key128 h = pkeys[pi];
        make_any_hash(p, moves[i], h, t);
        u64 v = 0;
        const bool probe = rc >= 2 && tt.base[rc];
        const bool hit = probe && tt_probe(tt, rc, h, &v);
        if (ctr) {
            ctr_count(ctr, rc, 0, probe);
            ctr_count(ctr, rc, 1, hit);
        }
        if (hit) {
            p.white = 0;
            cnt[i] = 0;
            state[i] = ST_HIT;
        } else {
            cnt[i] = count_any<V>(p, t);
            state[i] = ST_LIVE;
        }
        children[i] = p;
        ckeys[i] = h;
        cperft[i] = v;
        dup[i] = (u32)i;
    }
}

// The last materialized level of a table search (children of remaining
// depth 2): each child is made in registers only, for its key; the board and
// its move count are not written.  Equal keys are collapsed on the keys
// alone, and the fused leaf makes the children that remain from their parent
// and move (k_leaf_moves_per), so a collapsed child costs no move count and
// no 48-byte board.
template <int V>
__global__ __launch_bounds__(wg_size<V>()) void k_expand_keys(const Pos *__restrict__ parents,
                                                     const Key128 *__restrict__ pkeys, const u32 *__restrict__ pidx,
                                                     const u16 *__restrict__ moves, u64 m, Key128 *__restrict__ ckeys,
                                                     u64 *__restrict__ cperft, u8 *__restrict__ state, TTView tt,
                                                     int rc, TTCounters *ctr, Tables t) {
    stage_tables<V>(t);
    PG_GRID_LOOP(i, m) {
        const u32 pi = pidx[i];
        Pos p = parents[pi];
        Key128 h = pkeys[pi];
        make_any_hash(p, moves[i], h, t);
        u64 v = 0;
        const bool probe = rc >= 2 && tt.base[rc];
        const bool hit = probe && tt_probe(tt, rc, h, &v);
        if (ctr) {
            ctr_count(ctr, rc, 0, probe);
            ctr_count(ctr, rc, 1, hit);
        }
        ckeys[i] = h;
        if (hit) cperft[i] = v;  // the leaf writes the live children's, the resolve the duplicates'
        state[i] = hit ? ST_HIT : ST_LIVE;
    }
}

PG_HD bool same_board(const Pos &a, const Pos &b) {
    const u64 clk = ~(0x7FULL << 56);  // the half-move clock is not part of a position
    return a.white == b.white && (a.pawns & clk) == (b.pawns & clk) && a.knights == b.knights &&
           a.bishopQueens == b.bishopQueens && a.rookQueens == b.rookQueens && a.kings == b.kings;
}

// In-level duplicate removal of the table search: each live child inserts
// its index into an open-addressing set keyed by its Zobrist key; a child
// already in the set with the same 128-bit key becomes its owner and the
// child is resolved from it later.  Equal keys stand for equal positions
// here exactly as they do for a table hit, which is accepted on the same
// 128-bit key (the searches without a table compare boards too: kd_dedup).
// The set has at least twice as many slots as children, so a probe
// sequence always meets an empty slot; the loop is also capped at the set
// size.
// one child after its first probe: follow the probe sequence until the child
// owns a slot or finds its equal
__device__ __forceinline__ void dedup_finish(u64 i, bool live, u64 s, u32 cur, Pos *__restrict__ children,
                                             const Key128 *__restrict__ ckeys, u32 *__restrict__ cnt,
                                             u8 *__restrict__ state, u32 *__restrict__ dup, u32 *set, u64 setMask,
                                             u32 *err, TTCounters *ctr, int rc) {
    bool isDup = false;
    if (live) {
        for (u64 step = 0;; step++) {
            if (cur == 0) break;  // claimed: this child owns the position
            const u32 o = cur - 1;
            if (ckeys[o].lo == ckeys[i].lo && ckeys[o].hi == ckeys[i].hi) {
                dup[i] = o;
                isDup = true;
                break;
            }
            if (step >= setMask) { *err = 3; break; }
            s = (s + 1) & setMask;
            cur = atomicCAS(&set[s], 0u, (u32)i + 1);
        }
    }
    if (ctr) ctr_count(ctr, rc, 3, isDup);
    if (isDup) {
        state[i] = ST_DUP;
        if (cnt) cnt[i] = 0;  // a level whose boards and counts were not written (k_expand_keys) has neither
        if (children) children[i].white = 0;
    }
}

// Each thread takes two children per step and issues both first probes
// before waiting on either, so twice as many random accesses are in flight.
static __global__ __launch_bounds__(256) void k_dedup(Pos *__restrict__ children, const Key128 *__restrict__ ckeys, u64 m,
                                               u32 *__restrict__ cnt, u8 *__restrict__ state, u32 *__restrict__ dup,
                                               u32 *set, u64 setMask, u32 *err, TTCounters *ctr, int rc) {
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i0 = (u64)blockIdx.x * blockDim.x + threadIdx.x; i0 < m; i0 += 2 * stride) {
        const u64 i1 = i0 + stride;
        const bool has1 = i1 < m;
        const bool l0 = state[i0] == ST_LIVE, l1 = has1 && state[i1] == ST_LIVE;
        const u64 s0 = (ckeys[i0].lo >> 7) & setMask, s1 = l1 ? (ckeys[i1].lo >> 7) & setMask : 0;
        const u32 c0 = l0 ? atomicCAS(&set[s0], 0u, (u32)i0 + 1) : 0;
        const u32 c1 = l1 ? atomicCAS(&set[s1], 0u, (u32)i1 + 1) : 0;
        dedup_finish(i0, l0, s0, c0, children, ckeys, cnt, state, dup, set, setMask, err, ctr, rc);
        if (has1) dedup_finish(i1, l1, s1, c1, children, ckeys, cnt, state, dup, set, setMask, err, ctr, rc);
    }
}

// After the children's perfts are known: duplicates copy their owner's value,
// live children are stored in the table.
static __global__ __launch_bounds__(256) void k_resolve_store(const Key128 *__restrict__ ckeys, u64 m, u64 *cperft,
                                                       const u8 *__restrict__ state, const u32 *__restrict__ dup,
                                                       TTView tt, int rc, TTCounters *ctr) {
    PG_GRID_LOOP(i, m) {
        const u8 st = state[i];
        bool stored = false;
        if (st == ST_DUP) cperft[i] = cperft[dup[i]];
        else if (st == ST_LIVE && rc >= 2 && tt.base[rc]) stored = tt_store(tt, rc, ckeys[i], cperft[i]);  // no loads without a table
        if (ctr) ctr_count(ctr, rc, 2, stored);
    }
}

// P[i] = base + sum of the children's perfts of board i (children are
// contiguous; base is 0 except under the count-bias test hook); boards
// resolved before expansion keep their value.  A sum that wraps raises *ovf.
static __global__ __launch_bounds__(256) void k_segsum(const Pos *__restrict__ boards, u64 n, const u64 *__restrict__ off,
                                                const u64 *__restrict__ cperft, u64 *__restrict__ P, u64 base, u32 *ovf) {
    PG_GRID_LOOP(i, n) {
        if (!pos_valid(boards[i])) continue;
        P[i] = sum_checked(cperft, off[i], off[i + 1], base, ovf);
    }
}

// probe (store = 0) or store (store = 1) a single key, e.g. the root
static __global__ void k_tt_single(TTView tt, int r, Key128 k, u64 *val, u32 *hit, int store) {
    if (threadIdx.x || blockIdx.x) return;
    if (store) {
        tt_store(tt, r, k, *val);
    } else {
        u64 v = 0;
        const bool h = tt_probe(tt, r, k, &v);
        *hit = h ? 1u : 0u;
        *val = v;
    }
}

// fills n 16-byte words with splitmix64 output (test hook: garbage entries)
static __global__ __launch_bounds__(256) void k_fill_random(ulonglong2 *p, u64 n, u64 seed) {
    PG_GRID_LOOP(i, n) {
        u64 z = seed + (2 * i + 1) * 0x9E3779B97F4A7C15ULL;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        u64 w = z ^ (z >> 31);
        z = w * 0x9E3779B97F4A7C15ULL + i;
        z = (z ^ (z >> 29)) * 0xBF58476D1CE4E5B9ULL;
        ulonglong2 e;
        e.x = w;
        e.y = z ^ (z >> 32);
        p[i] = e;
    }
}

// perft(1) of valid boards = their move counts (+ base, see k_segsum)
static __global__ __launch_bounds__(256) void k_perft1(const Pos *__restrict__ boards, u64 n, const u32 *__restrict__ cnt, u64 *P,
                                                u64 base, u32 *ovf) {
    PG_GRID_LOOP(i, n) {
        if (!pos_valid(boards[i])) continue;
        const u64 v = base + cnt[i];
        if (v < base) *ovf = 1;
        P[i] = v;
    }
}

}  // namespace pg
