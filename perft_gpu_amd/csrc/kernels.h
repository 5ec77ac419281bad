// kernels.h -- gfx950 kernels of the breadth-first perft.
//
// One level of the tree is expanded by four kernels (the reference's level
// loop, perft_bb.h:1071-1135, without device-side launches):
//   k_count     moves per board                        (countMoves)
//   scan        exclusive prefix sum of the counts     (ScanD, scan.cuh:329)
//   k_generate  move list + parent index per child     (generateMoves +
//               IntervalExpandDGenValues folded together, intervalmove.cuh:404)
//   k_expand    child board = make(parent, move), and the child's move count
//               (makemove_and_count_moves_single_level, perft_bb.h:601)
// and the last two plies by
//   k_leaf      sum over (parent, move) of countMoves(make(parent, move))
//               (makeMove_and_perft_single_level_indices, perft_bb.h:412)
// All kernels are grid-stride loops sized to the CU count, so one launch
// covers a level of any size; indices are 64-bit.
#pragma once
#include "movegen.h"

namespace pg {

constexpr int kByteLdsBytes = FANCY_TABLE_SIZE + (BYTE_ROOK_SIZE + BYTE_BISHOP_SIZE) * 8 + 2 * 64 * 16 + 2 * 64 * 8;

__device__ __forceinline__ u64 wave_sum(u64 v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// *acc += s, raising *ovf when the 64-bit total wraps (a perft count that
// does not fit u64 is reported, never returned wrapped).  A wave's own sum
// cannot wrap: it counts at most a few million moves per lane.
__device__ __forceinline__ void add_checked(u64 *acc, u64 s, u32 *ovf) {
    const u64 old = atomicAdd(acc, s);
    if (old + s < old) *ovf = 1;
}

// sum of v[lo..hi) with a carry check
__device__ __forceinline__ u64 sum_checked(const u64 *__restrict__ v, u64 lo, u64 hi, u64 base, u32 *ovf) {
    u64 s = base;
    bool carry = false;
    for (u64 j = lo; j < hi; j++) {
        const u64 x = v[j];
        s += x;
        carry |= s < x;
    }
    if (carry) *ovf = 1;
    return s;
}

// per-call device state reset without memset nodes: total (with an optional
// starting bias), consumed leaf items, error and overflow flags
static __global__ void k_init_call(u64 *acc, u64 bias, u64 *items, u32 *err, u32 *ovf) {
    if (threadIdx.x || blockIdx.x) return;
    acc[0] = bias;
    acc[1] = 0;
    if (items) *items = 0;
    *err = 0;
    *ovf = 0;
}

// P[i] += base for n values, raising *ovf on a wrap (the count-bias test
// hook where a root's perft is not produced by a checked sum)
static __global__ void k_add_base(u64 *P, u64 n, u64 base, u32 *ovf) {
    for (u64 i = threadIdx.x; i < n; i += blockDim.x) {
        const u64 v = P[i] + base;
        if (v < base) *ovf = 1;
        P[i] = v;
    }
}

template <class T>
__device__ __forceinline__ void lds_copy(T *dst, const T *src, int count) {
    // 16-byte chunks; every table here is a multiple of 16 bytes
    const uint4 *s = reinterpret_cast<const uint4 *>(src);
    uint4 *d = reinterpret_cast<uint4 *>(dst);
    const int chunks = (int)(count * sizeof(T) / 16);
    for (int i = threadIdx.x; i < chunks; i += blockDim.x) d[i] = s[i];
}

// The byte-lookup variant keeps its whole working set (147 KiB) in LDS; the
// other variants read their tables through the L1/L2 caches.
template <int V>
__device__ __forceinline__ void stage_tables(Tables &t) {
    if constexpr (V == SL_BYTE) {
        __shared__ __align__(16) u8 sIdx[FANCY_TABLE_SIZE];
        __shared__ __align__(16) u64 sRook[BYTE_ROOK_SIZE];
        __shared__ __align__(16) u64 sBishop[BYTE_BISHOP_SIZE];
        __shared__ __align__(16) FancyEntry sRF[64], sBF[64];
        __shared__ __align__(16) u64 sRM[64], sBM[64];
        lds_copy(sIdx, t.byteIdx, FANCY_TABLE_SIZE);
        lds_copy(sRook, t.byteRook, BYTE_ROOK_SIZE);
        lds_copy(sBishop, t.byteBishop, BYTE_BISHOP_SIZE);
        lds_copy(sRF, t.rookFancy, 64);
        lds_copy(sBF, t.bishopFancy, 64);
        lds_copy(sRM, t.rookMask, 64);
        lds_copy(sBM, t.bishopMask, 64);
        __syncthreads();
        t.byteIdx = sIdx;
        t.byteRook = sRook;
        t.byteBishop = sBishop;
        t.rookFancy = sRF;
        t.bishopFancy = sBF;
        t.rookMask = sRM;
        t.bishopMask = sBM;
    }
}

template <int V> constexpr int wg_size() { return V == SL_BYTE ? 1024 : 256; }

#define PG_GRID_LOOP(i, n) \
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += (u64)gridDim.x * blockDim.x)

// number of legal moves of a board after one move, side-resolved once
template <int V>
__device__ __forceinline__ u32 count_after(Pos p, u16 mv, const Tables &t) {
    if (pos_side(p)) {
        make_move<BLACK, false>(p, mv, nullptr, t);
        return count_moves<WHITE, V>(p, t);
    }
    make_move<WHITE, false>(p, mv, nullptr, t);
    return count_moves<BLACK, V>(p, t);
}
// the same with the moving piece known
template <int V>
__device__ __forceinline__ u32 count_after_as(Pos p, u16 mv, int piece, const Tables &t) {
    if (pos_side(p)) {
        make_move_as<BLACK, false>(p, mv, piece, nullptr, t);
        return count_moves<WHITE, V>(p, t);
    }
    make_move_as<WHITE, false>(p, mv, piece, nullptr, t);
    return count_moves<BLACK, V>(p, t);
}

template <int V>
__global__ __launch_bounds__(wg_size<V>()) void k_count(const Pos *__restrict__ b, u64 n, u32 *__restrict__ cnt, Tables t) {
    stage_tables<V>(t);
    PG_GRID_LOOP(i, n) cnt[i] = count_any<V>(b[i], t);
}

template <int V>
__global__ __launch_bounds__(wg_size<V>()) void k_count_sum(const Pos *__restrict__ b, u64 n, u64 *acc, u32 *ovf, Tables t) {
    stage_tables<V>(t);
    u64 s = 0;
    PG_GRID_LOOP(i, n) s += count_any<V>(b[i], t);
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0 && s) add_checked(acc, s, ovf);
}

static __global__ __launch_bounds__(256) void k_sum_u32(const u32 *__restrict__ c, u64 n, u64 *acc, u32 *ovf) {
    u64 s = 0;
    PG_GRID_LOOP(i, n) s += c[i];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0 && s) add_checked(acc, s, ovf);
}

// Move lists and parent indices of boards [0, n) whose exclusive child
// offsets are off[0..n].  A work-group takes 256 consecutive boards, whose
// children form one contiguous output range: the boards generate into an LDS
// window (move, and owner | piece tag), then the work-group stores the window
// one lane per slot, so the HBM writes are coalesced and each byte is written
// once.  Boards that do not fit the window wait for the next pass.  err is
// set to 1 if a board generates a different number of moves than it counted
// (the two must agree for the level layout to be valid).
constexpr int GEN_WG = 256, GEN_WINDOW = 8192;

template <int V>
__device__ __forceinline__ void generate_tiles(const Pos *__restrict__ b, u64 n, const u64 *__restrict__ off,
                                               u16 *__restrict__ moves, u32 *__restrict__ parent, u32 *err,
                                               const Tables &t) {
    __shared__ u16 smv[GEN_WINDOW];
    __shared__ u16 stag[GEN_WINDOW];
    __shared__ u32 snext;
    const int tid = threadIdx.x;
    for (u64 tile = blockIdx.x; tile * GEN_WG < n; tile += gridDim.x) {
        const u64 first = tile * GEN_WG;
        const u64 last = first + GEN_WG < n ? first + GEN_WG : n;
        const u64 i = first + tid;
        const bool live = i < n;
        const u64 base = off[first];
        const u32 pos = live ? (u32)(off[i] - base) : 0;
        const u32 c = live ? (u32)(off[i + 1] - off[i]) : 0;  // 0: no moves, or resolved earlier
        const u32 total = (u32)(off[last] - base);
        u32 start = 0;
        while (start < total) {
            const u32 end = start + GEN_WINDOW;
            if (tid == 0) snext = total;
            __syncthreads();
            if (c && pos >= start && pos + c <= end) {
                const u32 g = generate_any<V>(b[i], t, smv + (pos - start), c, stag + (pos - start), (u32)tid);
                if (g != c) *err = 1;
            } else if (c && pos >= start) {
                atomicMin(&snext, pos);
            }
            __syncthreads();
            const u32 wend = snext;
            for (u32 j = start + tid; j < wend; j += GEN_WG) {
                moves[base + j] = smv[j - start];
                parent[base + j] = (u32)first + (stag[j - start] & 255u);
            }
            __syncthreads();
            start = wend;
        }
    }
}

template <int V>
__global__ __launch_bounds__(wg_size<V>()) void k_generate(const Pos *__restrict__ b, u64 n, const u64 *__restrict__ off,
                                                           u16 *__restrict__ moves, u32 *__restrict__ parent, u32 *err,
                                                           Tables t) {
    stage_tables<V>(t);
    if constexpr (V == SL_BYTE) {
        // the byte-lookup tables fill the LDS: each lane writes its own list
        PG_GRID_LOOP(i, n) {
            const u64 o = off[i];
            const u32 room = (u32)(off[i + 1] - o);
            if (room == 0) continue;
            const u32 c = generate_any<V>(b[i], t, moves + o, room);
            if (c != room) *err = 1;
            for (u32 j = 0; j < room; j++) parent[o + j] = (u32)i;
        }
    } else {
        generate_tiles<V>(b, n, off, moves, parent, err, t);
    }
}

template <int V>
__global__ __launch_bounds__(wg_size<V>()) void k_expand(const Pos *__restrict__ parents, const u32 *__restrict__ pidx,
                                                         const u16 *__restrict__ moves, u64 m, Pos *__restrict__ children,
                                                         u32 *__restrict__ cnt, Tables t) {
    stage_tables<V>(t);
    PG_GRID_LOOP(i, m) {
        Pos p = parents[pidx[i]];
        make_any(p, moves[i], t);
        children[i] = p;
        cnt[i] = count_any<V>(p, t);
    }
}

template <int V>
__global__ __launch_bounds__(wg_size<V>()) void k_leaf(const Pos *__restrict__ parents, const u32 *__restrict__ pidx,
                                                       const u16 *__restrict__ moves, u64 m, u64 *acc, u32 *ovf, Tables t) {
    stage_tables<V>(t);
    u64 s = 0;
    PG_GRID_LOOP(i, m) s += count_after<V>(parents[pidx[i]], moves[i], t);
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0 && s) add_checked(acc, s, ovf);
}

// ---------------------------------------------------------------- scan helpers
constexpr int SCAN_WG = 256, SCAN_IPT = 8, SCAN_TILE = SCAN_WG * SCAN_IPT;

__device__ __forceinline__ u64 wave_incl_scan(u64 v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const u64 y = __shfl_up(v, o, 64);
        if (lane >= o) v += y;
    }
    return v;
}

// exclusive scan across the work-group; *total gets the group sum
template <int WG>
__device__ __forceinline__ u64 block_excl_scan(u64 v, u64 *total) {
    __shared__ u64 waveTot[WG / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const u64 incl = wave_incl_scan(v);
    if (lane == 63) waveTot[w] = incl;
    __syncthreads();
    u64 base = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < WG / 64; i++) {
        const u64 x = waveTot[i];
        if (i < w) base += x;
        tot += x;
    }
    __syncthreads();
    *total = tot;
    return base + incl - v;
}

// ---------------------------------------------------------------- fused leaf
// The last two plies without touching HBM for the (parent, move) items: a
// work-group takes 256 boards of ply D-2 and their move counts, scans the
// counts in LDS, generates the moves into an LDS window of leaf_cap slots
// (boards that do not fit wait for the next window) and counts the legal
// moves of every child.  PER = also write each board's perft(2).
// The window is as large as four work-groups per CU allow (160 KB of LDS:
// at most 40 KB each with the 12 KB of boards, 4 bytes per slot, and 1 KB of
// per-board sums when PER): 256 boards average about 7000 children, so most
// tiles then take one pass.
constexpr int LEAF_WG = 256, LEAF_BLOCKS_PER_CU = 64;
template <bool PER> constexpr int leaf_cap() { return PER ? 6848 : 7040; }

// state of a board after expansion (TT and batch searches)
enum : u8 { ST_LIVE = 0, ST_HIT = 1, ST_DUP = 2 };

// The fused leaf over boards [0, n) with move counts cnt[]: PER = false adds
// every count to *acc; PER = true writes base + perft(2) to per[i] for live
// boards (state[i] == ST_LIVE when state is given, else a valid board).
// With idx, the boards are idx[0..n) (live boards with moves, listed by
// k_live_write) and per[] is written at those indices.
// MOVES: the n boards are not stored: board j is boards[par[j]] after move
// mv[j] (the children of the level before, as the generate kernel lists
// them); the work-group makes each one and counts its moves itself, so the
// level one above the leaf never goes to HBM.  With idx as well, the boards
// made are those listed (j = idx[t]), and with PER their sums go to per[j].
template <int V, bool PER, bool MOVES = false>
__device__ __forceinline__ void leaf_tiles(const Pos *__restrict__ boards, const u32 *__restrict__ cnt, u64 n, u64 *acc,
                                           u64 *items, u64 *__restrict__ per, const u8 *__restrict__ state, u64 base,
                                           u32 *err, u32 *ovf, const Tables &t, const u32 *__restrict__ idx = nullptr,
                                           const u32 *__restrict__ par = nullptr, const u16 *__restrict__ mv = nullptr) {
    static_assert(V != SL_BYTE, "the byte-lookup tables leave no LDS for the move window");
    __shared__ Pos sb[LEAF_WG];
    constexpr int LEAF_CAP = leaf_cap<PER>();
    __shared__ u16 smv[LEAF_CAP];
    __shared__ u16 stag[LEAF_CAP];  // owner | moving piece << 8
    __shared__ u32 sper[PER ? LEAF_WG : 1];
    __shared__ u32 snext;
    const int tid = threadIdx.x;
    const Tables &tl = t;
    u64 s = 0, itemSum = 0;
    for (u64 tile = blockIdx.x; tile * LEAF_WG < n; tile += gridDim.x) {
        const u64 t0 = tile * LEAF_WG + tid;
        const bool live = t0 < n;
        const u64 i = idx && live ? (u64)idx[t0] : t0;
        u32 c = 0;
        if constexpr (MOVES) {
            if (live) {
                Pos b = boards[par[i]];
                make_any(b, mv[i], tl);
                c = count_any<V>(b, tl);
                sb[tid] = b;
            }
        } else {
            c = live ? cnt[i] : 0;
            if (live) sb[tid] = boards[i];
        }
        if (PER) sper[tid] = 0;
        u64 total64;
        const u32 pos = (u32)block_excl_scan<LEAF_WG>(c, &total64);
        const u32 total = (u32)total64;
        itemSum += total64;
        u32 start = 0;
        while (start < total) {
            const u32 end = start + LEAF_CAP;
            if (tid == 0) snext = total;
            __syncthreads();
            const bool here = c && pos >= start && pos + c <= end;
            if (here) {
                const u32 g = generate_any<V>(sb[tid], tl, smv + (pos - start), c, stag + (pos - start), (u32)tid);
                if (g != c) *err = 2;
            } else if (c && pos >= start) {
                atomicMin(&snext, pos);
            }
            __syncthreads();
            const u32 wend = snext;
            for (u32 j = start + tid; j < wend; j += LEAF_WG) {
                const u32 k = j - start;
                const u32 tg = stag[k];
                const int b = tg & 255;
                const u32 m = count_after_as<V>(sb[b], smv[k], (int)(tg >> 8), tl);
                if (PER) atomicAdd(&sper[b], m);
                else s += m;
            }
            __syncthreads();
            start = wend;
        }
        if (PER && live && (idx || (state ? state[i] == ST_LIVE : pos_valid(sb[tid])))) {
            const u64 v = base + sper[tid];
            if (v < base) *ovf = 1;
            per[i] = v;
        }
        __syncthreads();
    }
    if (tid == 0 && items && itemSum) atomicAdd(items, itemSum);
    if (!PER) {
        s = wave_sum(s);
        if ((tid & 63) == 0 && s) add_checked(acc, s, ovf);
    }
}

// Occupancy targets: four waves per SIMD for Kogge-Stone, five for the magic
// variants (their table lookups wait on L2, more waves hide it).  Some magic
// instantiations do not fit five and the compiler says so (-Wpass-failed);
// they end at four, and asking them for four measured no faster.
template <int V, bool PER>
__global__ __launch_bounds__(LEAF_WG) __attribute__((amdgpu_waves_per_eu(V == SL_KOGGE ? 4 : 5, 8))) void k_leaf_fused(
    const Pos *__restrict__ boards, const u32 *__restrict__ cnt, u64 n, u64 *acc, u64 *items, u64 *__restrict__ per,
    u32 *err, u32 *ovf, Tables t, const u32 *__restrict__ idx) {
    leaf_tiles<V, PER>(boards, cnt, n, acc, items, per, nullptr, 0, err, ovf, t, idx);
}

// the leaf over the children of `gp` listed by par/mv (see leaf_tiles, MOVES)
template <int V>
__global__ __launch_bounds__(LEAF_WG) __attribute__((amdgpu_waves_per_eu(V == SL_KOGGE ? 4 : 5, 8))) void k_leaf_moves(
    const Pos *__restrict__ gp, const u32 *__restrict__ par, const u16 *__restrict__ mv, u64 n, u64 *acc, u64 *items,
    u32 *err, u32 *ovf, Tables t) {
    leaf_tiles<V, false, true>(gp, nullptr, n, acc, items, nullptr, nullptr, 0, err, ovf, t, nullptr, par, mv);
}

// per-child sums over the children of `gp` listed by idx (see leaf_tiles,
// MOVES with idx): perft(2) of each listed child into per[]
template <int V>
__global__ __launch_bounds__(LEAF_WG) __attribute__((amdgpu_waves_per_eu(V == SL_KOGGE ? 4 : 5, 8))) void k_leaf_moves_per(
    const Pos *__restrict__ gp, const u32 *__restrict__ par, const u16 *__restrict__ mv, const u32 *__restrict__ idx,
    u64 n, u64 *items, u64 *__restrict__ per, u32 *err, u32 *ovf, Tables t) {
    leaf_tiles<V, true, true>(gp, nullptr, n, nullptr, items, per, nullptr, 0, err, ovf, t, idx, par, mv);
}

// ---------------------------------------------------------------- live list
// The boards of a leaf level that have moves, in index order: a level whose
// other boards were resolved from the table or collapsed into an equal board
// (count 0) then fills the fused leaf's tiles with boards that have work.
// Each work-group takes LIVE_CHUNK consecutive boards: k_live_count counts
// its live boards, the counts are scanned, and k_live_write lists them at
// the chunk's offset.  With state given, a live board is one whose state is
// ST_LIVE (a level whose boards were not written, so not counted either).
constexpr int LIVE_WG = 256, LIVE_PER_THREAD = 16, LIVE_CHUNK = LIVE_WG * LIVE_PER_THREAD;

__device__ __forceinline__ bool live_at(const u32 *__restrict__ cnt, const u8 *__restrict__ state, u64 i) {
    return state ? state[i] == ST_LIVE : cnt[i] != 0;
}

static __global__ __launch_bounds__(LIVE_WG) void k_live_count(const u32 *__restrict__ cnt, const u8 *__restrict__ state,
                                                              u64 n, u32 *__restrict__ chunkCount) {
    const u64 base = (u64)blockIdx.x * LIVE_CHUNK;
    u32 c = 0;
#pragma unroll
    for (int k = 0; k < LIVE_PER_THREAD; k++) {
        const u64 i = base + (u64)k * LIVE_WG + threadIdx.x;
        c += i < n && live_at(cnt, state, i);
    }
    u64 total;
    (void)block_excl_scan<LIVE_WG>(c, &total);
    if (threadIdx.x == 0) chunkCount[blockIdx.x] = (u32)total;
}

static __global__ __launch_bounds__(LIVE_WG) void k_live_write(const u32 *__restrict__ cnt, const u8 *__restrict__ state,
                                                              u64 n, const u64 *__restrict__ chunkOff,
                                                              u32 *__restrict__ idx) {
    const u64 base = (u64)blockIdx.x * LIVE_CHUNK;
    u64 at = chunkOff[blockIdx.x];
    for (int k = 0; k < LIVE_PER_THREAD; k++) {
        const u64 i = base + (u64)k * LIVE_WG + threadIdx.x;
        const bool f = i < n && live_at(cnt, state, i);
        u64 total;
        const u64 pos = block_excl_scan<LIVE_WG>(f ? 1 : 0, &total);
        if (f) idx[at + pos] = (u32)i;
        at += total;
    }
}

// ---------------------------------------------------------------- scan
// Exclusive scan of u32 counts into u64 offsets; off[n] receives the total.
// Reduce-then-scan over 2048-item tiles: k_scan_reduce (tile sums),
// k_scan_spine (one work-group scans the tile sums), k_scan_down (each tile
// rescans itself from its base).
__device__ __forceinline__ void load_tile(const u32 *c, u64 n, u64 base, u32 v[SCAN_IPT]) {
    if (base + SCAN_IPT <= n && (reinterpret_cast<uintptr_t>(c + base) & 15) == 0) {
        const uint4 a = *reinterpret_cast<const uint4 *>(c + base);
        const uint4 b = *reinterpret_cast<const uint4 *>(c + base + 4);
        v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
        v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
    } else {
#pragma unroll
        for (int j = 0; j < SCAN_IPT; j++) v[j] = base + j < n ? c[base + j] : 0;
    }
}

static __global__ __launch_bounds__(SCAN_WG) void k_scan_reduce(const u32 *__restrict__ c, u64 n, u64 *__restrict__ tileSum) {
    const u64 base = (u64)blockIdx.x * SCAN_TILE + (u64)threadIdx.x * SCAN_IPT;
    u32 v[SCAN_IPT];
    load_tile(c, n, base, v);
    u64 s = 0;
#pragma unroll
    for (int j = 0; j < SCAN_IPT; j++) s += v[j];
    u64 tot;
    block_excl_scan<SCAN_WG>(s, &tot);
    if (threadIdx.x == 0) tileSum[blockIdx.x] = tot;
}

static __global__ __launch_bounds__(1024) void k_scan_spine(u64 *tileSum, u64 tiles, u64 *total) {
    u64 carry = 0;
    for (u64 s = 0; s < tiles; s += 1024) {
        const u64 i = s + threadIdx.x;
        const u64 v = i < tiles ? tileSum[i] : 0;
        u64 tot;
        const u64 ex = block_excl_scan<1024>(v, &tot);
        if (i < tiles) tileSum[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) *total = carry;
}

static __global__ __launch_bounds__(SCAN_WG) void k_scan_down(const u32 *__restrict__ c, u64 n, const u64 *__restrict__ tileSum,
                                                      u64 *__restrict__ off) {
    const u64 base = (u64)blockIdx.x * SCAN_TILE + (u64)threadIdx.x * SCAN_IPT;
    u32 v[SCAN_IPT];
    load_tile(c, n, base, v);
    u64 s = 0;
#pragma unroll
    for (int j = 0; j < SCAN_IPT; j++) s += v[j];
    u64 tot;
    u64 run = tileSum[blockIdx.x] + block_excl_scan<SCAN_WG>(s, &tot);
#pragma unroll
    for (int j = 0; j < SCAN_IPT; j++) {
        if (base + j < n) off[base + j] = run;
        run += v[j];
    }
}

// ---------------------------------------------------------------- probes

template <int V>
__global__ __launch_bounds__(wg_size<V>()) void k_probe_generate(const Pos *__restrict__ b, u64 n, u16 *__restrict__ moves,
                                                                 u32 *__restrict__ cnt, Tables t) {
    stage_tables<V>(t);
    PG_GRID_LOOP(i, n) cnt[i] = generate_any<V>(b[i], t, moves + i * 256);
}

static __global__ __launch_bounds__(256) void k_probe_make(const Pos *__restrict__ b, const u16 *__restrict__ mv, u64 n,
                                                    Pos *__restrict__ out, Key128 *keys, Tables t) {
    PG_GRID_LOOP(i, n) {
        Pos p = b[i];
        if (keys) {
            Key128 h = keys[i];
            make_any_hash(p, mv[i], h, t);
            keys[i] = h;
        } else {
            make_any(p, mv[i], t);
        }
        out[i] = p;
    }
}

template <int V>
__global__ __launch_bounds__(wg_size<V>()) void k_probe_slider(int rook, const u8 *__restrict__ sq, const u64 *__restrict__ occ,
                                                               u64 n, u64 *__restrict__ out, Tables t) {
    stage_tables<V>(t);
    PG_GRID_LOOP(i, n) out[i] = rook ? Sl<V>::rook(t, sq[i], occ[i]) : Sl<V>::bishop(t, sq[i], occ[i]);
}

}  // namespace pg
