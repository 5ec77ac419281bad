// engine.hip -- device context, tree arena and the breadth-first perft
// driver behind the C ABI (include/perft_gpu.h).
//
// Replaces the reference's CDP drivers (perft_bb_driver_gpu /
// perft_bb_gpu_simple, perft_bb.h:1027-1220) with a host-driven level loop on
// one HIP stream.  The tree lives in one pre-reserved HBM arena used as a
// stack (the reference bump-allocates a 3 GiB buffer with 32-bit offsets,
// perft_bb.h:113-139); a level whose children do not fit is split in halves
// and each half expanded on its own (the reference instead fixes a launch
// depth, perft.cu:117-120).
#include <hip/hip_runtime.h>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>
#include "../../include/perft_gpu.h"
#include "engine.h"
#include "kernels.h"

using namespace pg;

namespace pg {

thread_local char g_last_error[512];
void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
}

static double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

void *Engine::alloc(size_t bytes) {
    size_t a = (top + 255) & ~(size_t)255;
    if (a + bytes > arenaBytes) return nullptr;
    top = a + bytes;
    if (top > peak) peak = top;
    return arena + a;
}

u32 Engine::grid(u64 n, int wg) const {
    u64 want = (n + wg - 1) / wg;
    u64 cap = (u64)cus * (wg >= 1024 ? 1 : (u64)bpc);
    if (want > cap) want = cap;
    return (u32)(want ? want : 1);
}

// The fused leaf's tiles differ in work by their boards' move counts; with
// many more work-groups than fit at once the hardware hands tiles to CUs as
// they free up, which balances the tail (start perft(8): 51.7 ms with 8
// work-groups per CU, 47.0 ms with 64).
u32 Engine::leaf_grid(u64 n) const {
    u64 want = (n + LEAF_WG - 1) / LEAF_WG;
    const u64 cap = (u64)cus * LEAF_BLOCKS_PER_CU;
    if (want > cap) want = cap;
    return (u32)(want ? want : 1);
}

int Engine::read_u64(const u64 *d, u64 *h) {
    HIP_TRY(hipMemcpyAsync(hostBuf, d, sizeof(u64), hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    *h = *hostBuf;
    return PG_OK;
}

// ---------------------------------------------------------------- dispatch
template <int V> static void launch_count(Engine &c, const Pos *b, u64 n, u32 *cnt) {
    constexpr int W = wg_size<V>();
    k_count<V><<<c.grid(n, W), W, 0, c.stream>>>(b, n, cnt, c.dt);
}
template <int V> static void launch_count_sum(Engine &c, const Pos *b, u64 n, u64 *acc) {
    constexpr int W = wg_size<V>();
    k_count_sum<V><<<c.grid(n, W), W, 0, c.stream>>>(b, n, acc, c.dOvf, c.dt);
}
template <int V> static void launch_generate(Engine &c, const Pos *b, u64 n, const u64 *off, u16 *mv, u32 *par) {
    constexpr int W = wg_size<V>();
    k_generate<V><<<c.grid(n, W), W, 0, c.stream>>>(b, n, off, mv, par, c.dErr, c.dt);
}
template <int V> static void launch_expand(Engine &c, const Pos *b, const u32 *par, const u16 *mv, u64 m, Pos *ch, u32 *cnt) {
    constexpr int W = wg_size<V>();
    k_expand<V><<<c.grid(m, W), W, 0, c.stream>>>(b, par, mv, m, ch, cnt, c.dt);
}
template <int V> static void launch_leaf(Engine &c, const Pos *b, const u32 *par, const u16 *mv, u64 m, u64 *acc) {
    constexpr int W = wg_size<V>();
    k_leaf<V><<<c.grid(m, W), W, 0, c.stream>>>(b, par, mv, m, acc, c.dOvf, c.dt);
}

int Engine::scan(const u32 *cnt, u64 n, u64 *off, u64 *total) {
    const u64 tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
    const size_t mark = top;
    u64 *tileSum = (u64 *)alloc(tiles * sizeof(u64));
    if (!tileSum) return PG_ENOMEM;
    k_scan_reduce<<<(u32)tiles, SCAN_WG, 0, stream>>>(cnt, n, tileSum);
    k_scan_spine<<<1, 1024, 0, stream>>>(tileSum, tiles, off + n);
    k_scan_down<<<(u32)tiles, SCAN_WG, 0, stream>>>(cnt, n, tileSum, off);
    HIP_TRY(hipGetLastError());
    int rc = read_u64(off + n, total);
    top = mark;
    return rc;
}

int Engine::leaf(int v, const Pos *b, u64 n, const u32 *par, const u16 *mv, u64 m) {
    if (timing) HIP_TRY(hipEventRecord(ev0, stream));
    PG_VARIANT_CALL(v, launch_leaf, *this, b, par, mv, m, dAcc);
    HIP_TRY(hipGetLastError());
    if (timing) {
        HIP_TRY(hipEventRecord(ev1, stream));
        HIP_TRY(hipEventSynchronize(ev1));
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, ev0, ev1));
        stats.leaf_kernel_ms += ms;
    }
    stats.leaf_launches++;
    stats.leaf_items += m;
    stats.leaf_parents += n;
    return PG_OK;
}

template <int V> static void launch_leaf_fused(Engine &c, const Pos *b, const u32 *cnt, u64 n) {
    if constexpr (V == SL_BYTE) {
        (void)c; (void)b; (void)cnt; (void)n;
    } else {
        k_leaf_fused<V, false><<<c.leaf_grid(n), LEAF_WG, 0, c.stream>>>(b, cnt, n, c.dAcc, c.dItems, nullptr, c.dErr,
                                                                         c.dOvf, c.dt, nullptr);
    }
}

template <int V>
static void launch_leaf_moves(Engine &c, const Pos *gp, const u32 *par, const u16 *mv, u64 n) {
    if constexpr (V == SL_BYTE) {
        (void)c; (void)gp; (void)par; (void)mv; (void)n;
    } else {
        k_leaf_moves<V><<<c.leaf_grid(n), LEAF_WG, 0, c.stream>>>(gp, par, mv, n, c.dAcc, c.dItems, c.dErr, c.dOvf,
                                                                  c.dt);
    }
}

// the last three plies in one kernel: the n children of gp (par/mv) are made
// and counted in the leaf's work-groups
int Engine::leaf_moves(int v, const Pos *gp, const u32 *par, const u16 *mv, u64 n) {
    if (timing) HIP_TRY(hipEventRecord(ev0, stream));
    PG_VARIANT_CALL(v, launch_leaf_moves, *this, gp, par, mv, n);
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

int Engine::leaf_fused(int v, const Pos *b, const u32 *cnt, u64 n) {
    if (timing) HIP_TRY(hipEventRecord(ev0, stream));
    PG_VARIANT_CALL(v, launch_leaf_fused, *this, b, cnt, n);
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

// Sum of perft(r) over n boards that sit at `ply` below the root.  cnt, when
// given, already holds each board's move count.
int Engine::bfs(int v, const Pos *boards, u64 n, const u32 *cnt, int r, int ply) {
    if (n == 0) return PG_OK;
    if (r == 1) {
        if (cnt) k_sum_u32<<<grid(n, 256), 256, 0, stream>>>(cnt, n, dAcc, dOvf);
        else PG_VARIANT_CALL(v, launch_count_sum, *this, boards, n, dAcc);
        HIP_TRY(hipGetLastError());
        return PG_OK;
    }
    if (n > 0xFFFFFFFFULL) return split(v, boards, n, cnt, r, ply);
    const size_t mark = top;
    u32 *c = const_cast<u32 *>(cnt);
    if (!c) {
        c = (u32 *)alloc(n * sizeof(u32));
        if (!c) { top = mark; return split(v, boards, n, cnt, r, ply); }
        PG_VARIANT_CALL(v, launch_count, *this, boards, n, c);
    }
    if (r == 2 && leafMode == 0 && v != SL_BYTE) {
        // last two plies in one kernel; the items never reach HBM
        const int rc = leaf_fused(v, boards, c, n);
        top = mark;
        return rc;
    }
    u64 *off = (u64 *)alloc((n + 1) * sizeof(u64));
    u64 total = 0;
    int rc = off ? scan(c, n, off, &total) : PG_ENOMEM;
    if (rc == PG_ENOMEM) { top = mark; return split(v, boards, n, cnt, r, ply); }
    if (rc) return rc;
    if (total == 0) { top = mark; return PG_OK; }

    // three plies left with the fused leaf: the children are made inside the
    // leaf from the move list, so only moves and parents are stored
    const bool leafMoves = r == 3 && leafMode == 0 && v != SL_BYTE;
    Pos *children = nullptr;
    u32 *ccnt = nullptr;
    if (r > 2 && !leafMoves) {
        children = (Pos *)alloc(total * sizeof(Pos));
        ccnt = children ? (u32 *)alloc(total * sizeof(u32)) : nullptr;
    }
    u16 *mv = (r == 2 || ccnt || leafMoves) ? (u16 *)alloc(total * sizeof(u16)) : nullptr;
    u32 *par = mv ? (u32 *)alloc(total * sizeof(u32)) : nullptr;
    if (!par) {
        top = mark;
        if (n == 1) { set_error("one board's children do not fit the arena (ply %d)", ply); return PG_ENOMEM; }
        return split(v, boards, n, cnt, r, ply);
    }
    if (ply + 1 < 16) stats.nodes[ply + 1] += total;
    PG_VARIANT_CALL(v, launch_generate, *this, boards, n, off, mv, par);
    HIP_TRY(hipGetLastError());
    if (r == 2) {
        rc = leaf(v, boards, n, par, mv, total);
        top = mark;
        return rc;
    }
    if (leafMoves) {
        rc = leaf_moves(v, boards, par, mv, total);
        top = mark;
        return rc;
    }
    PG_VARIANT_CALL(v, launch_expand, *this, boards, par, mv, total, children, ccnt);
    HIP_TRY(hipGetLastError());
    top = (size_t)((u8 *)ccnt - arena) + total * sizeof(u32);  // drop moves and parents
    rc = bfs(v, children, total, ccnt, r - 1, ply + 1);
    top = mark;
    return rc;
}

// PG_EOVERFLOW if a 64-bit total wrapped, PG_EINVAL if a device
// consistency check failed (the stream has been synchronized by the caller's
// read of the result).
int Engine::check_flags(const char *what) {
    u32 *f = reinterpret_cast<u32 *>(hostBuf + 1);  // pinned staging: err, ovf
    HIP_TRY(hipMemcpyAsync(f, dErr, 2 * sizeof(u32), hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    if (f[1]) {
        set_error("%s: a perft total does not fit 64 bits", what);
        return PG_EOVERFLOW;
    }
    if (f[0]) {
        set_error("%s: device consistency check failed (flag %u)", what, f[0]);
        return PG_EINVAL;
    }
    return PG_OK;
}

int Engine::split(int v, const Pos *boards, u64 n, const u32 *cnt, int r, int ply) {
    if (n <= 1) {
        set_error("arena too small for a single board at ply %d", ply);
        return PG_ENOMEM;
    }
    stats.splits++;
    const u64 h = n / 2;
    int rc = bfs(v, boards, h, cnt, r, ply);
    if (rc) return rc;
    return bfs(v, boards + h, n - h, cnt ? cnt + h : nullptr, r, ply);
}

int Engine::perft(const Pos &root, u32 depth, int v, u64 *out) {
    memset(&stats, 0, sizeof(stats));
    const double t0 = now_ms();
    top = 0;
    peak = 0;
    if (depth == 0) { *out = 1; return PG_OK; }
    // Whole search as one graph launch (no host round trip between levels)
    // unless each leaf launch is to be timed with events, the materialized
    // leaf is asked for, or the call must not allocate (its plan and graphs
    // are created on first use).  A search whose levels do not fit the arena
    // falls back to the host-driven loop below, which splits levels.
    if (!timing && leafMode == 0 && graphPerft) {
        bool fits = false;
        const int rc = perft_graph(*this, root, depth, v, out, &fits);
        if (rc) return rc;
        if (fits) {
            stats.leaves = *out;
            stats.total_ms = now_ms() - t0;
            return PG_OK;
        }
        memset(&stats, 0, sizeof(stats));
    }
    Pos *d = (Pos *)alloc(sizeof(Pos));
    if (!d) return PG_ENOMEM;
    Pos *staged = reinterpret_cast<Pos *>(hostBuf + 8);  // pinned, so the copy needs no bounce buffer
    *staged = root;
    HIP_TRY(hipMemcpyAsync(d, staged, sizeof(Pos), hipMemcpyHostToDevice, stream));
    k_init_call<<<1, 64, 0, stream>>>(dAcc, bias, dItems, dErr, dOvf);
    int rc = bfs(v, d, 1, nullptr, (int)depth, 0);
    if (rc) return rc;
    u64 res = 0;
    rc = read_u64(dAcc, &res);
    if (rc) return rc;
    rc = check_flags("perft");
    if (rc) return rc;
    if (leafMode == 0) {
        u64 items = 0;
        rc = read_u64(dItems, &items);
        if (rc) return rc;
        stats.leaf_items += items;
        if (depth >= 2 && depth - 1 < 16) stats.nodes[depth - 1] += items;  // fused launches all sit at ply depth-2
    }
    stats.leaves = res;
    stats.peak_arena = peak;
    stats.total_ms = now_ms() - t0;
    *out = res;
    return PG_OK;
}

}  // namespace pg

// ============================================================ static scratch
// Device memory for pg_perft_noalloc, a variant of pg_perft for callers that
// cannot allocate (code under HIP graph capture, tight serving loops): the
// tree arena and the per-call totals and flags of that entry point live in
// this file-scope array instead of the context's arena.  Every GPU the library
// runs on holds its own copy until the process ends.  In host code `scratch`
// names a host-side stand-in, never the device memory: its device address is
// taken once per device with hipGetSymbolAddress (scratch_setup, called from
// pg_init with the device current) and only that address is used.
namespace pg {
constexpr size_t NOALLOC_SCRATCH_BYTES = 64ull << 20;  // 256 B of totals/flags, the rest tree arena
constexpr size_t NOALLOC_HEAD = 256;
static __device__ __align__(256) u8 scratch[NOALLOC_SCRATCH_BYTES];

static std::mutex g_scratch_mu;
static u8 *g_scratch_ptr[64];  // per device index

int scratch_setup(int device, u8 **out) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (device < 0 || device >= 64) return PG_EINVAL;
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error("static scratch of device %d: hipGetSymbolAddress failed (%s)", device, hipGetErrorString(e));
            return PG_EHIP;
        }
        g_scratch_ptr[device] = static_cast<u8 *>(p);
    }
    *out = g_scratch_ptr[device];
    return PG_OK;
}
}  // namespace pg

// ====================================================================== C ABI
struct pg_ctx : public pg::Engine {};

static int resolve_slider(const pg_ctx *c, int s) {
    if (s == PG_SLIDER_AUTO) return c->slider;
    if (s < 0 || s >= SL_COUNT) return -1;
    return s;
}

// ------------------------------------------------------------------ probes
template <int V> static void launch_probe_gen(Engine &c, const Pos *b, u64 n, u16 *mv, u32 *cnt) {
    constexpr int W = wg_size<V>();
    k_probe_generate<V><<<c.grid(n, W), W, 0, c.stream>>>(b, n, mv, cnt, c.dt);
}
template <int V> static void launch_probe_slider(Engine &c, int rook, const u8 *sq, const u64 *occ, u64 n, u64 *out) {
    constexpr int W = wg_size<V>();
    k_probe_slider<V><<<c.grid(n, W), W, 0, c.stream>>>(rook, sq, occ, n, out, c.dt);
}

extern "C" {

int pg_init(int device, const pg_config *cfg, pg_ctx **out) {
    if (!out) return PG_EINVAL;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        set_error("no GPU visible (%d devices)", ndev);
        return PG_ENODEV;
    }
    if (device < 0 || device >= ndev) return PG_EINVAL;
    const char *terr = nullptr;
    const HostTables *ht = host_tables(&terr);
    if (!ht) {
        set_error("table build failed: %s", terr ? terr : "?");
        return PG_ETABLE;
    }
    pg_ctx *c = new pg_ctx();
    c->device = device;
    pg_config def = {0, PG_SLIDER_AUTO, 0, 0, 0, 0};
    c->cfg = cfg ? *cfg : def;
    c->slider = (c->cfg.slider >= 0 && c->cfg.slider < SL_COUNT) ? c->cfg.slider : SL_KOGGE;
    c->bpc = c->cfg.blocks_per_cu > 0 ? c->cfg.blocks_per_cu : 8;
    c->timing = c->cfg.timing != 0;
    auto fail = [&](int rc) { pg_shutdown(c); return rc; };
    if (hipSetDevice(device) != hipSuccess) return fail(PG_EHIP);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return fail(PG_EHIP);
    c->cus = prop.multiProcessorCount;
    c->ldsPerBlock = prop.sharedMemPerBlock;
    c->clockKHz = prop.clockRate;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return fail(PG_EHIP);
    if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) return fail(PG_EHIP);
    std::vector<u8> blob;
    pack_tables(*ht, blob);
    if (hipMalloc(&c->dTables, blob.size()) != hipSuccess) return fail(PG_ENOMEM);
    if (hipMemcpy(c->dTables, blob.data(), blob.size(), hipMemcpyHostToDevice) != hipSuccess) return fail(PG_EHIP);
    c->dt = tables_at(c->dTables);
    if (hipMalloc(&c->dSmall, 256) != hipSuccess) return fail(PG_ENOMEM);
    c->dAcc = (u64 *)c->dSmall;
    c->dItems = (u64 *)(c->dSmall + 8);
    c->leafMode = c->cfg.leaf_mode;
    c->dErr = (u32 *)(c->dSmall + 64);
    c->dOvf = (u32 *)(c->dSmall + 68);
    if (hipHostMalloc(&c->hostBuf, 256) != hipSuccess) return fail(PG_ENOMEM);
    {
        const int rc = scratch_setup(device, &c->scratch);
        if (rc) return fail(rc);
    }
    size_t freeB = 0, totalB = 0;
    (void)hipMemGetInfo(&freeB, &totalB);
    size_t want = c->cfg.arena_bytes ? c->cfg.arena_bytes : (48ull << 30);
    if (!c->cfg.arena_bytes && want > freeB / 2) want = freeB / 2;
    if (want < (4ull << 20)) want = 4ull << 20;
    if (hipMalloc(&c->arena, want) != hipSuccess) {
        set_error("arena reservation of %zu MiB failed", want >> 20);
        return fail(PG_ENOMEM);
    }
    c->arenaBytes = want;
    c->cfg.arena_bytes = want;
    c->cfg.slider = c->slider;
    *out = c;
    return PG_OK;
}

void pg_shutdown(pg_ctx *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    // slot servers and tables still alive belong to this context
    while (!c->servers.empty()) server_free(c->servers.back());
    while (!c->tts.empty()) tt_free(c->tts.back());
    slot_delete(c->batchSlot);
    c->batchSlot = nullptr;
    if (c->arena) (void)hipFree(c->arena);
    if (c->dTables) (void)hipFree(c->dTables);
    if (c->dSmall) (void)hipFree(c->dSmall);
    if (c->hostBuf) (void)hipHostFree(c->hostBuf);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int pg_get_config(pg_ctx *c, pg_config *out) {
    if (!c || !out) return PG_EINVAL;
    *out = c->cfg;
    return PG_OK;
}

int pg_set_slider(pg_ctx *c, int s) {
    if (!c || s < 0 || s >= SL_COUNT) return PG_EINVAL;
    c->slider = s;
    c->cfg.slider = s;
    return PG_OK;
}

int pg_perft(pg_ctx *c, const pg_pos48 *pos, uint32_t depth, int slider, uint64_t *out) {
    if (!c || !pos || !out || depth > 15) return PG_EINVAL;
    const int v = resolve_slider(c, slider);
    if (v < 0) return PG_EINVAL;
    if (hipSetDevice(c->device) != hipSuccess) return PG_EHIP;
    Pos p;
    memcpy(&p, pos, sizeof(p));
    if (!pos_sane(p)) {
        set_error("position needs exactly one king per side");
        return PG_EINVAL;
    }
    return c->perft(p, depth, v, reinterpret_cast<u64 *>(out));
}

int pg_perft_noalloc(pg_ctx *c, const pg_pos48 *pos, uint32_t depth, int slider, uint64_t *out) {
    if (!c || !pos || !out) return PG_EINVAL;
    if (depth > 15) {
        set_error("pg_perft_noalloc: depth %u exceeds the maximum of 15", depth);
        return PG_EINVAL;
    }
    const int v = resolve_slider(c, slider);
    if (v < 0) return PG_EINVAL;
    if (!c->scratch) {
        set_error("pg_perft_noalloc: no static scratch on device %d", c->device);
        return PG_ENODEV;
    }
    if (hipSetDevice(c->device) != hipSuccess) return PG_EHIP;
    Pos p;
    memcpy(&p, pos, sizeof(p));
    if (!pos_sane(p)) {
        set_error("position needs exactly one king per side");
        return PG_EINVAL;
    }
    // run the plain search with its arena and per-call words in the static
    // scratch, then give the context its own buffers back
    u8 *const arena0 = c->arena;
    const size_t bytes0 = c->arenaBytes, top0 = c->top, peak0 = c->peak;
    u64 *const acc0 = c->dAcc, *const items0 = c->dItems;
    u32 *const err0 = c->dErr, *const ovf0 = c->dOvf;
    const int leafMode0 = c->leafMode;
    c->graphPerft = false;  // the graph path's plan and graphs are allocated on first use
    c->dAcc = reinterpret_cast<u64 *>(c->scratch);
    c->dItems = reinterpret_cast<u64 *>(c->scratch + 16);
    c->dErr = reinterpret_cast<u32 *>(c->scratch + 64);
    c->dOvf = reinterpret_cast<u32 *>(c->scratch + 68);
    c->arena = c->scratch + NOALLOC_HEAD;
    c->arenaBytes = NOALLOC_SCRATCH_BYTES - NOALLOC_HEAD;
    c->leafMode = 0;
    const int rc = c->perft(p, depth, v, reinterpret_cast<u64 *>(out));
    c->arena = arena0;
    c->arenaBytes = bytes0;
    c->top = top0;
    c->peak = peak0;
    c->dAcc = acc0;
    c->dItems = items0;
    c->dErr = err0;
    c->dOvf = ovf0;
    c->leafMode = leafMode0;
    c->graphPerft = true;
    if (rc == PG_ENOMEM)
        set_error("pg_perft_noalloc: one board's children do not fit the %zu MiB static scratch",
                  NOALLOC_SCRATCH_BYTES >> 20);
    return rc;
}

int pg_test_count_bias(pg_ctx *c, uint64_t bias) {
    if (!c) return PG_EINVAL;
    c->bias = bias;
    return PG_OK;
}

int pg_last_stats(pg_ctx *c, pg_stats *out) {
    if (!c || !out) return PG_EINVAL;
    *out = c->stats;
    return PG_OK;
}

const char *pg_last_error(void) { return g_last_error; }

static bool boards_sane(const pg_pos48 *pos, uint32_t n) {
    for (uint32_t i = 0; i < n; i++) {
        Pos p;
        memcpy(&p, &pos[i], sizeof(p));
        if (!pos_sane(p)) {
            set_error("board %u needs exactly one king per side", i);
            return false;
        }
    }
    return true;
}

static int stage_in(pg_ctx *c, const void *h, size_t bytes, void **d) {
    *d = c->alloc(bytes);
    if (!*d) return PG_ENOMEM;
    if (h && hipMemcpyAsync(*d, h, bytes, hipMemcpyHostToDevice, c->stream) != hipSuccess) return PG_EHIP;
    return PG_OK;
}
static int stage_out(pg_ctx *c, void *h, const void *d, size_t bytes) {
    if (hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, c->stream) != hipSuccess) return PG_EHIP;
    if (hipStreamSynchronize(c->stream) != hipSuccess) return PG_EHIP;
    return PG_OK;
}

#define PROBE_TRY(x) do { int r_ = (x); if (r_) { c->top = 0; return r_; } } while (0)

int pg_device_count_moves(pg_ctx *c, const pg_pos48 *pos, uint32_t n, int slider, uint32_t *out) {
    if (!c || (!pos && n) || (!out && n)) return PG_EINVAL;
    const int v = resolve_slider(c, slider);
    if (v < 0) return PG_EINVAL;
    if (!boards_sane(pos, n)) return PG_EINVAL;
    if (!n) return PG_OK;
    (void)hipSetDevice(c->device);
    c->top = 0;
    void *db, *dc;
    PROBE_TRY(stage_in(c, pos, n * sizeof(Pos), &db));
    PROBE_TRY(stage_in(c, nullptr, n * sizeof(u32), &dc));
    PG_VARIANT_CALL(v, launch_count, *c, (const Pos *)db, n, (u32 *)dc);
    if (hipGetLastError() != hipSuccess) return PG_EHIP;
    PROBE_TRY(stage_out(c, out, dc, n * sizeof(u32)));
    c->top = 0;
    return PG_OK;
}

int pg_device_generate_moves(pg_ctx *c, const pg_pos48 *pos, uint32_t n, int slider, pg_move *moves, uint32_t *counts) {
    if (!c || (n && (!pos || !moves || !counts))) return PG_EINVAL;
    const int v = resolve_slider(c, slider);
    if (v < 0) return PG_EINVAL;
    if (!boards_sane(pos, n)) return PG_EINVAL;
    if (!n) return PG_OK;
    (void)hipSetDevice(c->device);
    c->top = 0;
    void *db, *dm, *dc;
    PROBE_TRY(stage_in(c, pos, n * sizeof(Pos), &db));
    PROBE_TRY(stage_in(c, nullptr, (size_t)n * 256 * sizeof(u16), &dm));
    PROBE_TRY(stage_in(c, nullptr, n * sizeof(u32), &dc));
    PG_VARIANT_CALL(v, launch_probe_gen, *c, (const Pos *)db, n, (u16 *)dm, (u32 *)dc);
    if (hipGetLastError() != hipSuccess) return PG_EHIP;
    PROBE_TRY(stage_out(c, moves, dm, (size_t)n * 256 * sizeof(u16)));
    PROBE_TRY(stage_out(c, counts, dc, n * sizeof(u32)));
    c->top = 0;
    return PG_OK;
}

int pg_device_make_move(pg_ctx *c, const pg_pos48 *pos, const pg_move *moves, uint32_t n, pg_pos48 *out, pg_key128 *keys) {
    if (!c || (n && (!pos || !moves || !out))) return PG_EINVAL;
    if (!boards_sane(pos, n)) return PG_EINVAL;
    if (!n) return PG_OK;
    (void)hipSetDevice(c->device);
    c->top = 0;
    void *db, *dm, *dout, *dk = nullptr;
    PROBE_TRY(stage_in(c, pos, n * sizeof(Pos), &db));
    PROBE_TRY(stage_in(c, moves, n * sizeof(u16), &dm));
    PROBE_TRY(stage_in(c, nullptr, n * sizeof(Pos), &dout));
    if (keys) PROBE_TRY(stage_in(c, keys, n * sizeof(Key128), &dk));
    k_probe_make<<<c->grid(n, 256), 256, 0, c->stream>>>((const Pos *)db, (const u16 *)dm, n, (Pos *)dout, (Key128 *)dk, c->dt);
    if (hipGetLastError() != hipSuccess) return PG_EHIP;
    PROBE_TRY(stage_out(c, out, dout, n * sizeof(Pos)));
    if (keys) PROBE_TRY(stage_out(c, keys, dk, n * sizeof(Key128)));
    c->top = 0;
    return PG_OK;
}

int pg_device_scan(pg_ctx *c, const uint32_t *counts, uint64_t n, uint32_t misalign, uint64_t *offsets) {
    if (!c || !counts || !offsets || n == 0 || n > 0xFFFFFFFFULL || misalign > 3) return PG_EINVAL;
    (void)hipSetDevice(c->device);
    c->top = 0;
    void *dc, *doff;
    PROBE_TRY(stage_in(c, nullptr, (n + misalign) * sizeof(u32), &dc));
    u32 *cnt = (u32 *)dc + misalign;  // the arena hands out 256-byte aligned blocks
    if (hipMemcpyAsync(cnt, counts, n * sizeof(u32), hipMemcpyHostToDevice, c->stream) != hipSuccess) return PG_EHIP;
    PROBE_TRY(stage_in(c, nullptr, (n + 1) * sizeof(u64), &doff));
    u64 total = 0;
    PROBE_TRY(c->scan(cnt, n, (u64 *)doff, &total));
    PROBE_TRY(stage_out(c, offsets, doff, (n + 1) * sizeof(u64)));
    c->top = 0;
    return PG_OK;
}

int pg_device_slider_attacks(pg_ctx *c, int slider, int rook, const uint8_t *sq, const uint64_t *occ, uint32_t n, uint64_t *out) {
    if (!c || (n && (!sq || !occ || !out))) return PG_EINVAL;
    const int v = resolve_slider(c, slider);
    if (v < 0) return PG_EINVAL;
    for (uint32_t i = 0; i < n; i++)
        if (sq[i] > 63) return PG_EINVAL;
    if (!n) return PG_OK;
    (void)hipSetDevice(c->device);
    c->top = 0;
    void *ds, *dO, *dout;
    PROBE_TRY(stage_in(c, sq, n, &ds));
    PROBE_TRY(stage_in(c, occ, n * sizeof(u64), &dO));
    PROBE_TRY(stage_in(c, nullptr, n * sizeof(u64), &dout));
    PG_VARIANT_CALL(v, launch_probe_slider, *c, rook, (const u8 *)ds, (const u64 *)dO, n, (u64 *)dout);
    if (hipGetLastError() != hipSuccess) return PG_EHIP;
    PROBE_TRY(stage_out(c, out, dout, n * sizeof(u64)));
    c->top = 0;
    return PG_OK;
}

}  // extern "C"
