// graph.hip -- batched perft: many roots (mixed side to move allowed), one
// breadth-first search over all of them, captured once per shape into a HIP
// graph and replayed per batch.
//
// Prior art in the reference: sibling batching of up to 256 children with a
// quarter-size retry on OOM (launcher.h:966-1179) and the serial root loop of
// perft_bb_gpu_simple_hash, which needs every root to share the side to move
// (perft_bb.h:1296).  Here roots of both colours share the levels; they are
// ordered by side so that waves stay uniform.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>
#include "engine.h"
#include "dloop.h"

using namespace pg;

namespace pg {

struct GraphShape {
    int depth, v, keys, dedup, bpc;
    const TTable *tt;
    int total;  // 1: one root, only the sum is wanted (no per-board perft arrays)
    int cap;    // > 0: at most this many work-groups per CU in every kernel (leaves CUs to other streams)
    int gmax;   // > 0: at most this many work-groups in any kernel (a small search's whole need)
    bool operator==(const GraphShape &o) const {
        return depth == o.depth && v == o.v && keys == o.keys && dedup == o.dedup && bpc == o.bpc && tt == o.tt &&
               total == o.total && cap == o.cap && gmax == o.gmax;
    }
};

// The most work-groups any kernel of a search of n roots to `depth` can use:
// its largest materialized level holds at most about n * 48^(depth-2)
// boards, 256 per work-group.  Rounded up to a power of two (at least 64) so
// that few distinct graphs are captured; 0 when the whole chip is wanted.
// Every kernel is a grid-stride loop, so any grid size is correct; a small
// grid only spares the dispatcher the empty work-groups of a small search.
static int grid_limit(const Engine &e, u64 n, int depth) {
    const int full = e.cus * LEAF_BLOCKS_PER_CU;
    double wgs = (double)n / 256.0;
    for (int l = 2; l < depth && wgs < full; l++) wgs *= 48.0;
    if (wgs >= full) return 0;
    int g = 64;
    while (g < wgs) g <<= 1;
    return g >= full ? 0 : g;
}

// Grid sizes of one captured search
struct Grids {
    int work, scan, byteV, leaf, v;  // grid-stride kernels, scans, byte-lookup kernels, fused leaf, per-variant
};
static Grids grids_for(const Engine &e, const GraphShape &sh) {
    // a capped search keeps at most sh.cap work-groups per CU of each kernel
    // resident (four fit per CU), so batches on other streams find room
    const int wpc = sh.cap > 0 && sh.cap < sh.bpc ? sh.cap : sh.bpc;
    Grids g;
    g.work = e.cus * wpc;
    g.scan = e.cus * (sh.cap > 0 && sh.cap < 4 ? sh.cap : 4);
    g.byteV = sh.cap > 0 && sh.cap < 4 ? (e.cus * sh.cap + 3) / 4 : e.cus;
    g.leaf = e.cus * (sh.cap > 0 ? sh.cap : LEAF_BLOCKS_PER_CU);
    if (sh.gmax > 0) {
        g.work = std::min(g.work, sh.gmax);
        g.scan = std::min(g.scan, sh.gmax);
        g.byteV = std::min(g.byteV, sh.gmax);
        g.leaf = std::min(g.leaf, sh.gmax);
    }
    g.v = sh.v == SL_BYTE ? g.byteV : g.work;
    return g;
}

struct BatchGraph {
    GraphShape shape;
    hipGraphExec_t exec = nullptr;
};

// One batch slot: a plan in device memory and the arena region it carves.
struct Slot {
    Plan *dPlan = nullptr;
    Plan *hPlan = nullptr;  // pinned staging copy
    u8 *region = nullptr;
    size_t regionBytes = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<BatchGraph> graphs;
};

template <int V> static void cap_count(hipStream_t s, Plan *P, int g, const Tables &t) {
    kd_count<V><<<g, wg_size<V>(), 0, s>>>(P, t);
}
template <int V> static void cap_generate(hipStream_t s, Plan *P, int l, int g, const Tables &t) {
    kd_generate<V><<<g, wg_size<V>(), 0, s>>>(P, l, t);
}
template <int V> static void cap_expand(hipStream_t s, Plan *P, int l, const TTView &tt, int rc, TTCounters *ctr, int g,
                                        const Tables &t) {
    kd_expand<V><<<g, wg_size<V>(), 0, s>>>(P, l, tt, rc, ctr, t);
}
template <int V> static void cap_leaf(hipStream_t s, Plan *P, int l, int g, const Tables &t) {
    if constexpr (V != SL_BYTE) kd_leaf<V, false><<<g, LEAF_WG, 0, s>>>(P, l, t);
}
template <int V> static void cap_leaf_total(hipStream_t s, Plan *P, int l, int g, const Tables &t) {
    if constexpr (V != SL_BYTE) kd_leaf<V, true><<<g, LEAF_WG, 0, s>>>(P, l, t);
}
template <int V> static void cap_expand_keys(hipStream_t s, Plan *P, int l, const TTView &tt, int rc, TTCounters *ctr,
                                             int g, const Tables &t) {
    if constexpr (V != SL_BYTE) kd_expand_keys<V><<<g, wg_size<V>(), 0, s>>>(P, l, tt, rc, ctr, t);
}
template <int V> static void cap_leaf_keys(hipStream_t s, Plan *P, int l, int g, const Tables &t) {
    if constexpr (V != SL_BYTE) kd_leaf_keys<V><<<g, LEAF_WG, 0, s>>>(P, l, t);
}
template <int V> static void cap_leaf_moves(hipStream_t s, Plan *P, int l, int g, const Tables &t) {
    if constexpr (V != SL_BYTE) kd_leaf_moves<V><<<g, LEAF_WG, 0, s>>>(P, l, t);
}

// The search of one root when only its perft is wanted: per level just
// boards and counts (no keys, table or per-board sums), the leaf adds into the
// plan's total.
static void enqueue_total(const Engine &e, hipStream_t s, Plan *P, const GraphShape &sh) {
    const int D = sh.depth, v = sh.v;
    const Grids G = grids_for(e, sh);
    const int gWork = G.work, gScan = G.scan, gv = G.v;
    PG_VARIANT_CALL(v, cap_count, s, P, gv, e.dt);
    const int lastExpand = v == SL_BYTE ? D - 2 : D - 3;
    for (int l = 0; l <= lastExpand; l++) {
        // with the fused leaf the last expansion lists moves only: the leaf
        // makes and counts those children itself (kd_leaf_moves)
        const bool movesOnly = v != SL_BYTE && l == lastExpand;
        kd_plan_scan<<<1, 64, 0, s>>>(P, l);
        kd_scan_reduce<<<gScan, SCAN_WG, 0, s>>>(P, l);
        kd_scan_spine<<<1, 1024, 0, s>>>(P, l, 0, 0, movesOnly ? 2 : 1);
        kd_scan_down<<<gScan, SCAN_WG, 0, s>>>(P, l);
        PG_VARIANT_CALL(v, cap_generate, s, P, l, gv, e.dt);
        if (movesOnly) {
            PG_VARIANT_CALL(v, cap_leaf_moves, s, P, l, G.leaf, e.dt);
            return;
        }
        PG_VARIANT_CALL(v, cap_expand, s, P, l, TTView{}, D - l - 1, (TTCounters *)nullptr, gv, e.dt);
    }
    const int last = lastExpand + 1 < 0 ? 0 : lastExpand + 1;  // deepest materialized level
    if (D == 1 || v == SL_BYTE) kd_sum_counts<<<gWork, 256, 0, s>>>(P, last);
    else PG_VARIANT_CALL(v, cap_leaf_total, s, P, last, G.leaf, e.dt);
}

// Enqueues the whole search for `depth` on stream s (captured by the caller).
static void enqueue_search(const Engine &e, hipStream_t s, Plan *P, const GraphShape &sh) {
    const int D = sh.depth, v = sh.v;
    const Grids G = grids_for(e, sh);
    const int gWork = G.work, gScan = G.scan, gLeaf = G.leaf, gv = G.v;
    TTView tt{};
    TTCounters *ctr = nullptr;
    if (sh.tt) {
        tt = sh.tt->view;
        ctr = sh.tt->dCtr;
    }
    PG_VARIANT_CALL(v, cap_count, s, P, gv, e.dt);
    kd_root_probe<<<gWork, 256, 0, s>>>(P, tt, D, ctr);
    // levels expanded into children: remaining depth r = D - l >= 3, or >= 2
    // for the byte-lookup variant (no fused leaf, LDS is taken by its tables)
    const int lastExpand = v == SL_BYTE ? D - 2 : D - 3;
    // with keys, the deepest level keeps keys only (plan_children, total 3):
    // its children are made for their keys and collapsed, and the leaf makes
    // those that remain from their parent and move
    const bool keysLast = sh.keys && v != SL_BYTE && lastExpand >= 0;
    for (int l = 0; l <= lastExpand; l++) {
        const int rc = D - l - 1;
        const int wantSet = sh.dedup && rc >= 2;
        const bool kl = keysLast && l == lastExpand;
        kd_plan_scan<<<1, 64, 0, s>>>(P, l);
        kd_scan_reduce<<<gScan, SCAN_WG, 0, s>>>(P, l);
        kd_scan_spine<<<1, 1024, 0, s>>>(P, l, sh.keys, wantSet, kl ? 3 : 0);
        kd_scan_down<<<gScan, SCAN_WG, 0, s>>>(P, l);
        PG_VARIANT_CALL(v, cap_generate, s, P, l, gv, e.dt);
        if (kl) {
            PG_VARIANT_CALL(v, cap_expand_keys, s, P, l, tt, rc, ctr, gv, e.dt);
        } else {
            PG_VARIANT_CALL(v, cap_expand, s, P, l, tt, rc, ctr, gv, e.dt);
        }
        if (wantSet) {
            kd_clear_set<<<gWork, 256, 0, s>>>(P, l);
            kd_dedup<<<gWork, 256, 0, s>>>(P, l, ctr, rc, sh.tt != nullptr, e.dt);
        }
    }
    const int last = lastExpand + 1;  // deepest materialized level
    if (D == 1 || v == SL_BYTE) {
        kd_perft1<<<gWork, 256, 0, s>>>(P, last < 0 ? 0 : last);
    } else {
        // the leaf over the live boards (table hits and collapsed duplicates
        // are resolved already)
        kd_plan_live<<<1, 64, 0, s>>>(P, last);
        kd_live<<<gWork, LIVE_WG, 0, s>>>(P, last);
        if (keysLast) {
            PG_VARIANT_CALL(v, cap_leaf_keys, s, P, last, gLeaf, e.dt);
        } else {
            PG_VARIANT_CALL(v, cap_leaf, s, P, last, gLeaf, e.dt);
        }
    }
    for (int l = lastExpand; l >= 0; l--) {
        kd_resolve<<<gWork, 256, 0, s>>>(P, l, tt, D - l - 1, ctr);
        kd_segsum<<<gWork, 256, 0, s>>>(P, l);
    }
    kd_root_store<<<gWork, 256, 0, s>>>(P, tt, D, ctr);
}

static int get_graph(Engine &e, Slot &sl, const GraphShape &sh, hipGraphExec_t *out) {
    for (auto &g : sl.graphs)
        if (g.shape == sh) {
            *out = g.exec;
            return PG_OK;
        }
    hipGraph_t graph;
    HIP_TRY(hipStreamBeginCapture(sl.stream, hipStreamCaptureModeThreadLocal));
    if (sh.total) enqueue_total(e, sl.stream, sl.dPlan, sh);
    else enqueue_search(e, sl.stream, sl.dPlan, sh);
    HIP_TRY(hipStreamEndCapture(sl.stream, &graph));
    hipGraphExec_t exec;
    HIP_TRY(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    (void)hipGraphDestroy(graph);
    sl.graphs.push_back({sh, exec});
    *out = exec;
    return PG_OK;
}

// Runs one batch (sub-range of the level-0 arrays already on the device);
// sets *overflow if it did not fit.
static int run_range(Engine &e, Slot &sl, const GraphShape &sh, u64 first, u64 n, const DLevel &L0, size_t top0,
                     bool *overflow, u64 *leafItems, double *ms, u64 *peak) {
    Plan &h = *sl.hPlan;
    memset(&h, 0, sizeof(Plan));
    h.base = sl.region;
    h.cap = sl.regionBytes;
    h.top = top0;
    h.bias = e.bias;
    h.lv[0] = L0;
    h.lv[0].n = n;
    h.lv[0].b += first;
    if (h.lv[0].k) h.lv[0].k += first;
    h.lv[0].cnt += first;
    h.lv[0].perft += first;
    h.lv[0].state += first;
    h.lv[0].dup += first;
    hipGraphExec_t exec;
    int rc = get_graph(e, sl, sh, &exec);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(sl.dPlan, &h, sizeof(Plan), hipMemcpyHostToDevice, sl.stream));
    HIP_TRY(hipEventRecord(sl.ev0, sl.stream));
    HIP_TRY(hipGraphLaunch(exec, sl.stream));
    HIP_TRY(hipEventRecord(sl.ev1, sl.stream));
    HIP_TRY(hipMemcpyAsync(&h, sl.dPlan, sizeof(Plan), hipMemcpyDeviceToHost, sl.stream));
    HIP_TRY(hipStreamSynchronize(sl.stream));
    float t = 0;
    (void)hipEventElapsedTime(&t, sl.ev0, sl.ev1);
    *ms += t;
    *overflow = h.overflow != 0;
    *leafItems += h.leafItems;
    if (h.peak > *peak) *peak = h.peak;
    if (h.ovf) {
        set_error("batch: a perft total does not fit 64 bits");
        return PG_EOVERFLOW;
    }
    if (h.err) {
        set_error("device consistency check failed in batch (flag %u)", h.err);
        return PG_EINVAL;
    }
    return PG_OK;
}

static int run_split(Engine &e, Slot &sl, const GraphShape &sh, u64 first, u64 n, const DLevel &L0, size_t top0,
                     u64 *leafItems, double *ms, u64 *launches, u64 *splits, u64 *peak) {
    bool over = false;
    (*launches)++;
    int rc = run_range(e, sl, sh, first, n, L0, top0, &over, leafItems, ms, peak);
    if (rc) return rc;
    if (!over) return PG_OK;
    if (n == 1) {
        set_error("one root does not fit the batch region (%zu bytes)", sl.regionBytes);
        return PG_ENOMEM;
    }
    (*splits)++;
    const u64 h = n / 2;
    rc = run_split(e, sl, sh, first, h, L0, top0, leafItems, ms, launches, splits, peak);
    if (rc) return rc;
    return run_split(e, sl, sh, first + h, n - h, L0, top0, leafItems, ms, launches, splits, peak);
}

int slot_init(Engine &e, Slot &sl, u8 *region, size_t bytes, hipStream_t stream) {
    sl.region = region;
    sl.regionBytes = bytes;
    sl.stream = stream;
    HIP_TRY(hipMalloc(&sl.dPlan, sizeof(Plan)));
    HIP_TRY(hipHostMalloc(&sl.hPlan, sizeof(Plan)));
    HIP_TRY(hipEventCreate(&sl.ev0));
    HIP_TRY(hipEventCreate(&sl.ev1));
    (void)e;
    return PG_OK;
}

void slot_free(Slot &sl) {
    for (auto &g : sl.graphs) (void)hipGraphExecDestroy(g.exec);
    sl.graphs.clear();
    if (sl.dPlan) (void)hipFree(sl.dPlan);
    if (sl.hPlan) (void)hipHostFree(sl.hPlan);
    if (sl.ev0) (void)hipEventDestroy(sl.ev0);
    if (sl.ev1) (void)hipEventDestroy(sl.ev1);
    sl = Slot();
}

void slot_delete(Slot *sl) {
    if (!sl) return;
    slot_free(*sl);
    delete sl;
}

void slot_forget_tt(Slot &sl, const TTable *tt) {
    for (size_t i = 0; i < sl.graphs.size();) {
        if (sl.graphs[i].shape.tt == tt) {
            (void)hipGraphExecDestroy(sl.graphs[i].exec);
            sl.graphs.erase(sl.graphs.begin() + i);
        } else {
            i++;
        }
    }
}

// perft of one root through a total-only graph on the context's batch slot:
// no host round trip between levels.  *fits is false (and nothing is
// returned) when a level does not fit the arena; the caller then runs the
// host-driven search, which splits levels.
int perft_graph(Engine &e, const Pos &root, u32 depth, int v, u64 *out, bool *fits) {
    *fits = false;
    if (!e.batchSlot) {
        e.batchSlot = new Slot();
        const int rc = slot_init(e, *e.batchSlot, e.arena, e.arenaBytes, e.stream);
        if (rc) return rc;
    }
    Slot &sl = *e.batchSlot;
    Plan &h = *sl.hPlan;
    memset(&h, 0, sizeof(Plan));
    h.base = sl.region;
    h.cap = sl.regionBytes;
    // level 0: the root (staged through the pinned plan copy) and its count
    h.lv[0].n = 1;
    h.lv[0].b = (Pos *)sl.region;
    h.lv[0].cnt = (u32 *)(sl.region + 256);
    h.top = 512;
    h.acc = e.bias;
    GraphShape sh{(int)depth, v, 0, 0, e.bpc, nullptr, 1, 0, grid_limit(e, 1, (int)depth)};
    hipGraphExec_t exec;
    int rc = get_graph(e, sl, sh, &exec);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(sl.dPlan, &h, sizeof(Plan), hipMemcpyHostToDevice, sl.stream));
    Pos *staged = reinterpret_cast<Pos *>(e.hostBuf + 8);
    *staged = root;
    HIP_TRY(hipMemcpyAsync(h.lv[0].b, staged, sizeof(Pos), hipMemcpyHostToDevice, sl.stream));
    HIP_TRY(hipGraphLaunch(exec, sl.stream));
    HIP_TRY(hipMemcpyAsync(&h, sl.dPlan, sizeof(Plan), hipMemcpyDeviceToHost, sl.stream));
    HIP_TRY(hipStreamSynchronize(sl.stream));
    if (h.overflow) return PG_OK;
    if (h.ovf) {
        set_error("perft: a perft total does not fit 64 bits");
        return PG_EOVERFLOW;
    }
    if (h.err) {
        set_error("perft: device consistency check failed (flag %u)", h.err);
        return PG_EINVAL;
    }
    *fits = true;
    *out = h.acc;
    e.stats.leaf_items = h.leafItems;
    e.stats.leaf_launches = depth >= 2 && v != SL_BYTE ? 1 : 0;
    for (u32 l = 1; l < depth && l < 16; l++) e.stats.nodes[l] = h.lv[l].n;
    if (depth >= 2 && depth - 1 < 16 && v != SL_BYTE) e.stats.nodes[depth - 1] = h.leafItems;
    if (depth >= 2 && v != SL_BYTE) e.stats.leaf_parents = h.lv[depth - 2].n;
    return PG_OK;
}

// Batched perft of n roots on slot sl.  Results in out[] (host), in the
// caller's order.
int batch_perft(Engine &e, Slot &sl, TTable *tt, const Pos *roots, u64 n, u32 depth, int v, int dedup, int bpc,
                u64 *out, pg_batch_stats *bs) {
    const auto t0 = std::chrono::steady_clock::now();
    if (n == 0) return PG_OK;
    if (depth == 0) {
        for (u64 i = 0; i < n; i++) out[i] = 1;
        return PG_OK;
    }
    const HostTables *ht = host_tables(nullptr);
    if (!ht) return PG_ETABLE;
    const Tables hv = ht->view();
    // order by side to move so that waves see one colour
    std::vector<u64> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](u64 a, u64 b) { return pos_side(roots[a]) < pos_side(roots[b]); });
    const bool keys = tt != nullptr || dedup;
    std::vector<Pos> hb(n);
    std::vector<Key128> hk(keys ? n : 0);
    for (u64 i = 0; i < n; i++) {
        hb[i] = roots[order[i]];
        if (keys) hk[i] = zobrist128(hb[i], hv);
    }
    // level-0 arrays at the start of the region
    size_t top = 0;
    auto take = [&](size_t bytes) -> u8 * {
        size_t a = (top + 255) & ~(size_t)255;
        if (a + bytes > sl.regionBytes) return nullptr;
        top = a + bytes;
        return sl.region + a;
    };
    DLevel L0{};
    L0.n = n;
    L0.b = (Pos *)take(n * sizeof(Pos));
    L0.k = keys ? (Key128 *)take(n * sizeof(Key128)) : nullptr;
    L0.cnt = (u32 *)take(n * 4);
    L0.perft = (u64 *)take(n * 8);
    L0.state = (u8 *)take(n);
    L0.dup = (u32 *)take(n * 4);
    if (!L0.dup) {
        set_error("batch of %llu roots does not fit the region", (unsigned long long)n);
        return PG_ENOMEM;
    }
    HIP_TRY(hipMemcpyAsync(L0.b, hb.data(), n * sizeof(Pos), hipMemcpyHostToDevice, sl.stream));
    if (keys) HIP_TRY(hipMemcpyAsync(L0.k, hk.data(), n * sizeof(Key128), hipMemcpyHostToDevice, sl.stream));
    GraphShape sh{(int)depth, v, keys ? 1 : 0, dedup ? 1 : 0, bpc > 0 ? bpc : e.bpc, tt, 0, 0,
                  e.batchGrid >= 0 ? e.batchGrid : grid_limit(e, n, (int)depth)};
    u64 leafItems = 0, launches = 0, splits = 0, peak = 0;
    double ms = 0;
    int rc = run_split(e, sl, sh, 0, n, L0, top, &leafItems, &ms, &launches, &splits, &peak);
    if (rc) return rc;
    std::vector<u64> res(n);
    HIP_TRY(hipMemcpyAsync(res.data(), L0.perft, n * 8, hipMemcpyDeviceToHost, sl.stream));
    HIP_TRY(hipStreamSynchronize(sl.stream));
    for (u64 i = 0; i < n; i++) out[order[i]] = res[i];
    if (bs) {
        bs->graph_ms = ms;
        bs->launches = launches;
        bs->splits = splits;
        bs->leaf_items = leafItems;
        bs->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        bs->graphs_cached = sl.graphs.size();
        bs->peak_bytes = peak;
    }
    return PG_OK;
}

}  // namespace pg

// ====================================================================== C ABI
extern "C" {

int pg_set_batch_grid(pg_ctx *ctx, int max_groups) {
    if (!ctx || max_groups < -1) return PG_EINVAL;
    reinterpret_cast<Engine *>(ctx)->batchGrid = max_groups;
    return PG_OK;
}

int pg_perft_batch(pg_ctx *ctx, const pg_pos48 *pos, uint32_t n, uint32_t depth, int slider, int dedup,
                   int blocks_per_cu, pg_tt *tt, uint64_t *out, pg_batch_stats *st) {
    if (!ctx || (n && (!pos || !out)) || depth >= TT_MAX_DEPTH || blocks_per_cu < 0 || blocks_per_cu > 64) return PG_EINVAL;
    Engine *e = reinterpret_cast<Engine *>(ctx);
    TTable *t = reinterpret_cast<TTable *>(tt);
    if (t && t->eng != e) return PG_EINVAL;
    const int v = slider == PG_SLIDER_AUTO ? e->slider : slider;
    if (v < 0 || v >= SL_COUNT) return PG_EINVAL;
    if (st) memset(st, 0, sizeof(*st));
    if (!n) return PG_OK;
    std::vector<Pos> roots(n);
    memcpy(roots.data(), pos, n * sizeof(Pos));
    for (uint32_t i = 0; i < n; i++)
        if (!pos_sane(roots[i])) {
            set_error("board %u needs exactly one king per side", i);
            return PG_EINVAL;
        }
    (void)hipSetDevice(e->device);
    if (!e->batchSlot) {
        e->batchSlot = new Slot();
        int rc = slot_init(*e, *e->batchSlot, e->arena, e->arenaBytes, e->stream);
        if (rc) return rc;
    }
    return batch_perft(*e, *e->batchSlot, t, roots.data(), n, depth, v, dedup, blocks_per_cu,
                       reinterpret_cast<u64 *>(out), st);
}

// Roots of several depths: one pg_perft_batch per depth present (each level
// loop runs one remaining depth for all its boards), shallowest first.
int pg_perft_batch_mixed(pg_ctx *ctx, const pg_pos48 *pos, const uint32_t *depths, uint32_t n, int slider, int dedup,
                         int blocks_per_cu, pg_tt *tt, uint64_t *out, pg_batch_stats *st) {
    if (!ctx || (n && (!pos || !depths || !out))) return PG_EINVAL;
    for (uint32_t i = 0; i < n; i++)
        if (depths[i] >= TT_MAX_DEPTH) {
            set_error("root %u: depth %u is beyond the largest supported (%d)", i, depths[i], TT_MAX_DEPTH - 1);
            return PG_EINVAL;
        }
    const auto t0 = std::chrono::steady_clock::now();
    if (st) memset(st, 0, sizeof(*st));
    std::vector<uint32_t> order(n);
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return depths[a] < depths[b]; });
    std::vector<pg_pos48> group;
    std::vector<uint64_t> res;
    for (uint32_t a = 0; a < n;) {
        const uint32_t d = depths[order[a]];
        uint32_t b = a;
        while (b < n && depths[order[b]] == d) b++;
        group.resize(b - a);
        res.resize(b - a);
        for (uint32_t k = a; k < b; k++) group[k - a] = pos[order[k]];
        pg_batch_stats gs{};
        const int rc = pg_perft_batch(ctx, group.data(), b - a, d, slider, dedup, blocks_per_cu, tt, res.data(), &gs);
        if (rc) {
            const std::string why = pg_last_error();  // set_error writes the buffer it returns
            set_error("%s (the roots of depth %u)", why.c_str(), d);
            return rc;
        }
        for (uint32_t k = a; k < b; k++) out[order[k]] = res[k - a];
        if (st) {
            st->graph_ms += gs.graph_ms;
            st->launches += gs.launches;
            st->splits += gs.splits;
            st->leaf_items += gs.leaf_items;
            st->graphs_cached = gs.graphs_cached;
            st->peak_bytes = std::max(st->peak_bytes, gs.peak_bytes);
        }
        a = b;
    }
    if (st) st->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return PG_OK;
}

}  // extern "C"

// ===================================================================== serving
// A server owns `slots` batch slots, each with its own stream, plan, graph
// cache and HBM region, sharing one optional TT.  submit() stages a batch and
// launches its graph asynchronously; poll() reports completion.  Batches on
// different slots overlap on the device.
namespace pg {

struct ServerSlot {
    Slot slot;
    hipEvent_t done = nullptr;
    bool busy = false;
    u64 n = 0, first = 0;
    u32 depth = 0;
    int v = 0, dedup = 0, wgCap = 0;
    std::vector<u64> order;
    u64 *hOut = nullptr;  // pinned results, slot capacity
    u8 *hIn = nullptr;    // pinned roots then keys, slot capacity (a pageable source would make the copy block)
    u64 cap = 0;
    DLevel L0{};
    size_t top0 = 0;
    double t0 = 0;
};

struct Server {
    Engine *eng = nullptr;
    TTable *tt = nullptr;
    std::vector<ServerSlot> slots;
    u8 *mem = nullptr;
    size_t bytesPerSlot = 0;
};

}  // namespace pg

struct pg_server : public pg::Server {};

static double host_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

extern "C" {

void pg_server_destroy(pg_server *s);

int pg_server_create(pg_ctx *ctx, uint32_t nslots, uint64_t bytes_per_slot, uint32_t max_roots, pg_tt *tt,
                     pg_server **out) {
    if (!ctx || !out || nslots == 0 || nslots > 16 || bytes_per_slot < (16ull << 20) || max_roots == 0) return PG_EINVAL;
    *out = nullptr;
    Engine *e = reinterpret_cast<Engine *>(ctx);
    TTable *t = reinterpret_cast<TTable *>(tt);
    if (t && t->eng != e) return PG_EINVAL;
    (void)hipSetDevice(e->device);
    size_t freeB = 0, totalB = 0;
    (void)hipMemGetInfo(&freeB, &totalB);
    if ((u64)nslots * bytes_per_slot > freeB) {
        set_error("%u slots of %llu bytes exceed free HBM (%zu)", nslots, (unsigned long long)bytes_per_slot, freeB);
        return PG_ENOMEM;
    }
    pg_server *s = new pg_server();
    s->eng = e;
    s->tt = t;
    s->bytesPerSlot = bytes_per_slot;
    if (t) t->users++;
    e->servers.push_back(s);
    if (hipMalloc(&s->mem, (size_t)nslots * bytes_per_slot) != hipSuccess) {
        pg::server_free(s);
        return PG_ENOMEM;
    }
    s->slots.resize(nslots);
    for (uint32_t i = 0; i < nslots; i++) {
        ServerSlot &ss = s->slots[i];
        hipStream_t st;
        int rc = PG_OK;
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) rc = PG_EHIP;
        if (!rc) rc = slot_init(*e, ss.slot, s->mem + (size_t)i * bytes_per_slot, bytes_per_slot, st);
        if (!rc && hipEventCreateWithFlags(&ss.done, hipEventDisableTiming) != hipSuccess) rc = PG_EHIP;
        if (!rc && hipHostMalloc(&ss.hOut, (size_t)max_roots * 8) != hipSuccess) rc = PG_ENOMEM;
        if (!rc && hipHostMalloc(&ss.hIn, (size_t)max_roots * (sizeof(Pos) + sizeof(Key128))) != hipSuccess)
            rc = PG_ENOMEM;
        if (rc) {
            pg_server_destroy(s);
            return rc;
        }
        ss.cap = max_roots;
    }
    *out = s;
    return PG_OK;
}

void pg_server_destroy(pg_server *s) {
    if (s) pg::server_free(s);
}

}  // extern "C"

void pg::server_free(Server *srv) {
    pg_server *s = static_cast<pg_server *>(srv);
    Engine *e = s->eng;
    if (s->tt) s->tt->users--;
    for (size_t i = 0; i < e->servers.size(); i++)
        if (e->servers[i] == srv) {
            e->servers.erase(e->servers.begin() + i);
            break;
        }
    (void)hipSetDevice(e->device);
    for (auto &ss : s->slots) {
        if (ss.slot.stream) (void)hipStreamSynchronize(ss.slot.stream);
        hipStream_t st = ss.slot.stream;
        slot_free(ss.slot);
        if (st) (void)hipStreamDestroy(st);
        if (ss.done) (void)hipEventDestroy(ss.done);
        if (ss.hOut) (void)hipHostFree(ss.hOut);
        if (ss.hIn) (void)hipHostFree(ss.hIn);
    }
    if (s->mem) (void)hipFree(s->mem);
    delete s;
}

extern "C" {

// Stage a batch on an idle slot and launch its graph; returns at once.
int pg_server_submit(pg_server *s, uint32_t slot, const pg_pos48 *pos, uint32_t n, uint32_t depth, int slider,
                     int dedup) {
    if (!s) return PG_EINVAL;
    return pg::server_submit(*s, slot, pos, n, depth, slider, dedup, 0);
}

int pg_server_submit_capped(pg_server *s, uint32_t slot, const pg_pos48 *pos, uint32_t n, uint32_t depth, int slider,
                            int dedup, uint32_t wg_per_cu) {
    if (!s || wg_per_cu > 64) return PG_EINVAL;
    return pg::server_submit(*s, slot, pos, n, depth, slider, dedup, (int)wg_per_cu);
}

}  // extern "C"

// wgCap > 0: the batch's kernels keep at most wgCap work-groups per CU
// resident, so that batches on the other slots find room on every CU.
int pg::server_submit(Server &srv, uint32_t slot, const pg_pos48 *pos, uint32_t n, uint32_t depth, int slider,
                      int dedup, int wgCap) {
    pg_server *s = static_cast<pg_server *>(&srv);
    if (slot >= s->slots.size() || !pos || n == 0 || depth < 1 || depth >= TT_MAX_DEPTH || wgCap < 0) return PG_EINVAL;
    ServerSlot &ss = s->slots[slot];
    if (ss.busy) return PG_EINVAL;
    if (n > ss.cap) return PG_EINVAL;
    Engine &e = *s->eng;
    const int v = slider == PG_SLIDER_AUTO ? e.slider : slider;
    if (v < 0 || v >= SL_COUNT) return PG_EINVAL;
    const HostTables *ht = host_tables(nullptr);
    if (!ht) return PG_ETABLE;
    (void)hipSetDevice(e.device);
    std::vector<Pos> roots(n);
    memcpy(roots.data(), pos, n * sizeof(Pos));
    for (uint32_t i = 0; i < n; i++)
        if (!pos_sane(roots[i])) return PG_EINVAL;
    ss.order.resize(n);
    std::iota(ss.order.begin(), ss.order.end(), 0);
    std::stable_sort(ss.order.begin(), ss.order.end(), [&](u64 a, u64 b) { return pos_side(roots[a]) < pos_side(roots[b]); });
    Pos *hb = reinterpret_cast<Pos *>(ss.hIn);
    Key128 *hk = reinterpret_cast<Key128 *>(ss.hIn + (size_t)n * sizeof(Pos));
    for (uint32_t i = 0; i < n; i++) {
        hb[i] = roots[ss.order[i]];
        hk[i] = zobrist128(hb[i], ht->view());
    }
    Slot &sl = ss.slot;
    size_t top = 0;
    auto take = [&](size_t bytes) -> u8 * {
        size_t a = (top + 255) & ~(size_t)255;
        if (a + bytes > sl.regionBytes) return nullptr;
        top = a + bytes;
        return sl.region + a;
    };
    DLevel L0{};
    L0.n = n;
    L0.b = (Pos *)take(n * sizeof(Pos));
    L0.k = (Key128 *)take(n * sizeof(Key128));
    L0.cnt = (u32 *)take(n * 4);
    L0.perft = (u64 *)take(n * 8);
    L0.state = (u8 *)take(n);
    L0.dup = (u32 *)take(n * 4);
    if (!L0.dup) return PG_ENOMEM;
    HIP_TRY(hipMemcpyAsync(L0.b, hb, n * sizeof(Pos), hipMemcpyHostToDevice, sl.stream));
    HIP_TRY(hipMemcpyAsync(L0.k, hk, n * sizeof(Key128), hipMemcpyHostToDevice, sl.stream));
    GraphShape sh{(int)depth, v, 1, dedup ? 1 : 0, e.bpc, s->tt, 0, wgCap, grid_limit(e, n, (int)depth)};
    hipGraphExec_t exec;
    int rc = get_graph(e, sl, sh, &exec);
    if (rc) return rc;
    Plan &h = *sl.hPlan;
    memset(&h, 0, sizeof(Plan));
    h.base = sl.region;
    h.cap = sl.regionBytes;
    h.top = top;
    h.bias = e.bias;
    h.lv[0] = L0;
    HIP_TRY(hipMemcpyAsync(sl.dPlan, &h, sizeof(Plan), hipMemcpyHostToDevice, sl.stream));
    HIP_TRY(hipEventRecord(sl.ev0, sl.stream));  // the batch's device time (pg_batch_stats.graph_ms)
    HIP_TRY(hipGraphLaunch(exec, sl.stream));
    HIP_TRY(hipEventRecord(sl.ev1, sl.stream));
    HIP_TRY(hipMemcpyAsync(&h, sl.dPlan, sizeof(Plan), hipMemcpyDeviceToHost, sl.stream));
    HIP_TRY(hipMemcpyAsync(ss.hOut, L0.perft, n * 8, hipMemcpyDeviceToHost, sl.stream));
    HIP_TRY(hipEventRecord(ss.done, sl.stream));
    ss.busy = true;
    ss.n = n;
    ss.depth = depth;
    ss.v = v;
    ss.dedup = dedup;
    ss.wgCap = wgCap;
    ss.L0 = L0;
    ss.top0 = top;
    ss.t0 = host_ms();
    return PG_OK;
}

extern "C" {

// 1: done (out[] filled, in submission order), 0: still running, < 0: error.
// A batch that overflowed its region is finished here synchronously in
// halves.
int pg_server_poll(pg_server *s, uint32_t slot, uint64_t *out, pg_batch_stats *st) {
    if (!s || slot >= s->slots.size()) return PG_EINVAL;
    ServerSlot &ss = s->slots[slot];
    if (!ss.busy) return PG_EINVAL;
    const hipError_t q = hipEventQuery(ss.done);
    if (q == hipErrorNotReady) return 0;
    if (q != hipSuccess) return PG_EHIP;
    ss.busy = false;
    Engine &e = *s->eng;
    Slot &sl = ss.slot;
    const Plan &h = *sl.hPlan;
    if (h.ovf) {
        set_error("served batch: a perft total does not fit 64 bits");
        return PG_EOVERFLOW;
    }
    if (h.err) {
        set_error("device consistency check failed in served batch (flag %u)", h.err);
        return PG_EINVAL;
    }
    u64 items = h.leafItems, launches = 1, splits = 0, peak = h.peak;
    float first = 0;
    (void)hipEventElapsedTime(&first, sl.ev0, sl.ev1);  // read before a split reuses the events
    double ms = first;
    if (h.overflow) {
        GraphShape sh{(int)ss.depth, ss.v, 1, ss.dedup ? 1 : 0, e.bpc, s->tt, 0, ss.wgCap,
                      grid_limit(e, ss.n, (int)ss.depth)};
        (void)hipSetDevice(e.device);
        items = 0;
        launches = 0;
        int rc = run_split(e, sl, sh, 0, ss.n, ss.L0, ss.top0, &items, &ms, &launches, &splits, &peak);
        if (rc) return rc;
        HIP_TRY(hipMemcpyAsync(ss.hOut, ss.L0.perft, ss.n * 8, hipMemcpyDeviceToHost, sl.stream));
        HIP_TRY(hipStreamSynchronize(sl.stream));
    }
    if (out)
        for (u64 i = 0; i < ss.n; i++) out[ss.order[i]] = ss.hOut[i];
    if (st) {
        memset(st, 0, sizeof(*st));
        st->total_ms = host_ms() - ss.t0;
        st->graph_ms = ms;
        st->launches = launches;
        st->splits = splits;
        st->leaf_items = items;
        st->graphs_cached = sl.graphs.size();
        st->peak_bytes = peak;
    }
    return 1;
}

int pg_server_prepare(pg_server *s, uint32_t depth, uint32_t max_roots, int slider, int dedup, uint32_t wg_per_cu,
                      uint32_t *graphs) {
    if (graphs) *graphs = 0;
    if (!s || depth < 1 || depth >= TT_MAX_DEPTH || max_roots == 0 || wg_per_cu > 64) return PG_EINVAL;
    Engine &e = *s->eng;
    const int v = slider == PG_SLIDER_AUTO ? e.slider : slider;
    if (v < 0 || v >= SL_COUNT) return PG_EINVAL;
    (void)hipSetDevice(e.device);
    // the grids of the batch sizes 1..max_roots (grid_limit takes few values)
    std::vector<int> limits;
    for (uint32_t n = 1; n <= max_roots; n++) {
        const int g = grid_limit(e, n, (int)depth);
        if (std::find(limits.begin(), limits.end(), g) == limits.end()) limits.push_back(g);
    }
    uint32_t made = 0;
    for (auto &ss : s->slots) {
        if (ss.busy) continue;
        for (const int g : limits) {
            GraphShape sh{(int)depth, v, 1, dedup ? 1 : 0, e.bpc, s->tt, 0, (int)wg_per_cu, g};
            const size_t before = ss.slot.graphs.size();
            hipGraphExec_t exec;
            const int rc = get_graph(e, ss.slot, sh, &exec);
            if (rc) return rc;
            made += (uint32_t)(ss.slot.graphs.size() - before);
        }
    }
    if (graphs) *graphs = made;
    return PG_OK;
}

int pg_server_slots(pg_server *s, uint32_t *nslots, uint64_t *bytes_per_slot) {
    if (!s) return PG_EINVAL;
    if (nslots) *nslots = (uint32_t)s->slots.size();
    if (bytes_per_slot) *bytes_per_slot = s->bytesPerSlot;
    return PG_OK;
}

int pg_device_memory(pg_ctx *ctx, uint64_t *free_bytes, uint64_t *total_bytes) {
    if (!ctx) return PG_EINVAL;
    Engine *e = reinterpret_cast<Engine *>(ctx);
    (void)hipSetDevice(e->device);
    size_t f = 0, t = 0;
    HIP_TRY(hipMemGetInfo(&f, &t));
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
    return PG_OK;
}

}  // extern "C"
