// serve.cpp -- the serving loop: perft queries of mixed positions and depths
// arrive one at a time; the loop groups the waiting ones by depth, admits a
// group into an idle slot as one batched graph launch when its estimated
// level buffers fit the slot's HBM region, and hands back each query's result
// and latency.  Batches on different slots overlap on the device and share
// one optional transposition table.  No reference counterpart: the reference
// runs one perft at a time (perftLauncher, launcher.h:1455).
//
// The loop runs inside pg_serve_poll (the caller's thread drives it; no
// thread of its own).  Policy (pg_serve_policy, a pure function):
//  * among depths with waiting queries, the one whose oldest query arrived
//    first goes next;
//  * batches of depth >= long_depth may hold at most long_slots slots at once,
//    so short queries never wait behind a slot-full of long ones;
//  * a batch takes as many queries of that depth as max_batch and the slot's
//    region allow, by a per-depth estimate of region bytes per query, learned
//    from the peak bytes of the depth's completed batches (a batch that
//    still overflows its region is finished in halves by the slot);
//  * the kernels of a long batch keep at most long_wg_per_cu work-groups per
//    CU resident, so that short batches on other slots find free CUs; while
//    at most light_queue queries wait (a light load) and short queries are
//    about (one waiting or running, or one submitted in the last
//    kShortRecentMs) the limit is light_wg_per_cu, lower, since short
//    queries' latency then matters more than long ones' throughput.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <unordered_map>
#include <vector>
#include "../../include/perft_gpu.h"
#include "movegen.h"

namespace {

constexpr int kMaxDepth = 16;
constexpr double kShortRecentMs = 10.0;  // a short query this recent may have company on its way

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// region bytes per root before any batch of that depth has finished: the
// levels down to ply depth-2 (the leaf plies stay in LDS) at about 110 bytes
// per board (board, key, count, perft, state, owner, offsets, transient
// moves, parents and duplicate set), 40 children per board
double initial_estimate(uint32_t depth) {
    if (depth <= 2) return 4096;
    return 110.0 * std::pow(40.0, (double)(depth - 2)) + (1 << 20);
}

struct Query {
    uint64_t id, tag;
    pg_pos48 pos;
    uint32_t depth;
    double arrived;
    double admitted;  // batch launched
};

struct SlotState {
    bool busy = false;
    uint32_t depth = 0;
    uint64_t graphs = 0;        // graphs the slot holds (as of its last finished batch)
    std::vector<uint64_t> ids;  // queries of the running batch, submission order
};

}  // namespace

struct pg_serve {
    pg_ctx *ctx = nullptr;
    pg_server *srv = nullptr;
    pg_serve_config cfg{};
    uint64_t regionBytes = 0;
    std::vector<SlotState> slots;
    std::deque<uint64_t> waiting[kMaxDepth];
    std::unordered_map<uint64_t, Query> queries;  // waiting or running
    std::deque<pg_serve_result> done;
    double est[kMaxDepth];
    uint64_t nextId = 1;
    std::vector<double> latencies;
    std::vector<uint64_t> results;
    std::vector<pg_pos48> stage;
    pg_serve_stats st{};
    double firstSubmit = -1, lastDone = 0;
    double lastShort = -1e300;  // submit time of the latest short query
};

extern "C" {

int pg_serve_policy(const uint32_t waiting[16], const double oldest[16], uint32_t busy_long, uint32_t idle_slots,
                    const pg_serve_config *cfg, uint64_t region_bytes, const double est_bytes[16], uint32_t *depth,
                    uint32_t *count) {
    if (!waiting || !oldest || !cfg || !est_bytes || !depth || !count) return PG_EINVAL;
    *depth = 0;
    *count = 0;
    if (idle_slots == 0) return 0;
    int best = -1;
    for (int d = 1; d < kMaxDepth; d++) {
        if (!waiting[d]) continue;
        if ((uint32_t)d >= cfg->long_depth && busy_long >= cfg->long_slots) continue;
        if (best < 0 || oldest[d] < oldest[best]) best = d;
    }
    if (best < 0) return 0;
    const double per = est_bytes[best] > 1 ? est_bytes[best] : 1;
    uint64_t fit = (uint64_t)((double)region_bytes / per);
    if (fit < 1) fit = 1;  // a query too large for the estimate still runs, splitting inside its slot
    uint64_t k = waiting[best];
    if (k > cfg->max_batch) k = cfg->max_batch;
    if (k > fit) k = fit;
    *depth = (uint32_t)best;
    *count = (uint32_t)k;
    return 1;
}

uint32_t pg_serve_wg_cap(const pg_serve_config *cfg, uint32_t depth, uint32_t queued, int shorts_about) {
    if (!cfg || depth < cfg->long_depth) return 0;
    return shorts_about && queued <= cfg->light_queue ? cfg->light_wg_per_cu : cfg->long_wg_per_cu;
}

void pg_serve_destroy(pg_serve *s);

int pg_serve_create(pg_ctx *ctx, pg_tt *tt, const pg_serve_config *config, pg_serve **out) {
    if (!ctx || !out) return PG_EINVAL;
    *out = nullptr;
    pg_serve_config c{};
    if (config) c = *config;
    if (!c.slots) {
        // one hardware queue per slot stream: the runtime maps streams to
        // GPU_MAX_HW_QUEUES queues (4 unless set) round-robin, and the
        // context's and the null stream take two; a slot sharing a queue
        // would wait for the other stream's batch in order
        const char *q = getenv("GPU_MAX_HW_QUEUES");
        const int queues = q && atoi(q) > 0 ? atoi(q) : 4;
        c.slots = (uint32_t)std::max(1, std::min(4, queues - 2));
    }
    if (!c.max_batch) c.max_batch = 1024;
    if (!c.long_depth) c.long_depth = 6;
    if (!c.long_slots) c.long_slots = c.slots > 1 ? c.slots / 2 : 1;
    if (!c.long_wg_per_cu) c.long_wg_per_cu = 2;
    if (!c.light_wg_per_cu) c.light_wg_per_cu = std::min(1u, c.long_wg_per_cu);
    if (!c.light_queue) c.light_queue = 16;
    if (c.slots > 16 || c.long_slots > c.slots || c.long_wg_per_cu > 64 || c.light_wg_per_cu > 64) return PG_EINVAL;
    if (!c.bytes_per_slot) {
        // admission by free HBM: 60% of what the device has free now, over the slots
        uint64_t freeB = 0, totalB = 0;
        int rc = pg_device_memory(ctx, &freeB, &totalB);
        if (rc) return rc;
        c.bytes_per_slot = std::min<uint64_t>(24ull << 30, (uint64_t)(freeB * 0.6) / c.slots);
    }
    pg_serve *s = new pg_serve();
    s->ctx = ctx;
    s->cfg = c;
    int rc = pg_server_create(ctx, c.slots, c.bytes_per_slot, c.max_batch, tt, &s->srv);
    if (rc) {
        delete s;
        return rc;
    }
    s->regionBytes = c.bytes_per_slot;
    s->slots.resize(c.slots);
    for (int d = 0; d < kMaxDepth; d++) s->est[d] = initial_estimate((uint32_t)d);
    // every graph shape of the short depths and the first long ones captured
    // now: a capture during the run holds up admission for a millisecond or
    // two, long enough to show in short queries' tails
    uint64_t prepared = 0;
    for (uint32_t d = 1; d <= std::min<uint32_t>(c.long_depth + 1, kMaxDepth - 1) && rc == PG_OK; d++) {
        const uint32_t caps[2] = {d < c.long_depth ? 0u : c.light_wg_per_cu, c.long_wg_per_cu};
        for (int k = 0; k < (d < c.long_depth ? 1 : 2) && rc == PG_OK; k++) {
            uint32_t made = 0;
            rc = pg_server_prepare(s->srv, d, c.max_batch, c.slider, c.dedup, caps[k], &made);
            prepared += made;
        }
    }
    if (rc) {
        pg_serve_destroy(s);
        return rc;
    }
    s->st.graphs = prepared;
    for (auto &sl : s->slots) sl.graphs = prepared / c.slots;
    s->results.resize(c.max_batch);
    s->stage.resize(c.max_batch);
    *out = s;
    return PG_OK;
}

void pg_serve_destroy(pg_serve *s) {
    if (!s) return;
    pg_server_destroy(s->srv);
    delete s;
}

int pg_serve_submit(pg_serve *s, const pg_pos48 *pos, uint32_t depth, uint64_t tag, uint64_t *id) {
    if (!s || !pos || depth < 1 || depth >= kMaxDepth) return PG_EINVAL;
    pg::Pos p;
    memcpy(&p, pos, sizeof(p));
    if (!pg::pos_sane(p)) return PG_EINVAL;  // rejected here, so it cannot fail the batch it would join
    Query q;
    q.id = s->nextId++;
    q.tag = tag;
    q.pos = *pos;
    q.depth = depth;
    q.arrived = now_ms();
    q.admitted = q.arrived;
    if (s->firstSubmit < 0) s->firstSubmit = q.arrived;
    s->queries.emplace(q.id, q);
    s->waiting[depth].push_back(q.id);
    if (depth < s->cfg.long_depth) s->lastShort = q.arrived;
    s->st.submitted++;
    if (id) *id = q.id;
    return PG_OK;
}

// completes finished batches, admits waiting queries into idle slots, and
// copies up to max finished results to out
int pg_serve_poll(pg_serve *s, pg_serve_result *out, uint32_t max, uint32_t *n) {
    if (!s || (max && !out)) return PG_EINVAL;
    if (n) *n = 0;
    // 1. finished batches
    for (uint32_t i = 0; i < s->slots.size(); i++) {
        SlotState &sl = s->slots[i];
        if (!sl.busy) continue;
        pg_batch_stats bs{};
        const int r = pg_server_poll(s->srv, i, s->results.data(), &bs);
        if (r == 0) continue;
        const double t = now_ms();
        for (size_t j = 0; j < sl.ids.size(); j++) {
            auto it = s->queries.find(sl.ids[j]);
            if (it == s->queries.end()) continue;
            pg_serve_result res{};
            res.id = it->first;
            res.tag = it->second.tag;
            res.depth = it->second.depth;
            res.status = r == 1 ? PG_OK : r;
            res.perft = r == 1 ? s->results[j] : 0;
            res.latency_ms = t - it->second.arrived;
            res.queued_ms = it->second.admitted - it->second.arrived;
            s->latencies.push_back(res.latency_ms);
            if (r == 1) s->st.leaves += res.perft;
            s->done.push_back(res);
            s->queries.erase(it);
        }
        s->st.completed += sl.ids.size();
        s->st.graphs += bs.graphs_cached > sl.graphs ? bs.graphs_cached - sl.graphs : 0;
        sl.graphs = std::max(sl.graphs, bs.graphs_cached);
        s->st.launches += bs.launches;
        s->st.splits += bs.splits;
        s->st.device_ms += bs.graph_ms;
        s->lastDone = t;
        if (r == 1 && !sl.ids.empty() && bs.peak_bytes) {
            // learned region bytes per query of this depth, with headroom
            const double per = 1.25 * (double)bs.peak_bytes / (double)sl.ids.size();
            double &e = s->est[sl.depth];
            e = bs.splits ? std::max(e, 2.0 * per) : std::max(per, 0.5 * e);
        }
        sl.busy = false;
        sl.ids.clear();
    }
    // 2. admission
    for (;;) {
        uint32_t idle = 0, busyLong = 0, slot = 0;
        bool found = false;
        for (uint32_t i = 0; i < s->slots.size(); i++) {
            if (s->slots[i].busy) {
                if (s->slots[i].depth >= s->cfg.long_depth) busyLong++;
            } else {
                idle++;
                if (!found) slot = i, found = true;
            }
        }
        uint32_t w[kMaxDepth];
        double oldest[kMaxDepth];
        for (int d = 0; d < kMaxDepth; d++) {
            w[d] = (uint32_t)s->waiting[d].size();
            oldest[d] = w[d] ? s->queries[s->waiting[d].front()].arrived : 0;
        }
        uint32_t depth = 0, count = 0;
        if (pg_serve_policy(w, oldest, busyLong, idle, &s->cfg, s->regionBytes, s->est, &depth, &count) != 1) break;
        SlotState &sl = s->slots[slot];
        sl.ids.clear();
        const double admitted = now_ms();
        for (uint32_t j = 0; j < count; j++) {
            const uint64_t id = s->waiting[depth].front();
            s->waiting[depth].pop_front();
            sl.ids.push_back(id);
            Query &q = s->queries[id];
            q.admitted = admitted;
            s->stage[j] = q.pos;
        }
        uint32_t queued = 0;
        bool shortsAbout = admitted - s->lastShort < kShortRecentMs;
        for (int d = 0; d < kMaxDepth; d++) {
            queued += (uint32_t)s->waiting[d].size();
            if (d < (int)s->cfg.long_depth && !s->waiting[d].empty()) shortsAbout = true;
        }
        for (auto &x : s->slots)
            if (x.busy && x.depth < s->cfg.long_depth) shortsAbout = true;
        const uint32_t cap = pg_serve_wg_cap(&s->cfg, depth, queued, shortsAbout);
        const int rc = pg_server_submit_capped(s->srv, slot, s->stage.data(), count, depth, s->cfg.slider, s->cfg.dedup,
                                               cap);
        const double took = now_ms() - admitted;
        s->st.submit_ms_max = std::max(s->st.submit_ms_max, took);
        s->st.submit_ms_total += took;
        if (rc != PG_OK) {
            // the batch could not start: its queries finish with the error
            const double t = now_ms();
            for (uint64_t id : sl.ids) {
                auto it = s->queries.find(id);
                pg_serve_result res{};
                res.id = id;
                res.tag = it->second.tag;
                res.depth = depth;
                res.status = rc;
                res.latency_ms = t - it->second.arrived;
                res.queued_ms = res.latency_ms;
                s->done.push_back(res);
                s->queries.erase(it);
            }
            s->st.completed += sl.ids.size();
            sl.ids.clear();
            continue;
        }
        sl.busy = true;
        sl.depth = depth;
        s->st.batches++;
        uint32_t busy = 0, longs = 0;
        for (auto &x : s->slots) {
            busy += x.busy;
            longs += x.busy && x.depth >= s->cfg.long_depth;
        }
        s->st.max_busy_slots = std::max(s->st.max_busy_slots, busy);
        s->st.max_long_slots = std::max(s->st.max_long_slots, longs);
    }
    // 3. hand out finished queries
    uint32_t k = 0;
    while (k < max && !s->done.empty()) {
        out[k++] = s->done.front();
        s->done.pop_front();
    }
    if (n) *n = k;
    return PG_OK;
}

int pg_serve_get_stats(pg_serve *s, pg_serve_stats *out) {
    if (!s || !out) return PG_EINVAL;
    pg_serve_stats st = s->st;
    st.waiting = 0;
    for (int d = 0; d < kMaxDepth; d++) st.waiting += (uint32_t)s->waiting[d].size();
    st.busy_slots = 0;
    for (auto &x : s->slots) st.busy_slots += x.busy;
    if (!s->latencies.empty()) {
        std::vector<double> v = s->latencies;
        std::sort(v.begin(), v.end());
        auto pct = [&](double p) {
            // linear interpolation between closest ranks
            const double r = p / 100.0 * (double)(v.size() - 1);
            const size_t lo = (size_t)r, hi = std::min(lo + 1, v.size() - 1);
            return v[lo] + (v[hi] - v[lo]) * (r - (double)lo);
        };
        st.p50_ms = pct(50);
        st.p99_ms = pct(99);
        double sum = 0;
        for (double x : v) sum += x;
        st.mean_ms = sum / (double)v.size();
        st.max_ms = v.back();
    }
    st.elapsed_ms = s->firstSubmit >= 0 ? s->lastDone - s->firstSubmit : 0;
    for (int d = 0; d < kMaxDepth; d++) st.est_bytes_per_query[d] = s->est[d];
    *out = st;
    return PG_OK;
}

}  // extern "C"
