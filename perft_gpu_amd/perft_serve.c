/* perft_serve.c -- a plain C host serving a stream of perft queries through
 * the C ABI's serving loop (pg_serve_*), with no Python in between.
 *
 *   perft_serve <queries> <rate_per_s> [tt_gigabytes] [seed]
 *
 * Draws seeded random positions (pg_random_positions, ply 0-60) with depths
 * 3/4/5/6 at 40/30/20/10%, submits them at Poisson arrival times, drives the
 * loop with pg_serve_poll until every query is answered, then checks each
 * answer against pg_perft of the same query alone.  Prints throughput and
 * p50/p99 latency; exits non-zero on any error or mismatch. */
#define _POSIX_C_SOURCE 200112L
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <time.h>

#include "../include/perft_gpu.h"

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + ts.tv_nsec * 1e-9;
}

static uint64_t rng_state;
static double uniform(void) { /* splitmix64 -> [0, 1) */
    uint64_t z = (rng_state += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return (double)((z ^ (z >> 31)) >> 11) / 9007199254740992.0;
}

int main(int argc, char **argv) {
    if (argc < 3) {
        fprintf(stderr, "usage: %s <queries> <rate_per_s> [tt_gigabytes] [seed]\n", argv[0]);
        return 2;
    }
    const uint32_t n = (uint32_t)atoi(argv[1]);
    const double rate = atof(argv[2]);
    const double ttGB = argc > 3 ? atof(argv[3]) : 0.0;
    rng_state = argc > 4 ? strtoull(argv[4], NULL, 10) : 12345;
    if (n == 0 || rate <= 0) return 2;
    pg_pos48 *pos = malloc(n * sizeof(pg_pos48));
    uint32_t *plies = malloc(n * sizeof(uint32_t)), *depth = malloc(n * sizeof(uint32_t));
    uint64_t *answer = calloc(n, sizeof(uint64_t));
    double *arrive = malloc(n * sizeof(double));
    if (!pos || !plies || !depth || !answer || !arrive) return 1;
    if (pg_random_positions(rng_state, n, 0, 60, pos, plies) != PG_OK) return 1;
    double t = 0;
    for (uint32_t i = 0; i < n; i++) {
        const double u = uniform();
        depth[i] = u < 0.4 ? 3 : u < 0.7 ? 4 : u < 0.9 ? 5 : 6;
        t += -log(1.0 - uniform()) / rate;
        arrive[i] = t;
    }
    /* four slot streams beside the context's stream and the runtime's null
     * stream: with fewer hardware queues than that, short batches queue
     * behind long ones on a shared queue.  The HIP runtime reads this once,
     * at its initialization (the pg_init below). */
    const char *hwq = getenv("GPU_MAX_HW_QUEUES");
    if (!hwq || atoi(hwq) < 8) setenv("GPU_MAX_HW_QUEUES", "8", 1);
    pg_ctx *ctx = NULL;
    if (pg_init(0, NULL, &ctx) != PG_OK) {
        fprintf(stderr, "pg_init failed: %s\n", pg_last_error());
        return 1;
    }
    pg_tt *tt = NULL;
    if (ttGB > 0 && pg_tt_create(ctx, (uint64_t)(ttGB * (double)(1ull << 30)), &tt) != PG_OK) {
        fprintf(stderr, "pg_tt_create failed: %s\n", pg_last_error());
        return 1;
    }
    pg_serve_config cfg = {0};
    cfg.slots = 4;
    cfg.bytes_per_slot = 4ull << 30;
    cfg.dedup = 1;
    pg_serve *srv = NULL;
    if (pg_serve_create(ctx, tt, &cfg, &srv) != PG_OK) {
        fprintf(stderr, "pg_serve_create failed: %s\n", pg_last_error());
        return 1;
    }
    int rc = 0;
    uint32_t next = 0, finished = 0;
    pg_serve_result res[256];
    const double t0 = now_s();
    while (finished < n && rc == 0) {
        const double el = now_s() - t0;
        while (next < n && arrive[next] <= el) {
            if (pg_serve_submit(srv, &pos[next], depth[next], next, NULL) != PG_OK) {
                fprintf(stderr, "submit failed: %s\n", pg_last_error());
                rc = 1;
                break;
            }
            next++;
        }
        uint32_t got = 0;
        if (pg_serve_poll(srv, res, 256, &got) != PG_OK) rc = 1;
        for (uint32_t k = 0; k < got; k++) {
            if (res[k].status != PG_OK) {
                fprintf(stderr, "query %llu failed: %d\n", (unsigned long long)res[k].tag, res[k].status);
                rc = 1;
            }
            answer[res[k].tag] = res[k].perft;
            finished++;
        }
    }
    const double elapsed = now_s() - t0;
    pg_serve_stats st;
    pg_serve_get_stats(srv, &st);
    uint64_t leaves = 0, bad = 0;
    for (uint32_t i = 0; i < n && rc == 0; i++) { /* the same query alone */
        uint64_t alone = 0;
        if (pg_perft(ctx, &pos[i], depth[i], PG_SLIDER_AUTO, &alone) != PG_OK || alone != answer[i]) bad++;
        leaves += answer[i];
    }
    printf("served %u queries at %.0f/s in %.3f s: %.1f G leaves/s, %llu batches, %u slots busy at most, "
           "%.2f batches on the device on average, p50 %.3f ms, p99 %.3f ms, mismatches %llu\n",
           n, rate, elapsed, leaves / elapsed / 1e9, (unsigned long long)st.batches, st.max_busy_slots,
           st.elapsed_ms > 0 ? st.device_ms / st.elapsed_ms : 0.0, st.p50_ms, st.p99_ms, (unsigned long long)bad);
    pg_serve_destroy(srv);
    if (tt) pg_tt_destroy(tt);
    pg_shutdown(ctx);
    free(pos);
    free(plies);
    free(depth);
    free(answer);
    free(arrive);
    return rc || bad ? 1 : 0;
}
