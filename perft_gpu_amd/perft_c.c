/* perft_c.c -- a plain C caller of the C ABI in include/perft_gpu.h: what a
 * C or C++ host program (the reference's perft.cu main, or a cgo/JNI shim)
 * does to run perft on the GPU, with no Python in between.
 *
 *   perft_c "<fen>" <depth> [tt_gigabytes]
 *
 * Prints one line per depth from 1 to <depth> in the reference's format
 * (launcher.h:1496-1498) and exits non-zero on any library error. */
#define _POSIX_C_SOURCE 199309L
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

int main(int argc, char **argv) {
    if (argc < 3) {
        fprintf(stderr, "usage: %s \"<fen>\" <depth> [tt_gigabytes]\n", argv[0]);
        return 2;
    }
    const unsigned depth = (unsigned)atoi(argv[2]);
    const double ttGB = argc > 3 ? atof(argv[3]) : 0.0;
    pg_pos48 pos;
    if (pg_fen_to_pos(argv[1], &pos) != PG_OK) {
        fprintf(stderr, "bad FEN: %s\n", pg_last_error());
        return 1;
    }
    pg_ctx *ctx = NULL;
    pg_config cfg = {0};
    cfg.slider = PG_SLIDER_AUTO;
    if (pg_init(0, &cfg, &ctx) != PG_OK) {
        fprintf(stderr, "pg_init failed: %s\n", pg_last_error());
        return 1;
    }
    pg_tt *tt = NULL;
    if (ttGB > 0 && pg_tt_create(ctx, (uint64_t)(ttGB * (double)(1ull << 30)), &tt) != PG_OK) {
        fprintf(stderr, "pg_tt_create failed: %s\n", pg_last_error());
        pg_shutdown(ctx);
        return 1;
    }
    int rc = 0;
    for (unsigned d = 1; d <= depth && rc == 0; d++) {
        uint64_t n = 0;
        const double t0 = now_s();
        const int r = tt ? pg_perft_tt(ctx, tt, &pos, d, PG_SLIDER_AUTO, &n) : pg_perft(ctx, &pos, d, PG_SLIDER_AUTO, &n);
        const double dt = now_s() - t0;
        if (r != PG_OK) {
            fprintf(stderr, "perft(%u) failed: %s\n", d, pg_last_error());
            rc = 1;
            break;
        }
        printf("\nGPU Perft %u: %llu,   Time taken: %g seconds, nps: %llu\n", d, (unsigned long long)n, dt,
               (unsigned long long)(dt > 0 ? n / dt : 0));
    }
    if (tt) pg_tt_destroy(tt);
    pg_shutdown(ctx);
    return rc;
}
