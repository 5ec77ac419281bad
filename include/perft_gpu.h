/*
 * perft_gpu.h -- C ABI of the MI355X (gfx950) perft engine.
 *
 * Plain pointers and sizes only.  Every entry point returns PG_OK (0) or a
 * negative PG_E* code and never exits the process.  Output buffers are
 * caller-owned host memory; the library owns all device memory.
 *
 * The reference (ankan-ban/perft_gpu) has no plugin API: its boundary is the
 * set of C++ functions main() and launcher.h call.  Each entry below names
 * the reference interface it replaces.
 */
#ifndef PERFT_GPU_H
#define PERFT_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Byte-identical to the reference's HexaBitBoardPosition (chess.h:208-230):
 * six bitboards, game state packed in the rank-1/rank-8 bytes of `pawns`
 * (bits 0-1 white castle, 2-3 black castle, 4-7 en-passant file + 1,
 * 56-62 half-move counter, 63 side to move). */
typedef struct pg_pos48 {
    uint64_t whitePieces, pawns, knights, bishopQueens, rookQueens, kings;
} pg_pos48;

/* HashKey128b (chess.h:411-452): lowPart, highPart */
typedef struct pg_key128 {
    uint64_t lo, hi;
} pg_key128;

/* Moves are the reference's 16-bit CMove (chess.h:234-293):
 * from | to << 6 | flags << 12. */
typedef uint16_t pg_move;

enum {
    PG_OK = 0,
    PG_EINVAL = -1,   /* bad argument */
    PG_ENOMEM = -2,   /* device memory budget exhausted */
    PG_EHIP = -3,     /* HIP runtime error */
    PG_ENODEV = -4,   /* no usable GPU */
    PG_ETABLE = -5,   /* lookup-table construction failed */
    PG_EOVERFLOW = -6 /* a result does not fit the requested type */
};

/* Slider-attack variants (switches.h:102-125 selects one at compile time in
 * the reference; here each is a kernel instantiation). */
enum {
    PG_SLIDER_AUTO = -1,         /* the context's default (autotuned) */
    PG_SLIDER_KOGGE_STONE = 0,   /* USE_SLIDING_LUT 0 */
    PG_SLIDER_MAGIC_PLAIN = 1,   /* USE_FANCY_MAGICS 0 */
    PG_SLIDER_MAGIC_FANCY = 2,   /* USE_FANCY_MAGICS 1 (reference default) */
    PG_SLIDER_MAGIC_BYTE = 3     /* USE_BYTE_LOOKUP_FANCY 1, tables in LDS */
};

typedef struct pg_config {
    uint64_t arena_bytes;  /* device tree arena; 0 = default (48 GiB, capped by free HBM) */
    int slider;            /* default slider variant, PG_SLIDER_AUTO = Kogge-Stone */
    int block_size;        /* work-group size for the BFS kernels, 0 = 256 */
    int blocks_per_cu;     /* resident work-groups per CU for grid-stride kernels, 0 = 8 */
    int timing;            /* 1: time each leaf-kernel launch with HIP events */
    int leaf_mode;         /* 0: last two plies fused in LDS (default), 1: materialized item list */
} pg_config;

typedef struct pg_stats {
    double total_ms;          /* wall time of the last perft call */
    double leaf_kernel_ms;    /* summed duration of leaf-kernel launches (timing=1) */
    uint64_t leaf_launches;   /* leaf-kernel launches in the last call */
    uint64_t leaf_items;      /* (parent, move) items the leaf kernels consumed */
    uint64_t leaf_parents;    /* parent boards the leaf kernels read */
    uint64_t leaves;          /* the perft result */
    uint64_t splits;          /* level splits forced by the arena budget */
    uint64_t peak_arena;      /* peak arena bytes in use */
    uint64_t nodes[16];       /* boards materialized per ply below the root */
} pg_stats;

/* ------------------------------------------------------------ host side */

/* Utils::readFENString + Utils::board088ToHexBB (util.cpp:189, :92).
 * Clocks are not read, as in the reference. */
int pg_fen_to_pos(const char *fen, pg_pos48 *out);
/* Inverse, for display and fixtures (Utils::dispBoard has no FEN writer). */
int pg_pos_to_fen(const pg_pos48 *pos, char *buf, int buflen);
/* MoveGeneratorBitboard::computeZobristKey128b (MoveGeneratorBitboard.h:3914)
 * with the reference's key sets (randoms.cpp via :985-986). */
int pg_zobrist128(const pg_pos48 *pos, pg_key128 *out);
/* MoveGeneratorBitboard::generateMoves on the host (:2471); up to 218 moves. */
int pg_host_generate_moves(const pg_pos48 *pos, pg_move *out, uint32_t *n);
/* MoveGeneratorBitboard::countMoves (:3010) of n boards on the host, with
 * one slider variant (the same source the kernels compile). */
int pg_host_count_moves(const pg_pos48 *pos, uint32_t n, int slider, uint32_t *out);
/* MoveGeneratorBitboard::generateBoards (:1302-1738) of n boards on the host:
 * the children of board i follow those of board i-1 in out (at most cap of
 * them in all; PG_ENOMEM beyond), counts[i] = their number. */
int pg_host_children(const pg_pos48 *pos, uint32_t n, pg_pos48 *out, uint64_t cap, uint32_t *counts);
/* perft_bb (perft_bb.h:79-101): the reference's CPU perft -- one host
 * thread, generateBoards-style recursion with the legal-move count at depth
 * 1, fixed-shift fancy magics as the reference's CPU build uses.  A separate
 * entry point for callers that want the CPU path; no device entry point
 * ever calls it. */
int pg_cpu_perft(const pg_pos48 *pos, uint32_t depth, uint64_t *out);
/* MoveGeneratorBitboard::makeMove<chance, updateHash> (:3575); key may be NULL. */
int pg_host_make_move(pg_pos48 *pos, pg_move move, pg_key128 *key);
/* Slider lookups of one variant on the host tables (the device copies are
 * probed by pg_device_slider_attacks). */
int pg_host_slider_attacks(int slider, int rook, const uint8_t *squares, const uint64_t *occupancy,
                           uint32_t n, uint64_t *out);
/* Seeded workload generator (no reference counterpart): start position plus
 * a uniform-random number of plies in [min_ply, max_ply], each ply a uniform
 * pick among the legal moves sorted by (from, to, promotion); games that end
 * early or final positions without a legal move are redrawn. */
int pg_random_positions(uint64_t seed, uint32_t n, uint32_t min_ply, uint32_t max_ply,
                        pg_pos48 *out, uint32_t *plies);

/* ---------------------------------------------------------- device side */

typedef struct pg_ctx pg_ctx;

/* initGPU (launcher.h:1506) + MoveGeneratorBitboard::init (:982):
 * selects the device, builds and uploads tables, reserves the arena. */
int pg_init(int device, const pg_config *cfg, pg_ctx **ctx);
/* Releases the context, and with it every transposition table and slot
 * server created on it that is still alive (their handles become invalid).
 * Serving loops (pg_serve_create) must be destroyed before, since each holds
 * a slot server: the order is pg_serve_destroy, pg_server_destroy,
 * pg_tt_destroy, pg_shutdown. */
void pg_shutdown(pg_ctx *ctx);
int pg_get_config(pg_ctx *ctx, pg_config *out);
int pg_set_slider(pg_ctx *ctx, int slider);

/* perftLauncher, no-TT branch (launcher.h:1455-1503) driving
 * perft_bb_driver_gpu / perft_bb_gpu_simple (perft_bb.h:1027-1220):
 * breadth-first perft of one position; returns the leaf count. */
int pg_perft(pg_ctx *ctx, const pg_pos48 *pos, uint32_t depth, int slider, uint64_t *out);
/* No counterpart in the reference.  pg_perft for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): the same breadth-first search and
 * kernels, with the tree arena and the call's totals and flags in a static
 * device array of 64 MiB per GPU (held until the process ends) instead of the
 * context's arena.  Calls no hipMalloc, hipFree or other allocating routine;
 * levels split to fit the 64 MiB, and a board whose children alone do not fit
 * returns PG_ENOMEM with an error text.  All calls on one GPU share that
 * array: calls that overlap on one GPU race (use one caller thread per GPU).
 * Depth at most 15. */
int pg_perft_noalloc(pg_ctx *ctx, const pg_pos48 *pos, uint32_t depth, int slider, uint64_t *out);
int pg_last_stats(pg_ctx *ctx, pg_stats *out);
/* Test hook (no reference counterpart; the reference sums deep totals in
 * InfInt, launcher.h:1181): every later perft call of this context starts its
 * root total at `bias` instead of 0, until set back to 0.  A total that then
 * passes 2^64 must make the call return PG_EOVERFLOW, never a wrapped count;
 * biased totals are never stored in a transposition table. */
int pg_test_count_bias(pg_ctx *ctx, uint64_t bias);
/* Text of the last error on the calling thread (empty if none). */
const char *pg_last_error(void);

/* Device move-generator probes (one thread per board); the reference runs
 * the same functions inside its kernels (countMoves :3010, generateMoves
 * :2471, makeMove :3575, slider lookups :731-838).  moves: n*256 slots. */
int pg_device_count_moves(pg_ctx *ctx, const pg_pos48 *pos, uint32_t n, int slider, uint32_t *out);
int pg_device_generate_moves(pg_ctx *ctx, const pg_pos48 *pos, uint32_t n, int slider,
                             pg_move *moves, uint32_t *counts);
int pg_device_make_move(pg_ctx *ctx, const pg_pos48 *pos, const pg_move *moves, uint32_t n,
                        pg_pos48 *out, pg_key128 *keys /* in: parent keys, out: child keys; may be NULL */);
int pg_device_slider_attacks(pg_ctx *ctx, int slider, int rook, const uint8_t *squares,
                             const uint64_t *occupancy, uint32_t n, uint64_t *out);
/* The level scan (ScanD, scan.cuh:329-375, the reference's exclusive prefix
 * sum between levels, perft_bb.h:1092) on n u32 counts: offsets[i] = sum of
 * counts[0..i), offsets[n] = the total; the counts are placed misalign (0-3)
 * elements past a 16-byte boundary on the device. */
int pg_device_scan(pg_ctx *ctx, const uint32_t *counts, uint64_t n, uint32_t misalign, uint64_t *offsets);

/* ------------------------------------------------- transposition table */

typedef struct pg_tt pg_tt;

/* setupHashTables128b / allocAndClearMem (launcher.h:155-320): one HBM
 * allocation of at most `bytes`, split into one table per remaining depth
 * (16-byte entries for depths 3-4, 32-byte entries above; none for depth 2),
 * zeroed. */
int pg_tt_create(pg_ctx *ctx, uint64_t bytes, pg_tt **tt);
int pg_tt_clear(pg_tt *tt);
/* Checkpoint and resume of a table's contents (the reference keeps its
 * transpositions across runs in the disk hash perfthash.dat, launcher.h:700-789
 * and :1711): pg_tt_save writes a header with the table layout and a checksum,
 * then the entries; pg_tt_load accepts only a file written from a table of the
 * same size and layout, and clears the table if the data do not match. */
int pg_tt_save(pg_tt *tt, const char *path);
int pg_tt_load(pg_tt *tt, const char *path);
/* freeHashTables (launcher.h:322).  PG_EINVAL, and the table is kept, while
 * a slot server (pg_server_create, pg_serve_create) still shares it. */
int pg_tt_destroy(pg_tt *tt);
/* allocated bytes and log2(slots) per remaining depth (-1: no table) */
int pg_tt_info(pg_tt *tt, uint64_t *bytes, int32_t *bits_per_depth /* [16] */);
/* PRINT_HASH_STATS counters (perft_bb.h:24-32), per remaining depth:
 * probes, hits, stores, in-level duplicates -> out[16][4] */
int pg_tt_counters(pg_tt *tt, uint64_t *out);
/* perftLauncher, TT branch, driving perft_bb_gpu_simple_hash
 * (launcher.h:1457-1466, perft_bb.h:1242-1589) for one position. */
int pg_perft_tt(pg_ctx *ctx, pg_tt *tt, const pg_pos48 *pos, uint32_t depth, int slider, uint64_t *out);
/* perft_bb_cpu_launcher + dividedPerft (launcher.h:1181-1389, :1392-1450):
 * perft of any depth, with a 128-bit total out[0] (low) / out[1] (high) --
 * the reference keeps these in InfInt.  Nodes deeper than launch_depth are
 * walked on the host (moves ordered quiet first, as sortMoves,
 * launcher.h:367) and memoised in a host table keyed by (128-bit key,
 * depth), the reference's complete TT (:598-697); each node at launch_depth
 * is one device call (pg_perft_tt if tt is given, else pg_perft).  For every
 * move of a host node deeper than divided_depth, cb (if not NULL) receives
 * the move path ("..e2-e4 e7*d6 ", Utils::getCompactMoveString) and that
 * child's total, in the order the reference prints them (launcher.h:1346). */
typedef void (*pg_divide_cb)(void *user, const char *path, uint64_t lo, uint64_t hi);
typedef struct pg_deep_stats {
    double total_ms;        /* wall time of the call */
    double device_ms;       /* time inside device perft calls (the reference's gpuTime) */
    uint64_t device_calls;  /* device perft calls made */
    uint64_t host_nodes;    /* host nodes visited */
    uint64_t host_tt_hits;  /* host nodes answered by the host table */
    uint64_t disk_hits;     /* nodes answered by the disk table (pg_perft_deep_disk) */
    uint64_t disk_stores;   /* nodes written to it */
} pg_deep_stats;
int pg_perft_deep(pg_ctx *ctx, pg_tt *tt, const pg_pos48 *pos, uint32_t depth, uint32_t launch_depth,
                  uint32_t divided_depth, int slider, pg_divide_cb cb, void *user, uint64_t out[2],
                  pg_deep_stats *stats);
/* checkAndCreateDiskHash, diskTTProbe, diskTTStore (launcher.h:1698-1722,
 * :700-789): pg_perft_deep with a disk table.  hash_path is created with
 * 2^12 blank 32-byte DiskHashEntry slots (chess.h:518-525) when it does not
 * exist; nodes of remaining depth disk_depth (0: depth - 2, DISK_HASH_LEVEL
 * launcher.h:40) are looked up there first, and each one computed is written
 * and synced (totals below 2^64 only), so a run stopped part way resumes
 * from the nodes it finished.  One process at a time per file. */
int pg_perft_deep_disk(pg_ctx *ctx, pg_tt *tt, const pg_pos48 *pos, uint32_t depth, uint32_t launch_depth,
                       uint32_t divided_depth, int slider, pg_divide_cb cb, void *user, const char *hash_path,
                       uint32_t disk_depth, uint64_t out[2], pg_deep_stats *stats);
/* processPerftRecords (launcher.h:1578-1695): perft(depth) of every FEN line
 * of in_path, whose last number is an occurrence count; appends
 * "<line> <perft> <perft * count>" per line to <in_path>.op.  Lines the .op
 * file already answers are skipped (resume, :1627-1631); the rest go to the
 * device `chunk` lines per batched launch (pg_perft_batch, tt optional),
 * each chunk flushed before the next. */
int pg_perft_records(pg_ctx *ctx, pg_tt *tt, const char *in_path, uint32_t depth, int slider, uint32_t chunk,
                     uint32_t *done, uint32_t *skipped);
/* Test hook (no reference counterpart): overwrite every entry with seeded
 * pseudo-random words; lookups must then miss, never return garbage. */
int pg_tt_poison(pg_tt *tt, uint64_t seed);
/* FIND_DUPLICATES_IN_BFS (switches.h:52): 1 (default) collapses identical
 * positions within a level before expanding them. */
int pg_set_dedup(pg_ctx *ctx, int on);

/* ------------------------------------------------------------- batches */

typedef struct pg_batch_stats {
    double total_ms;        /* host wall time of the call */
    double graph_ms;        /* HIP-event time of the graph launches */
    uint64_t launches;      /* graph launches (1 unless the batch had to split) */
    uint64_t splits;        /* halvings forced by the region budget */
    uint64_t leaf_items;    /* (parent, move) items of the leaf level */
    uint64_t graphs_cached; /* instantiated graphs held by the context */
    uint64_t peak_bytes;    /* most bytes of the region the search used at once */
} pg_batch_stats;

/* Batched perft (no reference counterpart beyond launcher.h:966-1179's
 * sibling batches): perft(depth) of n roots, mixed side to move allowed, in
 * one breadth-first search over all roots, launched as a single HIP graph
 * (captured on first use per shape, then replayed).  dedup: collapse
 * identical positions within a level; blocks_per_cu: persistent grid size
 * (0 = context default); tt: optional shared table.  out[i] = perft of
 * pos[i]. */
int pg_perft_batch(pg_ctx *ctx, const pg_pos48 *pos, uint32_t n, uint32_t depth, int slider, int dedup,
                   int blocks_per_cu, pg_tt *tt, uint64_t *out, pg_batch_stats *stats);
/* Batched perft with a depth per root (launcher.h:966-1179's sibling
 * batches, as pg_perft_batch): one pg_perft_batch per depth present,
 * shallowest first; out[i] = perft of pos[i] to depths[i].  stats sums the
 * calls (peak_bytes: their maximum, total_ms: the whole call). */
int pg_perft_batch_mixed(pg_ctx *ctx, const pg_pos48 *pos, const uint32_t *depths, uint32_t n, int slider, int dedup,
                         int blocks_per_cu, pg_tt *tt, uint64_t *out, pg_batch_stats *stats);
/* Work-groups per kernel of later pg_perft_batch calls (the threads that
 * share a batch's positions): -1 (default) sized to the batch -- its largest
 * level holds at most about n * 48^(depth-2) boards, 256 per work-group,
 * rounded up to a power of two --, 0 the whole chip (blocks_per_cu per CU),
 * > 0 at most that many.  Every kernel is a grid-stride loop, so the counts
 * never depend on it. */
int pg_set_batch_grid(pg_ctx *ctx, int max_groups);

/* ------------------------------------------------------------- serving */

/* No reference counterpart.  A server holds `slots` batch slots, each with
 * its own stream, graph cache and HBM region of bytes_per_slot (carved from
 * free HBM at creation), sharing the optional table tt.  submit() stages up
 * to max_roots roots of one depth on an idle slot and launches the batch
 * asynchronously; poll() returns 1 and fills out[] (submission order) when
 * it is done, 0 while it runs.  Batches on different slots overlap. */
typedef struct pg_server pg_server;
int pg_server_create(pg_ctx *ctx, uint32_t slots, uint64_t bytes_per_slot, uint32_t max_roots, pg_tt *tt,
                     pg_server **out);
void pg_server_destroy(pg_server *srv);
int pg_server_submit(pg_server *srv, uint32_t slot, const pg_pos48 *pos, uint32_t n, uint32_t depth, int slider,
                     int dedup);
int pg_server_poll(pg_server *srv, uint32_t slot, uint64_t *out, pg_batch_stats *stats);
/* pg_server_submit whose kernels keep at most wg_per_cu work-groups per CU
 * resident (four fit per CU; 0 = no limit), leaving CUs to the other slots. */
int pg_server_submit_capped(pg_server *srv, uint32_t slot, const pg_pos48 *pos, uint32_t n, uint32_t depth,
                            int slider, int dedup, uint32_t wg_per_cu);
int pg_server_slots(pg_server *srv, uint32_t *slots, uint64_t *bytes_per_slot);
/* Capture ahead of time, on every idle slot, the graphs that batches of
 * `depth` with 1..max_roots roots and this slider / dedup / wg_per_cu will
 * replay (a graph is captured on a slot's first batch of each shape, which
 * holds up the caller for a millisecond or two).  *graphs (may be NULL)
 * gets the number captured. */
int pg_server_prepare(pg_server *srv, uint32_t depth, uint32_t max_roots, int slider, int dedup, uint32_t wg_per_cu,
                      uint32_t *graphs);
/* hipMemGetInfo of the context's device (admission by free HBM). */
int pg_device_memory(pg_ctx *ctx, uint64_t *free_bytes, uint64_t *total_bytes);

/* The serving loop (no reference counterpart; the reference runs one perft
 * at a time, perftLauncher launcher.h:1455).  Queries (position, depth)
 * arrive with pg_serve_submit; pg_serve_poll, called by the caller's thread,
 * finishes completed batches, admits waiting queries into idle slots and
 * returns finished queries with their latency (submit to completion seen).
 * Admission: waiting queries are grouped by depth; the depth whose oldest
 * query arrived first goes next, as many of its queries as max_batch and the
 * slot's HBM region allow (region bytes per query of each depth are learned
 * from the peak use of that depth's finished batches); depths >= long_depth
 * hold at most long_slots slots, and their kernels at most long_wg_per_cu
 * work-groups per CU (light_wg_per_cu when few queries wait), so short
 * queries keep finding free slots and CUs.
 * Zero fields take the defaults in brackets. */
typedef struct pg_serve_config {
    uint32_t slots;           /* concurrent batches, one stream each [GPU_MAX_HW_QUEUES - 2, 1..4]; each
                                 slot's stream needs a hardware queue of its own (the context's and the
                                 runtime's null stream take two), or its batches wait behind another's */
    uint64_t bytes_per_slot;  /* HBM region per slot [60% of free HBM / slots, at most 24 GiB] */
    uint32_t max_batch;       /* most queries per batch [1024] */
    uint32_t long_depth;      /* queries of this depth or deeper are long [6] */
    uint32_t long_slots;      /* most slots long batches hold at once [slots / 2] */
    uint32_t long_wg_per_cu;  /* work-groups per CU a long batch's kernels keep resident [2] */
    int slider;               /* slider variant [0 = Kogge-Stone; PG_SLIDER_AUTO = the context default] */
    int dedup;                /* collapse identical positions within a level */
    uint32_t light_wg_per_cu; /* ... while at most light_queue queries wait and short queries are about
                                 (waiting, running or submitted in the last 10 ms): a light load, where
                                 their latency matters more than long ones' throughput [1] */
    uint32_t light_queue;     /* [16] */
} pg_serve_config;
typedef struct pg_serve_result {
    uint64_t id;        /* from pg_serve_submit */
    uint64_t tag;       /* the caller's tag */
    uint64_t perft;     /* the perft count (when status is PG_OK) */
    uint32_t depth;
    int status;         /* PG_OK or the error of the query's batch */
    double latency_ms;  /* submit to completion seen by pg_serve_poll */
    double queued_ms;   /* of which waiting before its batch was launched */
} pg_serve_result;
typedef struct pg_serve_stats {
    uint64_t submitted, completed, batches, launches, splits, leaves;
    uint32_t waiting, busy_slots, max_busy_slots, max_long_slots;
    double p50_ms, p99_ms, mean_ms, max_ms;  /* latency of the completed queries */
    double elapsed_ms;                       /* first submit to last completion */
    double est_bytes_per_query[16];          /* current admission estimate per depth */
    double submit_ms_max, submit_ms_total;   /* host time spent launching batches */
    uint64_t graphs;                         /* graphs instantiated by the slots so far */
    double device_ms;                        /* device time of the finished batches, summed (HIP
                                                events around each batch's graph): over elapsed_ms,
                                                the mean number of batches running at once */
} pg_serve_stats;
typedef struct pg_serve pg_serve;
int pg_serve_create(pg_ctx *ctx, pg_tt *tt /* may be NULL */, const pg_serve_config *cfg, pg_serve **out);
void pg_serve_destroy(pg_serve *srv);
int pg_serve_submit(pg_serve *srv, const pg_pos48 *pos, uint32_t depth, uint64_t tag, uint64_t *id);
int pg_serve_poll(pg_serve *srv, pg_serve_result *out, uint32_t max, uint32_t *n);
int pg_serve_get_stats(pg_serve *srv, pg_serve_stats *out);
/* The admission decision alone (host only, no device): given waiting counts
 * and oldest arrival time per depth, long batches running, idle slots, the
 * region size and the per-depth estimates, returns 1 with the depth and the
 * number of queries of the next batch, or 0 if none may start. */
int pg_serve_policy(const uint32_t waiting[16], const double oldest[16], uint32_t busy_long, uint32_t idle_slots,
                    const pg_serve_config *cfg, uint64_t region_bytes, const double est_bytes[16], uint32_t *depth,
                    uint32_t *count);
/* The resident work-group cap of the next batch (host only, no device; 0 =
 * uncapped): 0 below long_depth; light_wg_per_cu while at most light_queue
 * queries wait and short queries are about (shorts_about != 0); else
 * long_wg_per_cu.  cfg's fields are used as given (no defaults filled). */
uint32_t pg_serve_wg_cap(const pg_serve_config *cfg, uint32_t depth, uint32_t queued, int shorts_about);

/* ---------------------------------------------------------- diagnostics */
/* Measurement reference points (no reference counterpart; the reference's
 * only timing is the wall clock around a perft, launcher.h:1486-1498).
 *
 * pg_device_copy: plain device-to-device copy of `bytes` (16-byte lanes),
 * `iters` times; *ms gets the mean launch time.  Its FETCH_SIZE / WRITE_SIZE
 * counters calibrate those of the perft kernels, its time the achievable
 * copy bandwidth.
 *
 * pg_device_issue_probe: one launch of an unrolled stream of one VALU
 * instruction class -- kind 0 v_bitop3_b32, 1 v_xor_b32 (two-operand
 * encoding), 2 v_lshlrev_b64, 3 v_bcnt_u32_b32, 4 v_and_b32, 5 v_or3_b32,
 * 6 v_lshlrev_b32, 7 v_alignbit_b32, 8 v_add_u32, 9 v_lshl_add_u64,
 * 10 v_bfrev_b32, 11 v_ffbl_b32, 12 v_mov_b32, 13 v_lshl_or_b32,
 * 14 v_and_or_b32, 15 v_not_b32 -- with eight waves per SIMD;
 * returns its time, the wave instructions it issued, the engine clock (kHz)
 * and the CU count, from which the issue rate per CU per clock follows. */
int pg_device_copy(pg_ctx *ctx, uint64_t bytes, uint32_t iters, double *ms);
int pg_device_issue_probe(pg_ctx *ctx, int kind, uint32_t iters, double *ms, uint64_t *wave_insts,
                          uint32_t *clock_khz, uint32_t *cus);

#ifdef __cplusplus
}
#endif
#endif
