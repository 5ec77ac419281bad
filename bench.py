#!/usr/bin/env python3
"""Perft throughput on one MI355X.

Workload (one step): plain breadth-first GPU perft without transposition
table of the start position to depth 7 and depth 8, plus 16 seeded random
positions (start position + 10..60 uniform-random legal plies) to depth 6 --
107.2e9 leaves per step, every count checked against the reference's values.

Prints one JSON line: leaves/s over the timed steps, per-variant timings,
the leaf kernel's roofline (integer VALU issue against the chip's vector
peak, HBM bytes beside it) and the CPU baseline (the reference's own CPU
perft compiled from its sources, timed on this host on a bounded sample).
Legs beside the headline (--legs, comma list):
  counters   two rocprofv3 --pmc passes over one step of this workload plus
             the calib leg in a child process: VALU instructions per counted
             board, HBM bytes per launch (scaled by the copy kernel's
             counters), clock cycles -- measured by this run, not pasted
  calib      1 GiB copy bandwidth and the VALU issue-rate probes
  batch      4096 seeded random positions at depth 5 in one graph launch
             (autotuned), against the one-position-per-launch loop
  tt         transposition-table perft of the start position at depth 9 and
             10 with 16 GB and 64 GB tables, with per-depth table counters
  serve      served queries (distinct seeded positions, depths 3-7 drawn
             40/30/20/8/2%) at Poisson rates 100..10000/s and a burst: leaves/s
             and p50/p99 latency, with and without a shared table
  leafmodes  start perft(7)/(8) per slider variant, fused vs materialized leaf
"""
import argparse
import csv
import glob
import json
import os
import platform
import shutil
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "perft leaves/sec on 1 MI355X (exact counts); p50/p99 latency of served queries"
HBM_PEAK_GBS = 8000.0
# Used only when the counters leg cannot run (no rocprofv3): VALU
# instructions per counted board (SQ_INSTS_VALU x 64 / boards, one lane per
# board) of k_leaf_moves<kogge_stone> measured by rocprofv3 on this workload
# (profiles/k_leaf_moves_kogge_start_d7_d8_c2_d6.md)
PROFILE_VALU_INSTS_PER_BOARD = 652
PROFILE_SOURCE = "profiles/k_leaf_moves_kogge_start_d7_d8_c2_d6.md"
LEAF_KERNEL = "k_leaf_moves<0>"
COPY_BYTES = 1 << 30
COPY_ITERS = 3


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def load_workload(pg):
    from perft_gpu_amd import workloads as W
    with open(os.path.join(ROOT, "tests", "golden", "c2.json")) as f:
        c2 = json.load(f)
    start = pg.read_fen_string(pg.START_FEN)
    jobs = [("start_d7", start, 7, W.CPW_COUNTS["start"][7]),
            ("start_d8", start, 8, W.CPW_COUNTS["start"][8])]
    for i, item in enumerate(c2["positions"]):
        jobs.append(("c2_%02d_d6" % i, pg.pos_from_words(item["words"]), 6, item["perft"][6]))
    return jobs


def run_step(dev, jobs, variant, leaf_acc=None):
    leaves = 0
    for name, pos, depth, want in jobs:
        got = dev.perft(pos, depth, variant)
        if got != want:
            raise SystemExit("MISMATCH %s: gpu %d expected %d" % (name, got, want))
        leaves += got
        if leaf_acc is not None:
            st = dev.last_stats()
            leaf_acc["launches"] += st["leaf_launches"]
            leaf_acc["ms"] += st["leaf_kernel_ms"]
            leaf_acc["items"] += st["leaf_items"]
            leaf_acc["parents"] += st["leaf_parents"]
            # the leaf makes its parents from the boards one ply up (the root at depth 3)
            leaf_acc["grandparents"] += st["nodes"][depth - 3] if depth > 3 else 1
    return leaves


def golden(name):
    with open(os.path.join(ROOT, "tests", "golden", name)) as f:
        return json.load(f)


def leg_batch(pg, dev):
    import numpy as np
    from perft_gpu_amd import autotune
    c3 = golden("c3.json")
    arr = np.array(c3["words"], np.uint64)
    want = np.array(c3["perft"], np.uint64)
    total = int(want.sum())
    cfg, table = autotune.tune(dev, arr, 5, sample=512, log=log)
    runs = []
    for _ in range(3):
        t = time.perf_counter()
        got = autotune.run(dev, arr, 5, cfg)
        runs.append(time.perf_counter() - t)
        if not np.array_equal(got, want):
            raise SystemExit("MISMATCH in the 4096-position batch")
    dt = min(runs)
    got, st = dev.perft_batch(arr, 5, cfg["slider"], cfg["dedup"], cfg["blocks_per_cu"])   # stats of one launch
    if not np.array_equal(got, want):
        raise SystemExit("MISMATCH in the 4096-position batch")
    t = time.perf_counter()
    s = 0
    for w in arr:
        s += dev.perft(pg.pos_from_words(w), 5, cfg["slider"])
    loop = time.perf_counter() - t
    if s != total:
        raise SystemExit("MISMATCH in the one-position-per-launch loop")
    return {"workload": "4096 random positions (ply 5-80) at depth 5, one graph launch", "leaves": total,
            "ms": dt * 1e3, "graph_ms": st["graph_ms"], "leaves_per_s": total / dt, "launches": st["launches"],
            "autotuned": {"slider": pg.SLIDER_NAMES[cfg["slider"]], "dedup": cfg["dedup"],
                          "blocks_per_cu": cfg["blocks_per_cu"], "split_depth": cfg["split"],
                          "grid": ("%d threads per position" % cfg["tpp"]) if cfg.get("tpp")
                          else "sized to the batch" if cfg["grid"] < 0 else "whole chip"},
            "autotune_table": [{"slider": pg.SLIDER_NAMES[c["slider"]], "dedup": c["dedup"],
                                "blocks_per_cu": c["blocks_per_cu"], "split_depth": c["split"],
                                "grid": "tpp" if c.get("tpp") else "batch" if c["grid"] < 0 else "chip",
                                "threads_per_position": tpp,
                                "ms_on_512_sample": t_ * 1e3} for c, t_, tpp in table],
            "one_position_per_launch": {"ms": loop * 1e3, "leaves_per_s": total / loop},
            "speedup_vs_plain_loop": loop / dt, "exact": True}


def leg_tt(pg, sizes=(16, 64), depths=(9, 10)):
    """Table searches on a context of their own, whose arena is what the
    largest table leaves free (pg.table_search_arena): fewer level splits,
    so more equal boards are collapsed before the leaf."""
    from perft_gpu_amd import workloads as W
    out = []
    arena = pg.table_search_arena(max(sizes) << 30) or pg.DEFAULT_ARENA
    with pg.Device(0, arena_bytes=arena) as dev:
        for gb in sizes:
            out.extend(tt_runs(pg, W, dev, gb, depths))
    return {"workload": "start position, TT perft(%s) with %s GB tables, each table cleared at creation and "
                        "filled by the depths before" % ("/".join(map(str, depths)), "/".join(map(str, sizes))),
            "arena_gib": arena / (1 << 30), "runs": out, "exact": True}


TT_SHALLOW_MAX = 4          # remaining depths with 16-byte entries; deeper ones take 32 bytes
LINE = 64                   # bytes a random access moves when rounded to whole lines


def tt_access_bytes(nodes, ctr):
    """The table's own HBM traffic in one TT search, algorithmic and rounded
    to one 64-byte line per random access.  nodes[r]: boards made at
    remaining depth r; ctr[r] = (probes, hits, stores, duplicates) of the
    search.  Per probe and per store one entry (16 B at r <= 4, 32 B deeper);
    per board of a level with a duplicate set (r >= 2) that the table did not
    answer, one 4-byte set slot read and written; per duplicate its owner's
    16-byte key (compared) and 8-byte perft (copied).  The streams every
    level has anyway (boards, keys, counts, per-board perfts) are not
    included."""
    alg = rnd = 0
    acc = {"probes": 0, "stores": 0, "set_inserts": 0, "duplicates": 0}
    for r, (probes, hits, stores, dups) in ctr.items():
        e = 16 if r <= TT_SHALLOW_MAX else 32
        inserts = max(0, nodes.get(r, 0) - hits) if r >= 2 else 0
        alg += e * (probes + stores) + 8 * inserts + 24 * dups
        rnd += LINE * (probes + stores) + 2 * LINE * inserts + 2 * LINE * dups
        acc["probes"] += probes
        acc["stores"] += stores
        acc["set_inserts"] += inserts
        acc["duplicates"] += dups
    acc.update({"algorithmic_gb": alg / 1e9, "line_rounded_gb": rnd / 1e9})
    return acc


def tt_runs(pg, W, dev, gb, depths):
    """TT perft of the start position at each depth on one new table of gb GB."""
    import numpy as np
    out = []
    with pg.TranspositionTable(dev, gb << 30) as tt:
        for d in depths:
            c0 = tt.counters().astype(np.int64)
            t = time.perf_counter()
            n = tt.perft(pg.START_FEN, d, 0)
            dt = time.perf_counter() - t
            if n != W.CPW_COUNTS["start"][d]:
                raise SystemExit("MISMATCH TT perft(%d)" % d)
            st = dev.last_stats()
            c = tt.counters().astype(np.int64) - c0
            nodes = {d - p: st["nodes"][p] for p in range(1, d)}
            ctr = {r: tuple(int(x) for x in c[r]) for r in range(2, d + 1)}
            out.append({"tt_gb": gb, "depth": d, "perft": n, "s": dt, "effective_leaves_per_s": n / dt,
                        "peak_arena_gb": st["peak_arena"] / 1e9, "boards_per_ply": st["nodes"][1:d],
                        "leaf_items": st["leaf_items"], "splits": st["splits"],
                        "table_access": tt_access_bytes(nodes, ctr)})
            log("tt %d GB d%d %.3f s" % (gb, d, dt))
        c = tt.counters()
        out[-1]["counters_by_remaining_depth"] = {
            int(r): {"probes": int(c[r][0]), "hits": int(c[r][1]), "stores": int(c[r][2]),
                     "in_level_duplicates": int(c[r][3])} for r in range(2, max(depths) + 1) if c[r].any()}
    return out


def leg_calib(pg, dev):
    """Copy bandwidth of a 1 GiB copy and the sustained issue rate of four
    VALU instruction classes (wave instructions per CU per clock)."""
    ms, gbs = dev.copy_bandwidth(COPY_BYTES, COPY_ITERS)
    probes = {}
    for kind, name in dev.ISSUE_KINDS.items():
        rate, pms, n, clk, cus = dev.issue_rate(kind, 16384)
        probes[name] = {"wave_insts_per_cu_per_clock": rate, "ms": pms, "wave_insts": n}
        log("issue probe %-24s %.3f wave instructions per CU per clock" % (name, rate))
    log("copy %d MiB: %.3f ms, %.0f GB/s" % (COPY_BYTES >> 20, ms, gbs))
    return {"copy": {"bytes": COPY_BYTES, "iters": COPY_ITERS, "ms_per_launch": ms, "gb_per_s_read_plus_write": gbs},
            "issue_probes": probes, "clock_mhz": clk / 1e3, "cus": cus}


def leg_leafmodes(pg, dev):
    """Start perft(7)/(8) per slider variant, fused last two plies (leaf_mode
    0) against the materialized (parent, move) item list (leaf_mode 1)."""
    from perft_gpu_amd import workloads as W
    rows = []
    for mode in (0, 1):
        with pg.Device(0, arena_bytes=48 << 30, timing=True, leaf_mode=mode) as d:
            for v in range(4):
                for depth in (7, 8):
                    d.perft(pg.START_FEN, depth, v)   # warm
                    t = time.perf_counter()
                    n = d.perft(pg.START_FEN, depth, v)
                    dt = time.perf_counter() - t
                    if n != W.CPW_COUNTS["start"][depth]:
                        raise SystemExit("MISMATCH leaf mode %d variant %d d%d" % (mode, v, depth))
                    st = d.last_stats()
                    rows.append({"leaf_mode": mode, "slider": pg.SLIDER_NAMES[v], "depth": depth, "s": dt,
                                 "leaves_per_s": n / dt, "leaf_kernel_ms": st["leaf_kernel_ms"],
                                 "leaf_items": st["leaf_items"]})
                    log("leaf_mode %d %-14s d%d %.4f s %7.1f G leaves/s leaf kernel %.2f ms"
                        % (mode, pg.SLIDER_NAMES[v], depth, dt, n / dt / 1e9, st["leaf_kernel_ms"]))
    return {"workload": "start position perft(7) and perft(8), each timed after one warm run", "rows": rows,
            "exact": True}


# ---------------------------------------------------------------- counters
def _read_counters(outdir):
    """rocprofv3 counter CSV -> {kernel: {counter: sum}, '#dispatches': {kernel: n}}"""
    sums, disp = {}, {}
    for path in glob.glob(os.path.join(outdir, "**", "*counter_collection.csv"), recursive=True):
        with open(path, newline="") as f:
            for row in csv.DictReader(f):
                k = row.get("Kernel_Name", "")
                c = row.get("Counter_Name", "")
                v = float(row.get("Counter_Value", "0") or 0)
                sums.setdefault(k, {})
                sums[k][c] = sums[k].get(c, 0.0) + v
                disp.setdefault(k, set()).add(row.get("Dispatch_Id", row.get("Correlation_Id", "")))
    return sums, {k: len(v) for k, v in disp.items()}


def _find(sums, needle):
    hits = [k for k in sums if needle in k]
    return hits[0] if hits else None


def leg_counters(pg, dev=None):
    """Hardware counters of this workload, collected by this run: two
    rocprofv3 --pmc passes over a child process that runs one bench step
    (Kogge-Stone) and the calib leg.  Instructions per counted board are
    deterministic, so the parent's HIP-event timing turns them into a rate."""
    exe = shutil.which("rocprofv3")
    if exe is None:
        return {"measured": False, "reason": "rocprofv3 not found"}
    passes = [["SQ_INSTS_VALU", "SQ_WAVES", "GRBM_GUI_ACTIVE", "FETCH_SIZE", "SQ_ACTIVE_INST_VALU2",
               "SQ_THREAD_CYCLES_VALU", "SQ_ACTIVE_INST_VALU"],
              ["WRITE_SIZE", "SQ_INSTS_LDS", "SQ_LDS_BANK_CONFLICT", "SQ_INSTS_SALU"]]
    child = [sys.executable, os.path.join(ROOT, "bench.py"), "--steps", "1", "--warmup", "0", "--no-cpu",
             "--legs", "calib", "--variant", "0"]
    work = tempfile.mkdtemp(prefix="pg_counters_", dir="/tmp")
    sums, disp, child_line = {}, {}, None
    env = dict(os.environ, TMPDIR="/tmp")
    try:
        for i, ctrs in enumerate(passes):
            out = os.path.join(work, "pass%d" % i)
            cmd = ["timeout", "-s", "KILL", "240", exe, "--pmc"] + ctrs + ["-d", out, "-o", "pmc", "--output-format",
                                                                          "csv", "--"] + child
            log("counters pass %d: %s" % (i + 1, " ".join(ctrs)))
            r = subprocess.run(cmd, capture_output=True, text=True, cwd="/tmp", env=env)
            lines = [x for x in r.stdout.splitlines() if x.startswith("{")]
            if lines:
                child_line = json.loads(lines[-1])
            s_, d_ = _read_counters(out)
            if not s_:
                return {"measured": False, "reason": "rocprofv3 pass %d gave no counters (rc %d): %s"
                        % (i + 1, r.returncode, (r.stderr or "")[-300:])}
            for k, v in s_.items():
                sums.setdefault(k, {}).update(v)
            for k, v in d_.items():
                disp[k] = max(disp.get(k, 0), v)
    finally:
        shutil.rmtree(work, ignore_errors=True)
    leaf, copy = _find(sums, LEAF_KERNEL), _find(sums, "k_copy16")
    if not leaf or not copy or child_line is None:
        return {"measured": False, "reason": "leaf or copy kernel missing from the counter output"}
    L, C = sums[leaf], sums[copy]
    tot = child_line["leaf_totals"]
    nl, nc = disp[leaf], disp[copy]
    # the child's timed pass may launch the leaf more often than the step its
    # totals describe (a warm run first): scale the boards to the dispatches
    boards = tot["items"] * nl / max(1, tot["launches"])
    # FETCH_SIZE / WRITE_SIZE are KiB per dispatch; the copy reads and writes COPY_BYTES per dispatch
    ff = COPY_BYTES * nc / (C["FETCH_SIZE"] * 1024.0)
    wf = COPY_BYTES * nc / (C["WRITE_SIZE"] * 1024.0)
    traffic = (L["FETCH_SIZE"] * ff + L["WRITE_SIZE"] * wf) * 1024.0 / nl
    cus = child_line["calib"]["cus"]
    cycles = L["GRBM_GUI_ACTIVE"] / 8.0    # GRBM_GUI_ACTIVE sums the eight XCDs
    res = {"measured": True, "kernel": leaf, "dispatches": nl,
           "valu_wave_insts": L["SQ_INSTS_VALU"], "boards": boards,
           # each counted board is one lane: wave instructions x 64 / boards is
           # the per-lane program length
           "valu_insts_per_board": L["SQ_INSTS_VALU"] * 64.0 / boards,
           "valu_wave_insts_per_cu_per_cycle": L["SQ_INSTS_VALU"] / (cus * cycles),
           # SQ_ACTIVE_INST_VALU counts quad-cycles per wave: one per wave64
           # instruction, so each SIMD's VALU is busy this fraction of the
           # time (above 1 by the quad-cycles that issued two instructions)
           "valu_busy_per_simd": L.get("SQ_ACTIVE_INST_VALU", 0) * 4.0 / (cus * 4 * cycles),
           "dual_issue_fraction": L.get("SQ_ACTIVE_INST_VALU2", 0) / L["SQ_INSTS_VALU"],
           "lane_utilization": L.get("SQ_THREAD_CYCLES_VALU", 0) / (64.0 * max(1.0, L.get("SQ_ACTIVE_INST_VALU", 0))),
           "lds_insts": L.get("SQ_INSTS_LDS"), "lds_bank_conflict_cycles": L.get("SQ_LDS_BANK_CONFLICT"),
           "salu_insts": L.get("SQ_INSTS_SALU"), "waves": L.get("SQ_WAVES"),
           "hbm_bytes_per_launch": traffic,
           "algorithmic_bytes_per_launch": (6.0 * tot["parents"] + 48.0 * tot["grandparents"]) / tot["launches"],
           "copy_calibration": {"bytes_per_dispatch": COPY_BYTES, "dispatches": nc,
                                "fetch_size_kib_per_dispatch": C["FETCH_SIZE"] / nc,
                                "write_size_kib_per_dispatch": C["WRITE_SIZE"] / nc,
                                "fetch_factor": ff, "write_factor": wf}}
    probes = {}
    for kind, iname in pg.Device.ISSUE_KINDS.items():
        k = _find(sums, "k_issue<%d>" % kind)
        if k and sums[k].get("GRBM_GUI_ACTIVE"):
            probes[iname] = {"valu_wave_insts": sums[k]["SQ_INSTS_VALU"], "dispatches": disp[k],
                             "wave_insts_per_cu_per_cycle": sums[k]["SQ_INSTS_VALU"]
                             / (cus * sums[k]["GRBM_GUI_ACTIVE"] / 8.0),
                             "dual_issue_fraction": sums[k].get("SQ_ACTIVE_INST_VALU2", 0) / sums[k]["SQ_INSTS_VALU"]}
    res["issue_probes"] = probes
    log("counters: %.0f VALU instructions per board, %.3f per CU per cycle (dual issue %.1f%%, lanes active %.1f%%), "
        "%.1f B HBM per parent (algorithmic %.1f)" % (res["valu_insts_per_board"], res["valu_wave_insts_per_cu_per_cycle"],
                                                     100 * res["dual_issue_fraction"], 100 * res["lane_utilization"],
                                                     traffic * tot["launches"] / max(1, tot["parents"]),
                                                     res["algorithmic_bytes_per_launch"] * tot["launches"]
                                                     / max(1, tot["parents"])))
    return res


def leg_serve(pg, dev):
    """Served queries through the library's serving loop (pg_serve_*): Poisson
    streams at five rates and one burst, each over its own slice of a pool of
    distinct seeded queries (no query repeats, so table hits are genuine
    transpositions), with a shared 8 GB table; then the same burst without a
    table, so batching alone is compared with the plain path.  The first 512
    queries are the golden set; every served answer equals the plain
    one-query-per-launch GPU path, which is also the throughput baseline."""
    import numpy as np
    from perft_gpu_amd import workloads as W
    from perft_gpu_amd.serve import ServeLoop, serve, poisson_arrivals, latency_summary
    c5 = golden("c5.json")
    plan = [(100, 200), (300, 600), (1000, 2000), (3000, 2000), (10000, 2000), (None, 2000)]
    pool, depths = W.c5_queries(sum(n for _, n in plan))
    if not np.array_equal(pool[:len(c5["words"])], np.array(c5["words"], np.uint64)):
        raise SystemExit("served-query generator differs from the golden set")
    want = np.zeros(len(pool), np.uint64)
    took = np.zeros(len(pool))
    for i in range(len(pool)):
        t = time.perf_counter()
        want[i] = dev.perft(pg.pos_from_words(pool[i]), int(depths[i]), 0)
        took[i] = time.perf_counter() - t
    plain = float(took.sum())
    if [int(x) for x in want[:len(c5["perft"])]] != c5["perft"]:
        raise SystemExit("MISMATCH in plain served-query loop")
    log("plain one-query-per-launch: %d queries %.2f s" % (len(pool), plain))

    def row(loop, rate, idx):
        arr = np.zeros(len(idx)) if rate is None else poisson_arrivals(len(idx), float(rate), seed=rate)
        r = serve(loop, pool[idx], depths[idx], arr)
        if not np.array_equal(r["results"], want[idx]):
            raise SystemExit("MISMATCH in served results at rate %s" % rate)
        leaves = int(want[idx].sum())
        out = {"rate_per_s": rate if rate else "burst", "queries": len(idx), "elapsed_s": r["elapsed"],
               "leaves_per_s": leaves / r["elapsed"], "batches": r["batches"], "mean_batch": r["mean_batch"],
               "max_busy_slots": r["max_busy_slots"], "max_long_slots": r["max_long_slots"],
               "device_ms": r["device_ms"], "batches_in_flight": r["batches_in_flight"],
               "plain_leaves_per_s": leaves / float(took[idx].sum())}
        out.update(latency_summary(r["latency"]))
        lat, dq, qd = np.asarray(r["latency"]), depths[idx], np.asarray(r["queued"]) * 1e3
        out["by_depth"] = {int(d): dict(latency_summary(lat[dq == d]), queries=int((dq == d).sum()),
                                        queued_p99_ms=float(np.percentile(qd[dq == d], 99)))
                           for d in sorted(set(int(x) for x in dq))}
        out["submit_ms_max"] = r["stats"]["submit_ms_max"]
        log("serve rate %s: p50 %.2f ms p99 %.2f ms %.1f G leaves/s, %.2f batches in flight"
            % (rate, out["p50_ms"], out["p99_ms"], out["leaves_per_s"] / 1e9, out["batches_in_flight"]))
        return out

    rates, first = [], 0
    with pg.TranspositionTable(dev, 8 << 30) as tt, ServeLoop(dev, tt, slots=4, slot_bytes=16 << 30) as loop:
        for rate, n in plan:
            rates.append(row(loop, rate, np.arange(first, first + n)))
            first += n
    burst_idx = np.arange(first - plan[-1][1], first)
    with ServeLoop(dev, None, slots=4, slot_bytes=16 << 30) as loop:
        no_tt = row(loop, None, burst_idx)
    burst = rates[-1]
    return {"workload": "%d distinct seeded queries (ply 0-60; depth 3/4/5/6/7 at 40/30/20/8/2%%), one slice per "
                        "rate; Poisson arrivals at 100-10,000/s, then a burst of 2000 at once; 4 slots, shared 8 GB TT; "
                        "the burst again without a table" % len(pool), "rates": rates,
            "burst_without_table": no_tt,
            "plain_one_query_per_launch": {"queries": len(pool), "s": plain, "leaves_per_s": int(want.sum()) / plain},
            "burst_vs_plain": burst["leaves_per_s"] / burst["plain_leaves_per_s"],
            "burst_without_table_vs_plain": no_tt["leaves_per_s"] / no_tt["plain_leaves_per_s"],
            "exact": True}


def cpu_model():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name"):
                return line.split(":", 1)[1].strip()
    except Exception:
        pass
    return platform.processor()


def cpu_baseline(jobs, threads):
    """The reference's CPU perft_bb (perft_bb.h:79-101) on a bounded sample."""
    import oracle
    lib, kind = oracle.ref(), "reference"
    if lib is None:
        lib, kind = oracle.oracle(), "port"
    start = jobs[0][1].words()
    c2 = [j[1].words() for j in jobs if j[0].startswith("c2_")]
    # single thread: start d6 + the 16 random positions at d5
    t = time.perf_counter()
    n1 = lib.perft(start, 6) + sum(lib.perft(p, 5) for p in c2)
    st_rate = n1 / (time.perf_counter() - t)
    # all threads: start d7 + the 16 random positions at d6
    t = time.perf_counter()
    n2 = lib.perft_mt(start, 7, threads) + sum(lib.perft_mt(p, 6, threads) for p in c2)
    dt = time.perf_counter() - t
    want = jobs[0][3] + sum(j[3] for j in jobs if j[0].startswith("c2_"))
    if n2 != want:
        raise SystemExit("CPU baseline mismatch: %d vs %d" % (n2, want))
    return {"value": n2 / dt, "unit": "leaves/s", "cores": threads, "kind": kind,
            "sample": "start position d7 + the 16 random positions at d6 (%d leaves, %.1f s), "
                      "root split over %d threads" % (n2, dt, threads),
            "single_thread": {"value": st_rate, "sample": "start d6 + 16 random positions at d5 (%d leaves)" % n1},
            "cpu_model": cpu_model(), "nproc": os.cpu_count()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--variant", type=int, default=-2, help="slider variant; -2 = pick the fastest")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--legs", default="counters,calib,batch,tt,serve",
                    help="legs beside the headline (comma list, or none): counters, calib, batch, tt, serve, leafmodes")
    ap.add_argument("--tt-gb", default="16,64", help="table sizes of the tt leg (GB, comma list)")
    ap.add_argument("--tt-depths", default="9,10", help="depths of the tt leg (comma list)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build runs perft on a single MI355X "
              "(--gpus %d requested); multi-GPU sharding is a design note only." % args.gpus)
        sys.exit(2)
    legs_wanted = [x for x in args.legs.split(",") if x and x != "none"]
    # the serving leg runs four slot streams next to the context's and the
    # runtime's null stream: give each its own hardware queue (the HIP
    # runtime reads this once, at its initialization, which is still ahead)
    if int(os.environ.get("GPU_MAX_HW_QUEUES", "4") or 4) < 8:
        os.environ["GPU_MAX_HW_QUEUES"] = "8"

    import perft_gpu_amd as pg
    # the counter passes run in child processes before this one touches the GPU
    counters = leg_counters(pg) if "counters" in legs_wanted else None

    jobs = load_workload(pg)
    # headline steps: each perft is one graph launch (no per-kernel events)
    dev = pg.Device(0, arena_bytes=48 << 30)
    _, _, _, clk_khz, cus = dev.issue_rate(0, 16)

    # per-variant pass (the slider choice for this path); skipped when a
    # variant is forced (profiling runs)
    variants = {}
    for v in (range(4) if args.variant < 0 else []):
        t = time.perf_counter()
        n = run_step(dev, jobs, v)
        dt = time.perf_counter() - t
        variants[pg.SLIDER_NAMES[v]] = {"leaves_per_s": n / dt, "ms_per_step": dt * 1e3}
        log("variant %d: %.1f G leaves/s" % (v, n / dt / 1e9))
    best = args.variant if args.variant >= 0 else max(range(4), key=lambda v: variants[pg.SLIDER_NAMES[v]]["leaves_per_s"])
    log("slider variant", best)

    for _ in range(args.warmup):
        run_step(dev, jobs, best)
    t0 = time.perf_counter()
    leaves = 0
    for _ in range(args.steps):
        leaves += run_step(dev, jobs, best)
    elapsed = max(time.perf_counter() - t0, 1e-9)
    value = leaves / elapsed

    # roofline pass: the same step once more on a context that brackets every
    # leaf-kernel launch with HIP events on its launch stream (the host-driven
    # level loop, which launches the same fused leaf as a kernel of its own)
    acc = {"launches": 0, "ms": 0.0, "items": 0, "parents": 0, "grandparents": 0}
    if args.steps > 0:
        with pg.Device(0, arena_bytes=48 << 30, timing=True) as tdev:
            run_step(tdev, jobs, best)
            run_step(tdev, jobs, best, acc)

    launches = max(1, acc["launches"])
    avg_ms = acc["ms"] / launches
    fused = best != pg.SLIDER_MAGIC_BYTE   # the byte-lookup variant keeps the materialized leaf level
    if fused:
        # the leaf's parents are made in LDS from a move and a parent index
        # (6 B each) and the boards one ply up (48 B each); their moves stay in LDS
        bytes_per_launch = (6 * acc["parents"] + 48 * acc["grandparents"]) / launches
        kname, bnote = "k_leaf_moves", ("6 B per parent (move + index of the board it is made from) + 48 B per board "
                                        "one ply up; the parents and their (parent, move) items stay in LDS")
    else:
        bytes_per_launch = (6 * acc["items"] + 48 * acc["parents"]) / launches
        kname, bnote = "k_leaf", "6 B per (parent, move) item + 48 B per parent board"
    hbm_gbs = bytes_per_launch / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    boards_per_s = acc["items"] / (acc["ms"] * 1e-3) if acc["ms"] > 0 else 0.0
    # VALU roofline: every counted board costs a fixed number of VALU wave
    # instructions (64 lanes each); the vector peak is CUs x 128 lanes x clock,
    # the published 157.3 TFLOP/s fp32 vector rate counted in lane operations
    # (an FMA is two flops).  Instructions per board come from the counters leg
    # of this run; without rocprofv3 the committed profile's figure is used.
    peak_tops = cus * 128 * clk_khz * 1e3 / 1e12
    have_ctr = bool(counters and counters.get("measured")) and best == pg.SLIDER_KOGGE_STONE
    ipb = counters["valu_insts_per_board"] if have_ctr else (
        PROFILE_VALU_INSTS_PER_BOARD if best == pg.SLIDER_KOGGE_STONE else None)
    achieved = boards_per_s * ipb / 1e12 if ipb else None   # lane operations per second
    # (the child ran one step of the same workload: the same launches)
    traffic = counters["hbm_bytes_per_launch"] if have_ctr else None
    roofline = {"bound": "valu", "achieved": achieved, "peak": peak_tops, "unit": "T int32 lane-ops/s",
                "frac": achieved / peak_tops if achieved else None, "traffic": traffic,
                "kernel": "%s<%s>" % (kname, pg.SLIDER_NAMES[best]), "launches": acc["launches"],
                "avg_launch_ms": avg_ms, "boards_per_launch": acc["items"] / launches,
                "valu_insts_per_board": ipb,
                "valu_source": ("rocprofv3 --pmc SQ_INSTS_VALU in this run (counters leg)" if have_ctr
                                else "%s (counters leg not run)" % PROFILE_SOURCE),
                "wave_insts_per_cu_per_clock": achieved * 1e12 / 64 / (cus * clk_khz * 1e3) if achieved else None,
                "peak_note": "%d CUs x 128 lanes x %.0f MHz (= published 157.3 TFLOP/s fp32 vector / 2); the issue "
                             "probes (calib leg) show what each instruction class sustains" % (cus, clk_khz / 1e3),
                "hbm": {"achieved": hbm_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": hbm_gbs / HBM_PEAK_GBS,
                        "algorithmic_bytes_per_launch": bytes_per_launch, "note": bnote,
                        "measured_bytes_per_launch": traffic}}

    line = {
        "metric": METRIC, "value": value, "unit": "leaves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": elapsed / max(1, args.steps) * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u64", "data": "synthetic",
        "config": {"workload": "plain GPU BFS perft, no TT: start position d7 + d8 and 16 seeded random "
                               "positions (ply 10-60) at d6", "leaves_per_step": leaves // max(1, args.steps),
                   "slider": pg.SLIDER_NAMES[best], "exact": True},
        "variants": variants, "roofline": roofline,
        "leaf_totals": {"launches": acc["launches"], "items": acc["items"], "parents": acc["parents"],
                        "grandparents": acc["grandparents"], "ms": acc["ms"]},
    }
    if counters is not None:
        line["counters"] = counters
    for leg in legs_wanted:
        if leg == "counters":
            continue
        if leg == "tt":
            line["tt"] = leg_tt(pg, tuple(int(x) for x in args.tt_gb.split(",")),
                                tuple(int(x) for x in args.tt_depths.split(",")))
        else:
            line[leg] = {"batch": leg_batch, "serve": leg_serve, "calib": leg_calib,
                         "leafmodes": leg_leafmodes}[leg](pg, dev)
    dev.close()

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(jobs, args.cpu_threads)
        log("cpu baseline: %.2f G leaves/s on %d threads" % (cpu["value"] / 1e9, cpu["cores"]))
    line["cpu_baseline"] = cpu
    if "serve" in line:
        r1k = [r for r in line["serve"]["rates"] if r["rate_per_s"] == 1000][0]
        line["served_p50_ms"] = r1k["p50_ms"]
        line["served_p99_ms"] = r1k["p99_ms"]
    print(json.dumps(line))


if __name__ == "__main__":
    main()
