#!/usr/bin/env python3
"""Serving benchmark: updates/s at p99.9 <= 20 ms, batched serve vs plain GPU path.

Workload (the serving configuration the metric is quoted on): R-MAT scale 24
(16,777,216 vertices, 268,435,456 edges, a/b/c/d = 0.57/0.19/0.19/0.05,
u32 weights uniform in [1, 1000]), first 90% loaded, BFS / SSSP / SSWP / WCC
dependency trees all maintained on one MI355X.  The update stream is the
all_inc alternation (add edge 0.9E+k, delete edge k) over the circular
stream (positions mod E, so the sliding window can run past the end);
requests arrive by a Poisson process from 256 simulated clients.

Legs:
  plain   one update per device batch (rg_serve with batch cap 1), at the
          highest offered rate whose p99.9 stays within the target;
  serve   continuous batching with the safe/unsafe split (safe updates
          complete once their store batch is applied and classified; unsafe
          ones after the repair) and the adaptive admission cap, which counts
          unsafe updates (up to 2^20 updates per batch): offered rate
          ramped x1.5 per short calibration round while the server keeps up
          (>= 95% served, p99 within the target; two misses in a row end the
          ramp), then three bisection rounds; a candidate rate is prechecked
          for a few untimed steps, then the W warmup and K timed steps run
          at it (a step = one serving round of ~step_seconds of arrivals)
          and must themselves meet the target, else the rate backs off (x0.97
          when it is near the limit: its precheck passed, or missed by at most
          half the target; else x0.95) and both run again (up to 12
          attempts, within a wall-clock budget);
  check   every incremental tree is compared with a fresh full rebuild on the
          final graph (values bit-exact);
  cpu     the reference's rt serving loop restated on all host threads
          (OpenMP) on the same stream, after the GPU legs (rank 0, N=1
          only): the highest offered rate whose ~1 s round meets p99.9 <=
          20 ms, and the one-thread back-to-back rate beside it.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import risgraph_amd as rg  # noqa: E402

METRIC = "updates/s at p99.9 <= 20 ms (batched serve vs plain GPU path) + HBM GB/s"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5, help="timed serving steps")
    ap.add_argument("--warmup", type=int, default=1, help="untimed serving steps at the chosen rate")
    ap.add_argument("--scale", type=int, default=24)
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--target-ms", type=float, default=20.0)
    ap.add_argument("--step-seconds", type=float, default=0.5)
    ap.add_argument("--calib-seconds", type=float, default=0.3)
    ap.add_argument("--plain-updates", type=int, default=4000)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--json-out", default="")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="for rocprofv3 --pmc: after the setup, apply --probe-batches store batches of "
                         "--probe-pairs insert+delete pairs, then a plain 1 GiB device copy (counter calibration)")
    ap.add_argument("--probe-pairs", type=int, default=150000)
    ap.add_argument("--probe-batches", type=int, default=16)
    ap.add_argument("--serve-only", type=float, default=0.0, metavar="RATE",
                    help="for profiling: skip the plain leg and the ramp, serve the warmup and timed steps at RATE "
                         "updates/s (the line's plain and ramp fields are then empty)")
    ap.add_argument("--precheck-steps", type=int, default=6,
                    help="untimed serving steps at a candidate rate before its warmup and timed steps "
                         "(0: none); a candidate whose p99.9 misses the target backs off at once "
                         "(x0.97 when it missed by at most half the target, else x0.95)")
    ap.add_argument("--attempt-seconds", type=float, default=240.0,
                    help="wall-clock budget of the timed attempts; past it each back-off is x0.75")
    ap.add_argument("--no-split", action="store_true",
                    help="every update of a batch waits for the batch's repair (no safe/unsafe split)")
    ap.add_argument("--cap", choices=["unsafe", "total"], default="unsafe",
                    help="what the adaptive admission cap counts: unsafe updates (default) or all updates")
    ap.add_argument("--no-autotune", action="store_true",
                    help="skip the engine autotuner (rg_autotune) before serving")
    ap.add_argument("--batch-max", type=int, default=BATCH_MAX,
                    help="largest serving batch (updates); 0 = the autotuner's batch cap for the latency target")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch each batch's window, store and split kernels one by one instead of replaying "
                         "the captured hipGraph of the batch's size bucket")
    return ap.parse_args()


class Cursor:
    """Position in the all_inc stream: update u adds e[(ins0 + u/2) mod E] and
    deletes e[(u/2) mod E], a sliding window of ins0 edges that wraps around
    the circular stream, so the run length is not bounded by E - ins0."""

    def __init__(self, ins0, n_edges):
        self.ins0 = ins0
        self.n_edges = n_edges
        self.used = 0          # updates consumed

    def take(self, n):
        n -= n % 2
        ins_first = (self.ins0 + self.used // 2) % self.n_edges
        del_first = (self.used // 2) % self.n_edges
        self.used += n
        return ins_first, del_first, n


# Batched serving uses the library's adaptive batch cap (rg_serve): it
# starts at 16 unsafe updates, grows x1.5 while the backlog exceeds it and
# otherwise shrinks x0.7 when a batch's device time exceeds a quarter of the
# latency target, up to BATCH_MAX updates.  (Taking every arrived update per
# batch, with no cap, met the p99.9 target less often: 6 of 8 half-second
# rounds at 80M updates/s against 7 of 8.  Over 14 half-second windows at
# 250M updates/s offered, a cap of 2^20 gave p99.9 19.2 ms and 2^21 21.7 ms;
# at 235M both gave 16.5 ms.)
BATCH_MAX = 1 << 20


def serve(g, labels, stream, cur, n, rate, cap, cap_max, target_ms, split=True, cap_unsafe=False):
    ins_first, del_first, n = cur.take(n)
    res, lat = g.serve(labels, stream, ins_first, del_first, n, rate, batch_cap=cap, cap_max=cap_max,
                       target_ms=target_ms, latencies=True, split=split, cap_unsafe=cap_unsafe)
    return res, lat


LAT_BINS = 1 << 21   # 1-us latency bins up to ~2.1 s; the last bin holds everything above


def lat_hist(lat):
    """1-us histogram of one round's latencies (ms); unserved updates of an
    overloaded round (+inf) land in the last bin."""
    b = np.multiply(lat, 1000.0)
    np.clip(b, 0.0, LAT_BINS - 1, out=b)
    return np.bincount(b.astype(np.int64), minlength=LAT_BINS)


def hist_percentiles(h, qs):
    """Percentiles (ms) of the latencies counted in h: the upper edge of the
    1-us bin that holds rank ceil(q n) - 1, so at most 1 us above the exact
    order statistic and never below it (2097.152 ms stands for anything
    beyond the histogram)."""
    c = np.cumsum(h)
    n = int(c[-1])
    out = {}
    for q in qs:
        k = min(n - 1, int(np.ceil(q / 100.0 * n)) - 1)
        out[q] = (int(np.searchsorted(c, k + 1)) + 1) / 1000.0
    return out


def serve_batched(g, labels, stream, cur, n, rate, args):
    # with the unsafe-count cap the initial cap is in unsafe updates
    cap_unsafe = args.cap == "unsafe" and not args.no_split
    return serve(g, labels, stream, cur, n, rate, 16 if cap_unsafe else 64, args.batch_max, args.target_ms,
                 split=not args.no_split, cap_unsafe=cap_unsafe)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}): this build has no multi-GPU path; "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    rg.load_library()
    scale, V = args.scale, 1 << args.scale
    E = args.edge_factor * V
    n0 = int(0.9 * E)
    t_setup = time.time()
    stream = rg.Stream("rmat", scale, E, seed=1, wseed=2)
    g = rg.Graph(V, E, symmetric=False, dual=True, weighted=True)
    # HIP events around every engine team launch of this process; the store
    # kernels run in captured graphs here and are timed in the sample batches
    g.profile(2)
    g.capture(not args.no_graph)
    g.load_stream(stream, 0, n0, True)
    root = stream.max_degree_vertex(n0, V)
    names = ["bfs", "sssp", "sswp", "wcc"]
    labels = [g.alloc_vertex_tree_array(rg.ALGOS[nm], root, trace_modified=False) for nm in names]
    build_s = {}
    for nm, L in zip(names, labels):
        t = time.time()
        L.build_tree()
        build_s[nm] = round(time.time() - t, 3)
    log(f"setup {time.time() - t_setup:.1f}s V={V} E={E} loaded={n0} root={root} build={build_s}")
    info = g.info()
    if args.traffic_probe:
        traffic_probe(g, labels, stream, n0, E, args)
        return

    # CPU baseline: snapshot the built fixpoint and the stream before any
    # update runs; the (multi-threaded) host setup runs after the GPU
    # measurements, so it cannot steal host cores from the serving loop
    cpu = {}
    snap = None
    if not args.no_cpu:
        snap = [L.read() for L in labels]

    cur = Cursor(n0, E)
    # engine settings per batch size, timed on forward/inverse batches of the
    # current window (the graph and the values return to where they were)
    tune = None
    if not args.no_autotune:
        t = time.time()
        tune = g.autotune(labels, stream, n0, 0, target_ms=args.target_ms)
        tune["seconds"] = round(time.time() - t, 2)
        log(f"autotune {tune['seconds']}s: " + "; ".join(
            f"{b['ops']} ops cap x{b['cap_x4'] / 4:g} w {b['weight']} {b['default_ms']:.3f}->{b['best_ms']:.3f} ms"
            for b in tune["buckets"]) + f"; batch cap {tune['batch_cap']}")
    if args.batch_max == 0:
        args.batch_max = tune["batch_cap"] if tune else BATCH_MAX
    args.autotune = tune
    if args.serve_only:
        serve_only(g, labels, stream, cur, args, snap, names, root, V, E, n0, info, build_s)
        return

    # ---- plain path: one update per device batch ----
    res, _ = serve(g, labels, stream, cur, args.plain_updates, 1e9, 1, 1, 0.0)
    plain_capacity = res["updates_per_s"]
    plain = None
    rate = 0.9 * plain_capacity
    for _ in range(4):
        n = int(max(2000, min(60000, rate * args.calib_seconds)))
        res, _ = serve(g, labels, stream, cur, n, rate, 1, 1, 0.0)
        log(f"plain rate={rate:.0f} -> {res['updates_per_s']:.0f} upd/s p99.9={res['p999_ms']:.2f}ms")
        if res["p999_ms"] <= args.target_ms:
            plain = dict(res, offered=rate)
            break
        rate *= 0.7
    if plain is None:
        plain = dict(res, offered=rate)

    # ---- batched serve: calibration ramp (x1.5), then three bisection steps ----
    # a calibration round passes when the server keeps up: not overloaded,
    # >= 95% of the offered rate served and p99 within the target.  The p99.9
    # of a short round is set by one or two large deletion subtrees, so it
    # is left to the timed steps below, which must meet it on their own.
    def calibrate(rate):
        n = int(max(20000, rate * args.calib_seconds))
        res, _ = serve_batched(g, labels, stream, cur, n, rate, args)
        ok = (not res["overloaded"] and res["p99_ms"] <= args.target_ms
              and res["updates_per_s"] >= 0.95 * rate)
        log(f"serve rate={rate:.0f} -> {res['updates_per_s']:.0f} upd/s p99={res['p99_ms']:.2f}ms "
            f"p99.9={res['p999_ms']:.2f}ms mean_batch={res['mean_batch']:.1f} max_batch={res['max_batch']} "
            f"rebuilds={res['rebuilds']} {'ok' if ok else ('overloaded' if res['overloaded'] else 'over')}")
        return ok

    # the ramp stops at two failures in a row: one slow stretch of the stream
    # does not end it
    rate = max(2.0 * plain_capacity, 100000.0)
    best, fail, misses = None, None, 0
    for _ in range(40):
        if calibrate(rate):
            best, fail, misses = rate, None, 0
        else:
            fail = fail or rate
            misses += 1
            if misses == 2:
                break
        rate *= 1.5
    if best is not None and fail is not None:
        for _ in range(3):
            mid = (best * fail) ** 0.5
            if calibrate(mid):
                best = mid
            else:
                fail = mid
    if best is None:
        best = max(plain_capacity, 1000.0)
    rate = best

    # ---- warmup + timed steps at the chosen rate; the timed steps themselves
    # must meet the latency target, else the rate backs off and they rerun ----
    def timed(rate):
        n_step = int(max(20000, rate * args.step_seconds))
        for _ in range(args.warmup):
            serve_batched(g, labels, stream, cur, n_step, rate, args)
        hist = np.zeros(LAT_BINS, np.int64)
        acc = dict(wall=0.0, total=0, batches=0, grid_steps=0, gpu_ms=0.0, overloaded=0, unsafe=0,
                   safe_p50=[], unsafe_p50=[], unsafe_p999=[])
        tp0 = g.profile_read(1)   # the timed steps' engine team launches (HIP events)
        for k in range(args.steps):
            res, lat = serve_batched(g, labels, stream, cur, n_step, rate, args)
            hist += lat_hist(lat)
            del lat
            acc["wall"] += res["wall_s"]
            acc["total"] += res["served"]
            acc["batches"] += res["batches"]
            acc["grid_steps"] += res["grid_steps"]
            acc["gpu_ms"] += res["gpu_ms"]
            acc["overloaded"] += res["overloaded"]
            acc["unsafe"] += res["unsafe"]
            acc["safe_p50"].append(res["p50_safe_ms"])
            acc["unsafe_p50"].append(res["p50_unsafe_ms"])
            acc["unsafe_p999"].append(res["p999_unsafe_ms"])
            log(f"step {k}: {res['updates_per_s']:.0f} upd/s p99.9={res['p999_ms']:.2f}ms "
                f"mean_batch={res['mean_batch']:.1f} max_batch={res['max_batch']} rebuilds={res['rebuilds']} "
                f"gpu_busy={res['gpu_ms'] / 1e3 / max(res['wall_s'], 1e-9):.2f}")
        tp1 = g.profile_read(1)
        acc["team_launches"] = tp1["launches"] - tp0["launches"]
        acc["team_ms"] = tp1["ms"] - tp0["ms"]
        p = hist_percentiles(hist, (50, 99, 99.9))   # unserved updates of an overloaded step count as +inf
        value = acc["total"] / acc["wall"]
        log(f"timed {args.steps} steps at {rate:.0f}/s: {value:.0f} upd/s p50={p[50]:.2f} p99={p[99]:.2f} "
            f"p99.9={p[99.9]:.2f}ms")
        return value, p, acc, n_step

    # a short untimed check at the rate before the warmup and timed steps
    # are spent on it: the stream's slow regions recur every few seconds
    # (one circle of the stream), which a few steps already cover
    def precheck(rate):
        n_step = int(max(20000, rate * args.step_seconds))
        hist, over = np.zeros(LAT_BINS, np.int64), 0
        for _ in range(args.precheck_steps):
            res, lat = serve_batched(g, labels, stream, cur, n_step, rate, args)
            hist += lat_hist(lat)
            del lat
            over += res["overloaded"]
        p999 = hist_percentiles(hist, (99.9,))[99.9]
        ok = p999 <= args.target_ms and not over
        log(f"precheck {args.precheck_steps} steps at {rate:.0f}/s: p99.9={p999:.2f}ms {'ok' if ok else 'miss'}")
        return ok, p999

    attempts = []
    deadline = time.time() + args.attempt_seconds
    met, result = False, None
    near = False
    for i in range(12):
        if i:
            # a rate that missed narrowly (its precheck passed, or missed by at
            # most half the target) is near the limit: a finer step; past the
            # time budget the rate drops faster, so the run ends in time
            rate *= (0.97 if near else 0.95) if time.time() < deadline else 0.75
        near = False
        if args.precheck_steps and i < 11:
            ok, p999 = precheck(rate)
            if not ok:
                attempts.append({"offered": round(rate, 1), "precheck_p99.9_ms": round(p999, 3)})
                near = p999 <= 1.5 * args.target_ms
                continue
        result = timed(rate)
        value, p, acc, n_step = result
        met = p[99.9] <= args.target_ms and not acc["overloaded"]
        attempts.append({"offered": round(rate, 1), "value": round(value, 1), "p99.9_ms": round(p[99.9], 3)})
        if met:
            break
        near = True
    wall, total, batches = acc["wall"], acc["total"], acc["batches"]
    grid_steps, gpu_ms, overloaded = acc["grid_steps"], acc["gpu_ms"], acc["overloaded"]

    # ---- roofline: dominant kernel (store apply) timed live with HIP events ----
    roof, roof_store, batch_roof = roofline_sample(g, labels, stream, cur, rate, n_step, args,
                                                   int(total / max(batches, 1)),
                                                   (acc["team_launches"], acc["team_ms"]))

    # ---- correctness: incremental trees vs a fresh rebuild ----
    check = None
    if not args.no_check:
        check = {}
        for nm, L in zip(names, labels):
            fresh = g.alloc_vertex_tree_array(rg.ALGOS[nm], root, trace_modified=False)
            fresh.build_tree()
            a, _ = L.read()
            b, _ = fresh.read()
            check[nm] = bool(np.array_equal(a, b))
            fresh.close()
        log(f"values vs full rebuild: {check}")

    if snap is not None:
        src, dst, w = stream.read(0, E)
        cpu_baseline_setup(cpu, src, dst, w, snap, labels, V, E, n0, root)
        cpu_baseline_run(cpu, args.cpu_seconds)

    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * wall / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32/u64 integer",
        "data": "synthetic (seeded R-MAT stream generated on device)",
        "config": {
            "workload": f"serve all_inc mixed stream, R-MAT scale {scale}, 4 trees (bfs,sssp,sswp,wcc)",
            "vertices": V, "edges": E, "loaded": n0, "root": root, "clients": 256,
            "offered_rate": round(rate, 1), "target_p999_ms": args.target_ms,
            "step_updates": n_step,
        },
        "latency_ms": {"p50": round(p[50], 3), "p99": round(p[99], 3), "p99.9": round(p[99.9], 3),
                       "p50_safe_median_step": round(float(np.median(acc["safe_p50"])), 3),
                       "p50_unsafe_median_step": round(float(np.median(acc["unsafe_p50"])), 3),
                       "p99.9_unsafe_worst_step": round(float(max(acc["unsafe_p999"])), 3)},
        "split": not args.no_split,
        "graph_replay": g.capture_info(),
        "autotune": args.autotune,
        "batch_max": args.batch_max,
        "cap_counts": "unsafe updates" if (args.cap == "unsafe" and not args.no_split) else "all updates",
        "safe_frac": round(1.0 - acc["unsafe"] / max(total, 1), 4),
        "target_met": met,
        "timed_attempts": attempts,
        "plain_gpu": {"updates_per_s": round(plain["updates_per_s"], 1), "offered": round(plain["offered"], 1),
                      "capacity": round(plain_capacity, 1), "p50_ms": round(plain["p50_ms"], 3),
                      "p99.9_ms": round(plain["p999_ms"], 3)},
        "serve_vs_plain": round(value / max(plain["updates_per_s"], 1e-9), 2),
        "batches": batches, "mean_batch": round(total / max(batches, 1), 1), "grid_steps": grid_steps,
        "overloaded_steps": overloaded, "gpu_busy_frac": round(gpu_ms / 1e3 / max(wall, 1e-9), 3),
        "values_match_rebuild": check,
        "build_s": build_s,
        "store": info,
        "pool": g.pool_info(),
        "roofline": roof,
        "roofline_store": roof_store,
        "batch_roofline": batch_roof,
        "cpu_baseline": cpu.get("result"),
    }
    line = json.dumps(out)
    print(line, flush=True)
    if args.json_out:
        with open(args.json_out, "w") as f:
            f.write(line + "\n")


def serve_only(g, labels, stream, cur, args, snap, names, root, V, E, n0, info, build_s):
    """Profiling form of the run: no plain leg and no ramp; W warmup and K timed
    steps of serving at a fixed offered rate, then the roofline sample, so a
    kernel trace of the command holds the serving kernels alone."""
    rate = args.serve_only
    n_step = int(max(20000, rate * args.step_seconds))
    hist, wall, total, batches = np.zeros(LAT_BINS, np.int64), 0.0, 0, 0
    tp0 = None
    for k in range(args.warmup + args.steps):
        if k == args.warmup:
            tp0 = g.profile_read(1)
        res, lat = serve_batched(g, labels, stream, cur, n_step, rate, args)
        log(f"serve-only step {k}: {res['updates_per_s']:.0f} upd/s p99.9={res['p999_ms']:.2f}ms "
            f"mean_batch={res['mean_batch']:.1f}")
        if k >= args.warmup:
            hist += lat_hist(lat)
            wall += res["wall_s"]
            total += res["served"]
            batches += res["batches"]
    tp1 = g.profile_read(1)
    roof, roof_store, batch_roof = roofline_sample(g, labels, stream, cur, rate, n_step, args,
                                                   int(total / max(batches, 1)),
                                                   (tp1["launches"] - tp0["launches"], tp1["ms"] - tp0["ms"]))
    out = {"mode": "serve-only profiling run", "offered_rate": rate, "value": round(total / max(wall, 1e-9), 1),
           "p99.9_ms": round(hist_percentiles(hist, (99.9,))[99.9], 3), "mean_batch": round(total / max(batches, 1), 1),
           "roofline": roof, "roofline_store": roof_store, "batch_roofline": batch_roof,
           "autotune": args.autotune, "batch_max": args.batch_max}
    print(json.dumps(out), flush=True)
    if args.json_out:
        with open(args.json_out, "w") as f:
            f.write(json.dumps(out) + "\n")


# ---------------------------------------------------------------------------
# roofline
# ---------------------------------------------------------------------------
B_STORE = 32                      # bytes per directed edge op (out+in: key, weight, count RMW)
B_VERTEX = 32                     # per changed vertex: label word r/w + out/in list headers
B_EDGE = {"bfs": 12, "sssp": 16, "sswp": 16, "wcc": 12}   # nbr id (+ weight) + neighbour label word


def traffic_probe(g, labels, stream, n0, E, args):
    """Fixed-size store batches (the serve mix: inserts + deletes, all four
    trees repaired) and a plain device copy of known size, for the HBM
    FETCH_SIZE / WRITE_SIZE counter runs."""
    pos = 0
    m = args.probe_pairs
    names = ["bfs", "sssp", "sswp", "wcc"]
    out_deg, in_deg = g.degrees()
    deg = out_deg.astype(np.int64) + in_deg.astype(np.int64)
    eng_bytes, team_l, grid_l, changed = 0, 0, 0, 0
    for _ in range(args.probe_batches):
        before = [L.read()[0] for L in labels]
        t0, r0 = g.profile_read(1), g.profile_read(2)
        g.apply_stream(labels, stream, (n0 + pos) % E, m, pos % E, m, True)
        t1, r1 = g.profile_read(1), g.profile_read(2)
        b, c = delta_bytes(names, labels, before, deg)
        eng_bytes += b
        changed += c
        team_l += t1["launches"] - t0["launches"]
        grid_l += r1["launches"] - r0["launches"]
        pos += m
    # the HIP runtime's device-to-device copy (its blit kernel) of 1 GiB
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    nbytes = 1 << 30
    src_p, dst_p = ctypes.c_void_p(), ctypes.c_void_p()
    ok = hip.hipMalloc(ctypes.byref(src_p), ctypes.c_size_t(nbytes)) == 0
    ok = ok and hip.hipMalloc(ctypes.byref(dst_p), ctypes.c_size_t(nbytes)) == 0
    if ok:
        for _ in range(3):
            hip.hipMemcpy(dst_p, src_p, ctypes.c_size_t(nbytes), 3)   # hipMemcpyDeviceToDevice
        hip.hipDeviceSynchronize()
        hip.hipFree(src_p)
        hip.hipFree(dst_p)
    print(json.dumps({"probe": "traffic", "store_batch_edge_ops": 2 * m, "batches": args.probe_batches,
                      "copy_ok": bool(ok), "copy_bytes_read": nbytes, "copy_bytes_written": nbytes,
                      "engine_algorithmic_bytes": eng_bytes, "changed_vertices": changed,
                      "team_launches": team_l, "grid_steps": grid_l,
                      "engine_bytes_per_team_launch": round(eng_bytes / max(team_l, 1), 1)}),
          flush=True)


def measured_traffic():
    """HBM bytes per store-apply launch from the committed counter summary
    (profiles/k_apply_all_traffic_rmat24.md), with the batch size it was taken at."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "k_apply_all_traffic_rmat24.md")
    try:
        vals = {}
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_"):
                    k, v = line.split(":", 1)
                    vals[k.strip()] = float(v.strip())
        return vals or None
    except OSError:
        return None


def engine_traffic():
    """HBM bytes per engine team launch (rg::k_team) from the committed counter
    summary (profiles/k_team_traffic_rmat24.md), with the algorithmic bytes per
    launch of the same probe batches."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "k_team_traffic_rmat24.md")
    try:
        vals = {}
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_"):
                    k, v = line.split(":", 1)
                    vals[k.strip()] = float(v.strip())
        return vals or None
    except OSError:
        return None


def delta_bytes(names, labels, before, deg):
    """Algorithmic bytes of one batch's repair from its changed-vertex set:
    per changed vertex B_VERTEX + (out+in live degree) x B_EDGE[tree]."""
    b, changed = 0, 0
    for nm, L, v0 in zip(names, labels, before):
        d = np.nonzero(L.read()[0] != v0)[0]
        changed += len(d)
        b += len(d) * B_VERTEX + int(deg[d].sum()) * B_EDGE[nm]
    return b, changed


def roofline_sample(g, labels, stream, cur, rate, n_step, args, mean_batch, team_timed):
    """The dominant kernel is the persistent engine launch (rg::k_team): its
    algorithmic bytes are the changed-vertex bytes of whole batches (the
    store's 32 B per edge op left out), divided over the batches' team
    launches; its average duration is the HIP-event average over every team
    launch of the process (the dispatches rocprofv3 averages for this
    command).  Grid-wide passes of the same batches, which share that work,
    are added into a second, engine-level figure.  The store apply kernel
    (rg::k_apply_all) keeps its own line.  Batch level: all bytes over the
    batch's device time."""
    names = ["bfs", "sssp", "sswp", "wcc"]
    out_deg, in_deg = g.degrees()
    deg = out_deg.astype(np.int64) + in_deg.astype(np.int64)
    g.profile(7)
    tot_bytes, eng_bytes, tot_ms, changed = 0, 0, 0.0, 0
    team_l, team_ms, grid_ms = 0, 0.0, 0.0
    batch = max(2, mean_batch - mean_batch % 2)
    nb = 8   # change sets vary widely from batch to batch (3 batches gave 54-110 MB per launch)
    for _ in range(nb):
        before = [L.read()[0] for L in labels]
        t0, r0 = g.profile_read(1), g.profile_read(2)
        ins_first, del_first, m = cur.take(batch)
        g.apply_stream(labels, stream, ins_first, m // 2, del_first, m // 2, True)
        tot_ms += g.last_gpu_ms()
        t1, r1 = g.profile_read(1), g.profile_read(2)
        team_l += t1["launches"] - t0["launches"]
        team_ms += t1["ms"] - t0["ms"]
        grid_ms += r1["ms"] - r0["ms"]
        b, c = delta_bytes(names, labels, before, deg)
        changed += c
        eng_bytes += b
        tot_bytes += b + B_STORE * m
    g.profile(2)
    # the team launches of the timed serving steps: the dispatches that
    # rocprofv3 averages for the bench's serve-only form at the same rate
    team_l_timed, team_ms_timed = team_timed
    team_avg_ms = team_ms_timed / max(team_l_timed, 1)
    per_launch = eng_bytes / max(team_l, 1)
    achieved = per_launch / (team_avg_ms * 1e-3) / 1e9 if team_avg_ms > 0 else 0.0
    eng_ach = eng_bytes / ((team_ms + grid_ms) * 1e-3) / 1e9 if team_ms + grid_ms > 0 else 0.0
    roof = {
        "kernel": "rg::k_team (persistent engine launch: all four trees' repair passes of a batch)",
        "bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": None,
        "algorithmic_bytes_per_launch": round(per_launch, 1),
        "bytes_basis": f"changed-vertex set of {nb} sample batches at the operating batch size "
                       f"({batch} updates): {B_VERTEX} B per changed vertex + live degree x "
                       f"{B_EDGE} B per tree, over their {team_l} team launches",
        "avg_launch_ms": round(team_avg_ms, 5), "launches": team_l_timed,
        "avg_basis": "HIP events around every team launch of the timed serving steps",
        "engine_level": {"bytes": eng_bytes, "team_ms": round(team_ms, 4), "grid_ms": round(grid_ms, 4),
                         "achieved": round(eng_ach, 3), "frac": round(eng_ach / HBM_PEAK_GBS, 6)},
    }
    tr = engine_traffic()
    if tr and "traffic_bytes_per_launch" in tr:
        roof["traffic"] = tr["traffic_bytes_per_launch"]
        roof["traffic_algorithmic_bytes_per_launch"] = tr.get("traffic_algorithmic_bytes_per_launch")
        roof["traffic_batch_pairs"] = int(tr.get("traffic_batch_pairs", 0))
    # the store apply kernel (one launch per batch of more than 64 edge ops),
    # timed in the sample batches above: timing bit 0 launches the store
    # kernels of a batch one by one, outside a captured graph
    ap = g.profile_read(0)
    launch_ms = ap["ms"] / max(ap["launches"], 1)
    bytes_per_launch = B_STORE * ap["units"] / max(ap["launches"], 1)
    st_ach = bytes_per_launch / (launch_ms * 1e-3) / 1e9 if launch_ms > 0 else 0.0
    roof_store = {
        "kernel": "rg::k_apply_all (store insert/delete batch: light and heavy lists in one launch)",
        "bound": "hbm", "achieved": round(st_ach, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(st_ach / HBM_PEAK_GBS, 6), "traffic": None,
        "algorithmic_bytes_per_launch": round(bytes_per_launch, 1),
        "avg_launch_ms": round(launch_ms, 5), "launches": ap["launches"],
    }
    tr = measured_traffic()
    if tr and "traffic_bytes_per_launch" in tr:
        roof_store["traffic"] = tr["traffic_bytes_per_launch"]
        roof_store["traffic_batch_edge_ops"] = int(tr.get("traffic_batch_edge_ops", 0))
        roof_store["traffic_algorithmic_bytes"] = tr.get("traffic_algorithmic_bytes")
    ach = tot_bytes / (tot_ms * 1e-3) / 1e9 if tot_ms > 0 else 0.0
    batch_roof = {"batch_updates": batch, "algorithmic_bytes_per_batch": round(tot_bytes / nb, 1),
                  "changed_vertices_per_batch": round(changed / nb, 1),
                  "batch_gpu_ms": round(tot_ms / nb, 4), "achieved": round(ach, 3), "unit": "GB/s",
                  "frac": round(ach / HBM_PEAK_GBS, 6)}
    log(f"roofline: k_team {achieved:.2f} GB/s ({per_launch / 1e6:.2f} MB per launch, "
        f"{team_avg_ms * 1e3:.1f} us avg); engine {eng_ach:.2f} GB/s; k_apply_all {st_ach:.2f} GB/s; "
        f"batch {ach:.2f} GB/s over {tot_ms / nb:.3f} ms/batch")
    return roof, roof_store, batch_roof


# ---------------------------------------------------------------------------
# CPU baseline: the oracle's sequential port of the reference engine
# ---------------------------------------------------------------------------
def cpu_baseline_setup(cpu, src, dst, w, snap, labels, V, E, n0, root):
    try:
        import oracle
        t = time.time()
        w = w.astype(np.uint64)
        gd = oracle.RefGraph(V, E, symmetric=False, dual=True, weighted=True)
        gd.bulk_load(src[:n0], dst[:n0], w[:n0], True)
        gu = oracle.RefGraph(V, E, symmetric=True, dual=False, weighted=False)
        gu.bulk_load(src[:n0], dst[:n0], None, False)
        Ld = []
        for L, (v, p) in zip(labels[:3], snap[:3]):
            R = oracle.RefLabels(gd, L.algo, root)
            R.assign(v, p)          # start from the same fixpoint as the device
            Ld.append(R)
        Lu = oracle.RefLabels(gu, oracle.WCC, root)
        v, p = snap[3]
        Lu.assign(v, p)
        cpu.update(src=src, dst=dst, w=w, gd=gd, gu=gu, Ld=Ld, Lu=Lu, n0=n0, setup_s=time.time() - t)
        log(f"cpu baseline setup {time.time() - t:.1f}s")
    except Exception as e:  # reported, never silently replaced
        cpu["error"] = repr(e)
        log(f"cpu baseline setup failed: {e!r}")


def host_info():
    """Host facts the CPU baseline depends on: core count, model, sockets, RTM
    (without RTM the reference takes its lock path for weighted labels,
    graph.hpp:1781)."""
    info = {"nproc": os.cpu_count(), "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS")}
    try:
        with open("/proc/cpuinfo") as f:
            txt = f.read()
        models = [l.split(":", 1)[1].strip() for l in txt.splitlines() if l.startswith("model name")]
        phys = {l.split(":", 1)[1].strip() for l in txt.splitlines() if l.startswith("physical id")}
        flags = next((l.split(":", 1)[1].split() for l in txt.splitlines() if l.startswith("flags")), [])
        info.update(model=models[0] if models else None, sockets=len(phys) or 1, rtm="rtm" in flags)
    except OSError:
        pass
    return info


def cpu_baseline_run(cpu, seconds):
    """The reference's rt serving loop on all host threads (OpenMP restatement,
    oracle.serve_rt: parallel safe phase, serial unsafe phase, the
    num_changes_threshold cap; bfs_inc_rt.cpp:154-431) with Poisson arrivals:
    offered rate ramped x1.5 from 50k/s while a ~1 s round meets p99.9 <= 20 ms
    and keeps up; the answer is the best passing round's served rate.  Then the
    one-thread back-to-back rate of the same engine as a second figure."""
    if "error" in cpu:
        cpu["result"] = {"error": cpu["error"]}
        return
    import oracle
    n0, src, dst, w = cpu["n0"], cpu["src"], cpu["dst"], cpu["w"]
    rng = np.random.default_rng(4)
    k = 0            # stream updates consumed (even: whole add/del pairs)
    rate, best, trials = 50000.0, None, []
    t_start = time.time()
    while time.time() - t_start < seconds:
        n = int(max(20000, rate * 1.0))
        n -= n % 2
        if n0 + (k + n) // 2 >= len(src):
            break
        arrival = np.cumsum(rng.exponential(1.0 / rate, n))
        h = k // 2
        st, lat = oracle.serve_rt(cpu["gd"], cpu["Ld"], cpu["gu"], cpu["Lu"], n, src[n0 + h:], dst[n0 + h:],
                                  w[n0 + h:], src[h:], dst[h:], w[h:], arrival, timeout_ms=20.0)
        served = st["served"]
        wall = st["wall_ns"] / 1e9
        p999 = float(np.percentile(lat[:served], 99.9)) if served else float("inf")
        ok = served == n and not st["overloaded"] and p999 <= 20.0
        trials.append({"offered": round(rate, 1), "served_per_s": round(served / max(wall, 1e-9), 1),
                       "p99.9_ms": round(p999, 3), "unsafe_frac": round(st["unsafe"] / max(served, 1), 4),
                       "threshold": st["threshold"], "ok": ok})
        log(f"cpu rt rate={rate:.0f} -> {served / max(wall, 1e-9):.0f} upd/s p99.9={p999:.2f}ms "
            f"threads={st['threads']} {'ok' if ok else 'miss'}")
        threads = st["threads"]
        if not ok:
            break
        best = trials[-1]
        k += n
        rate *= 1.5
    # one thread, back to back (the previous baseline), a bounded sample
    h = k // 2
    t = time.time()
    n1 = 20000
    _, lat1 = oracle.stream_mixed(cpu["gd"], cpu["Ld"], cpu["gu"], cpu["Lu"], n1, src[n0 + h:], dst[n0 + h:],
                                  w[n0 + h:], src[h:], dst[h:], w[h:], latency=True)
    el1 = time.time() - t
    cpu["result"] = {
        "value": best["served_per_s"] if best else 0.0, "unit": "updates/s at p99.9 <= 20 ms",
        "cores": threads if trials else 0, "kind": "port",
        "sample": "the reference's rt serving loop (parallel safe phase, serial unsafe phase, unsafe-change cap) "
                  "restated in C with OpenMP, on the same all_inc stream and 4 trees at full scale, started "
                  "from the device's built fixpoint; Poisson arrivals, ~1 s rounds at rates x1.5 from 50k/s "
                  "until one misses p99.9 <= 20 ms",
        "trials": trials,
        "host": host_info(),
        "single_thread_back_to_back": {"value": round(n1 / el1, 1), "unit": "updates/s", "cores": 1,
                                       "p50_ms": round(float(np.percentile(lat1, 50)) / 1e6, 4),
                                       "p99.9_ms": round(float(np.percentile(lat1, 99.9)) / 1e6, 4)},
        "setup_s": round(cpu["setup_s"], 1),
    }


if __name__ == "__main__":
    main()
