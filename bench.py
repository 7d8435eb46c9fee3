"""Benchmark of the table-commitment hot path on one MI355X.

Workload (default): main-table commit at 2^20 rows x 379 Goldilocks columns,
coset LDE x4 (offset 7) to 2^22 rows, Tip5 hash of every row, Merkle tree over
the 2^22 row digests -- the GPU configuration the metric is quoted on.  One
step = LDE + row hashing + Merkle build with the trace already resident in
HBM and all outputs left on the device.  Inputs are synthetic uniform field
elements drawn on the host (numpy PCG64) and uploaded once, outside the timed
region.

Prints one JSON line.  `value` is output codeword cells per second
(rows_out * columns / step time).  `roofline` is the LDE (the HBM-bound NTT
passes): compulsory bytes 8*N*W + 8*M*W per launch over the measured LDE time
against 8 TB/s.  `tip5` reports the Tip5 phase (VALU-bound integer work) as
permutations per second.  `cpu_baseline` times the reference's own CPU code
(oracle/_ref, built from the reference sources) on a bounded sample on the host.
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
METRIC = "LDE+Tip5 Merkle commit cells/s & HBM GB/s at 2^20x379; p50/p99 job latency"
HBM_PEAK = 8.0e12


def _tvh():
    spec = importlib.util.spec_from_file_location(
        "tvh", os.path.join(ROOT, "triton-vm-prover_amd", "tvh.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["tvh"] = mod
    spec.loader.exec_module(mod)
    return mod


def random_field(numel: int, seed: int):
    """Uniform canonical Goldilocks values (numpy PCG64)."""
    import numpy as np

    rng = np.random.default_rng(seed)
    return rng.integers(0, 0xFFFFFFFF00000001, size=numel, dtype=np.uint64)


def measured_traffic(log_n: int, width: int, expansion: int):
    """HBM bytes per LDE launch from the committed rocprof counter summary
    (FETCH_SIZE and WRITE_SIZE passes, copy-calibrated), for the workload it covers."""
    path = os.path.join(ROOT, "profiles", f"commit_2p{log_n}x{width}_hbm_counters.md")
    if expansion != 4 or not os.path.exists(path):
        return None
    import re

    m = re.search(r"traffic_bytes_per_launch:\s*(\d+)", open(path).read())
    return int(m.group(1)) if m else None


def cpu_baseline(log_n: int = 12, width: int = 379, expansion: int = 4) -> dict:
    """Reference CPU commit (oracle/_ref) on a 2^log_n x width sample, 1 thread."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Oracle, Reference

    orc = Oracle()
    n, m = 1 << log_n, expansion << log_n
    trace = orc.table_rowmajor_draw(1, n, width)
    if Reference.available():
        eng, kind = Reference(), "reference"
        eng.set_threads(1)
    else:
        eng, kind = orc, "port"
    t0 = time.perf_counter()
    eng.commit(trace, n, width, m)
    dt = time.perf_counter() - t0
    return {"value": m * width / dt, "unit": "cells/s", "cores": 1, "kind": kind,
            "sample": f"commit of a 2^{log_n} x {width} trace (LDE x{expansion} to "
                      f"2^{log_n + expansion.bit_length() - 1} rows, Tip5 rows, Merkle), "
                      f"{dt:.2f} s on 1 host thread"}


def splitmix_stream(seed: int):
    s = seed
    while True:
        s = (s + 0x9E3779B97F4A7C15) & (2**64 - 1)
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & (2**64 - 1)
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & (2**64 - 1)
        yield z ^ (z >> 31)


def batch_shapes(count: int):
    """(log2 height, width) per table: two SplitMix64(7) draws per table,
    height 2^(12 + d1 mod 9), width [264, 156, 379][d2 mod 3]."""
    g = splitmix_stream(7)
    out = []
    for _ in range(count):
        d1, d2 = next(g), next(g)
        out.append((12 + d1 % 9, [264, 156, 379][d2 % 3]))
    return out


def run_batch(args, mod, lib) -> dict:
    """Mixed-height batch: per-table tvh_commit_dev vs one batch plan per wave
    (plain launch and hipGraph replay).  Waves bound HBM use; every root of the
    batched run is compared with the per-table run."""
    import numpy as np

    hip_rt = mod.hip_rt
    E = args.expansion
    shapes = batch_shapes(args.tables)
    order = sorted(range(len(shapes)), key=lambda i: -(shapes[i][1] << shapes[i][0]))
    budget = int(args.wave_gib * (1 << 30))

    def table_bytes(lg, w):
        n = 1 << lg
        return 8 * n * w + 8 * E * n * w + 40 * (2 * E * n)

    waves, cur, used = [], [], 0
    for i in order:
        b = table_bytes(*shapes[i])
        if cur and used + b > budget:
            waves.append(cur)
            cur, used = [], 0
        cur.append(i)
        used += b
    if cur:
        waves.append(cur)
    tuned = autotune(args, lib, [(1 << lg, w, E << lg) for lg, w in shapes])
    stream = hip_rt.Stream()
    t_plain = t_plan = t_graph = 0.0
    cells = 0
    mismatches = 0
    for wave in waves:
        tabs = []
        for i in wave:
            lg, w = shapes[i]
            n, m = 1 << lg, E << lg
            tr = hip_rt.DeviceBuffer(n * w)
            lib.fill_uniform(tr, n * w, 1000 + i, stream)
            tabs.append(dict(n=n, w=w, m=m, trace=tr, cw=hip_rt.DeviceBuffer(m * w),
                             tree=hip_rt.DeviceBuffer(5 * (2 * m - 1))))
            cells += m * w
        stream.synchronize()
        descs = [lib.table_desc(t["trace"], t["w"], t["n"], t["m"], t["cw"], t["tree"])
                 for t in tabs]
        plan = lib.batch_plan(descs)
        # warm-up of both paths (twiddle/coset tables, graph capture)
        lib.batch_launch_graph(plan, stream)
        for t in tabs:
            lib.commit(t["trace"], t["w"], t["n"], None, 0, 1, 7, t["m"], t["cw"], t["tree"], 0,
                       stream)
        stream.synchronize()
        ev = [hip_rt.Event() for _ in range(6)]
        ev[0].record(stream)
        for t in tabs:
            lib.commit(t["trace"], t["w"], t["n"], None, 0, 1, 7, t["m"], t["cw"], t["tree"], 0,
                       stream)
        ev[1].record(stream)
        stream.synchronize()
        roots = [t["tree"].download(5, 5 * (2 * t["m"] - 2)) for t in tabs]
        ev[2].record(stream)
        lib.batch_launch(plan, stream)
        ev[3].record(stream)
        ev[4].record(stream)
        lib.batch_launch_graph(plan, stream)
        ev[5].record(stream)
        stream.synchronize()
        for t, r in zip(tabs, roots):
            if not np.array_equal(t["tree"].download(5, 5 * (2 * t["m"] - 2)), r):
                mismatches += 1
        t_plain += ev[0].elapsed_ms(ev[1])
        t_plan += ev[2].elapsed_ms(ev[3])
        t_graph += ev[4].elapsed_ms(ev[5])
        lib.batch_destroy(plan)
        del tabs
    # host-buffer drop-in per table (trace up, codewords and digests down), each once
    t_host = 0.0
    if not args.no_host_baseline:
        for i, (lg, w) in enumerate(shapes):
            t_host += plain_host_seconds(lib, 1 << lg, w, E << lg, 5000 + i)
    return {
        "metric": "batched LDE+Tip5 Merkle commit cells/s (mixed-height tables)",
        "value": cells / (t_graph / 1000), "unit": "cells/s", "n_gpus": 1,
        "higher_is_better": True, "dtype": "u64", "data": "synthetic (tvh_fill_uniform_dev)",
        "config": {"workload": f"{len(shapes)} tables, log2 height 12..20, widths "
                               f"{{264,156,379}} (SplitMix64 seed 7), LDE x{E}, Tip5 rows, Merkle",
                   "tables": len(shapes), "waves": len(waves), "total_cells": cells},
        "ms": {"per_table_commits": t_plain, "batch_plan": t_plan, "batch_graph": t_graph},
        "cells_per_s": {"per_table": cells / (t_plain / 1000), "batch_plan": cells / (t_plan / 1000),
                        "batch_graph": cells / (t_graph / 1000)},
        "speedup_vs_per_table": t_plain / t_graph,
        "host_dropin_per_table": None if args.no_host_baseline else {
            "s": t_host, "cells_per_s": cells / t_host, "batch_graph_speedup": t_host / (t_graph / 1000)},
        "lde_forms": tuned,
        "root_mismatches": mismatches,
    }


def pct(xs, q):
    import numpy as np

    return float(np.percentile(np.asarray(xs, dtype=float), q)) if len(xs) else None


def serve_stream(args, count, seed):
    """Job stream: log2 trace height uniform in [serve_min_log, serve_max_log], width
    args.width, Poisson arrivals at args.rate jobs/s (all at t=0 when rate <= 0), 80
    random leaf queries per job, per-job seeds."""
    import numpy as np

    rng = np.random.default_rng(seed)
    lgs = rng.integers(args.serve_min_log, args.serve_max_log + 1, size=count)
    gaps = rng.exponential(1.0 / args.rate, size=count) if args.rate > 0 else np.zeros(count)
    arrivals = np.cumsum(gaps) - gaps[0]
    jobs = []
    for i in range(count):
        n = 1 << int(lgs[i])
        m = args.expansion * n
        jobs.append(dict(i=i, n=n, m=m, w=args.width, seed=(seed << 32) + i, t=float(arrivals[i]),
                         leaves=rng.integers(0, m, size=args.queries, dtype=np.uint64)))
    return jobs


def serve_run(args, lib, jobs, keep=()):
    """Submits `jobs` on their arrival schedule to one server; a second thread waits for
    each job in submission order, answers its queries (opened rows + authentication
    paths) and releases it.  Returns per-job times and the answers of jobs in `keep`."""
    import threading

    srv = lib.server_create(hbm_budget_bytes=int(args.hbm_gib * (1 << 30)),
                            max_inflight=args.inflight)
    ids = [None] * len(jobs)
    ready = [threading.Event() for _ in jobs]
    res = [None] * len(jobs)
    kept = {}
    err = []

    def consumer():
        try:
            for k, j in enumerate(jobs):
                ready[k].wait()
                if ids[k] is None:  # the submitter stopped
                    return
                root, (ts, ta, td) = lib.server_wait(srv, ids[k])
                rows, _ = lib.server_open(srv, ids[k], j["leaves"], j["w"], want_auth=False)
                paths = lib.server_paths(srv, ids[k], j["leaves"], j["m"])
                tq = time.monotonic()
                lib.server_release(srv, ids[k])
                res[k] = (ts, ta, td, tq, root)
                if j["i"] in keep:
                    kept[j["i"]] = (root, rows, paths)
                if (k + 1) % 250 == 0:
                    print(f"serve: {k + 1}/{len(jobs)} jobs answered", file=sys.stderr, flush=True)
        except Exception as e:  # surfaced by the caller
            err.append(e)

    th = threading.Thread(target=consumer)
    th.start()
    t0 = time.monotonic()
    try:
        for k, j in enumerate(jobs):
            delay = t0 + j["t"] - time.monotonic()
            if delay > 0:
                time.sleep(delay)
            ids[k] = lib.server_submit(srv, j["w"], j["n"], j["m"], seed=j["seed"])
            ready[k].set()
    finally:
        for e in ready:
            e.set()
        th.join()
        lib.server_destroy(srv)
    if err:
        raise err[0]
    return res, kept


def plain_job(lib, hip_rt, j, stream, want_rows=False):
    """The plain per-table drop-in path for one job: allocate, draw the trace,
    tvh_commit_dev, wait, gather the authentication paths, free.  Returns the
    seconds of that sequence; the queried rows (for checks) are read outside it."""
    import numpy as np

    t0 = time.perf_counter()
    n, m, w = j["n"], j["m"], j["w"]
    tr, cw = hip_rt.DeviceBuffer(n * w), hip_rt.DeviceBuffer(m * w)
    tree = hip_rt.DeviceBuffer(5 * (2 * m - 1))
    lib.fill_uniform(tr, n * w, j["seed"], stream)
    lib.commit(tr, w, n, None, 0, 1, 7, m, cw, tree, 0, stream)
    stream.synchronize()
    root = tree.download(5, 5 * (2 * m - 2))
    h = m.bit_length() - 1
    idx = np.concatenate([lib.auth_path_indices(m, int(x)) for x in j["leaves"]]).astype(np.uint64)
    d_idx = hip_rt.to_device(idx)
    d_out = hip_rt.DeviceBuffer(5 * idx.size)
    lib.merkle_gather(tree, m, 0, d_idx, idx.size, d_out, stream)
    stream.synchronize()
    paths = d_out.download().reshape(len(j["leaves"]), h, 5)
    dt = time.perf_counter() - t0
    rows = None
    if want_rows and w:
        rows = np.stack([cw.gather_strided(int(x), m, w) for x in j["leaves"]])
    del tr, cw, tree, d_idx, d_out
    return dt, [int(v) for v in root], rows, paths


def plain_host_seconds(lib, n, w, m, seed):
    """The host-buffer drop-in (tvh_commit_host, the CudaBackend-style path of the
    CPU prover: trace uploaded, codewords and digests downloaded) for one table;
    seconds of that call.  Inputs are uniform values from numpy, drawn untimed."""
    import numpy as np

    tr = random_field(n * w, seed)
    cw = np.empty(m * w, np.uint64)
    dg = np.empty(m * 5, np.uint64)
    root = np.empty(5, np.uint64)
    t0 = time.perf_counter()
    lib.commit_host(tr, w, n, m, codewords=cw, digests=dg, root=root)
    return time.perf_counter() - t0


def run_serve(args, mod, lib) -> dict:
    """Served stream (Poisson arrivals, latency) and saturated stream (capacity) through
    the commit server, against the plain per-table path timed per height."""
    import numpy as np

    hip_rt = mod.hip_rt
    stream = hip_rt.Stream()
    jobs = serve_stream(args, args.jobs, seed=2024)
    cells = sum(j["m"] * j["w"] for j in jobs)
    tuned = autotune(args, lib, [(j["n"], j["w"], j["m"]) for j in jobs])
    # warm-up: one job of every height through both paths (coset tables, code objects)
    warm = [dict(jobs[0], n=1 << lg, m=args.expansion << lg, seed=7, t=0.0, i=-1 - lg,
                 leaves=jobs[0]["leaves"] % np.uint64(args.expansion << lg))
            for lg in range(args.serve_min_log, args.serve_max_log + 1)]
    serve_run(args, lib, warm)
    # plain per-table path: each height timed twice, the faster counted
    plain_s = {}
    for wj in warm:
        best = None
        for _ in range(2):
            dt = plain_job(lib, hip_rt, wj, stream)[0]
            best = dt if best is None else min(best, dt)
        plain_s[wj["n"]] = best
    plain_total = sum(plain_s[j["n"]] for j in jobs)
    print(f"serve: plain per-height times {plain_s}", file=sys.stderr, flush=True)
    host_s = {wj["n"]: plain_host_seconds(lib, wj["n"], wj["w"], wj["m"], 11) for wj in warm}
    host_total = sum(host_s[j["n"]] for j in jobs)
    print(f"serve: host drop-in per-height times {host_s}", file=sys.stderr, flush=True)

    # bit-exactness sample: one job of every height plus the first 8, compared with the
    # plain path on the same seed (root, opened rows, authentication paths)
    sample = sorted({j["i"] for j in jobs[:8]} |
                    {next(j["i"] for j in jobs if j["n"] == 1 << lg)
                     for lg in range(args.serve_min_log, args.serve_max_log + 1)
                     if any(j["n"] == 1 << lg for j in jobs)})
    t0 = time.monotonic()
    res, kept = serve_run(args, lib, jobs, keep=set(sample))
    t_end = time.monotonic()
    lat = [r[2] - r[0] for r in res]
    ans = [r[3] - r[0] for r in res]
    qwait = [r[1] - r[0] for r in res]
    first = min(r[0] for r in res)
    span = max(r[3] for r in res) - first
    mism = 0
    for i in sample:
        _, root, rows, paths = plain_job(lib, hip_rt, jobs[i], stream, want_rows=True)
        sr, srows, spaths = kept[i]
        if sr != root or not np.array_equal(spaths, paths) or (
                rows is not None and not np.array_equal(srows, rows)):
            mism += 1

    # capacity: the same jobs all submitted at once (no arrival gaps)
    sat_args = argparse.Namespace(**{**vars(args), "rate": 0.0})
    sat_jobs = [dict(j, t=0.0) for j in jobs[: args.saturated_jobs]]
    sres, _ = serve_run(sat_args, lib, sat_jobs)
    sat_cells = sum(j["m"] * j["w"] for j in sat_jobs)
    sat_span = max(r[3] for r in sres) - min(r[0] for r in sres)
    sat_plain = sum(plain_s[j["n"]] for j in sat_jobs)
    sat_host = sum(host_s[j["n"]] for j in sat_jobs)
    return {
        "metric": "served LDE+commit jobs: cells/s and p50/p99 job latency",
        "value": cells / span, "unit": "cells/s", "n_gpus": 1, "higher_is_better": True,
        "dtype": "u64", "data": "synthetic (per-job seeds, tvh_fill_uniform_dev on the device)",
        "config": {"workload": f"{len(jobs)} commit jobs, Poisson arrivals at {args.rate} jobs/s, "
                               f"log2 trace height uniform {args.serve_min_log}..{args.serve_max_log}, "
                               f"width {args.width}, LDE x{args.expansion}, Tip5 rows, Merkle, "
                               f"{args.queries} queries (rows + authentication paths) per job",
                   "jobs": len(jobs), "inflight_slots": args.inflight, "hbm_budget_gib": args.hbm_gib,
                   "total_cells": cells},
        "latency_s": {"commit_p50": pct(lat, 50), "commit_p99": pct(lat, 99),
                      "answered_p50": pct(ans, 50), "answered_p99": pct(ans, 99),
                      "admission_wait_p50": pct(qwait, 50), "admission_wait_p99": pct(qwait, 99)},
        "stream_span_s": span, "offered_cells_per_s": cells / max(1e-9, jobs[-1]["t"]),
        "saturated": {"jobs": len(sat_jobs), "cells_per_s": sat_cells / sat_span,
                      "plain_cells_per_s": sat_cells / sat_plain,
                      "speedup_vs_plain": sat_plain / sat_span,
                      "host_dropin_cells_per_s": sat_cells / sat_host,
                      "speedup_vs_host_dropin": sat_host / sat_span},
        "plain_per_table": {"cells_per_s": cells / plain_total, "total_s": plain_total,
                            "s_per_job_by_log_height": {str(k.bit_length() - 1): v
                                                        for k, v in plain_s.items()}},
        "host_dropin_per_table": {"cells_per_s": cells / host_total, "total_s": host_total,
                                  "s_per_job_by_log_height": {str(k.bit_length() - 1): v
                                                              for k, v in host_s.items()}},
        "checked_jobs": len(sample), "mismatches_vs_plain": mism, "lde_forms": tuned,
        "wall_s": t_end - t0,
    }


def autotune(args, lib, shapes):
    """Runs the library's LDE autotuner once per (trace_len, target_len) with the
    width given; returns {log2 trace_len: {"variant": chosen, "ms": {variant: ms}}}.
    With --lde-variant the given variant is set for every shape instead."""
    out = {}
    if args.lde_variant:
        name = sys.modules["tvh"].lde_variant_name
        for n, w, m in sorted(set(shapes)):
            lib.lde_set_variant(n.bit_length() - 1, (m // n).bit_length() - 1, args.lde_variant)
            out[n.bit_length() - 1] = {"variant": name(args.lde_variant), "forced": True}
        return out
    if args.no_autotune:
        return out
    for n, w, m in sorted(set(shapes)):
        lg = n.bit_length() - 1
        if lg in out:
            continue
        chosen, ms = lib.lde_autotune(w, n, m)
        name = sys.modules["tvh"].lde_variant_name
        out[lg] = {"variant": name(chosen), "ms": {name(k): round(v, 3) for k, v in ms.items()}}
    return out


def lde_kernels(lib, n, m) -> str:
    """Names the kernels of the LDE variant in use for (n, m)."""
    v = lib.lde_get_variant(n.bit_length() - 1, (m // n).bit_length() - 1)
    tile = 14 - ((v >> 2) & 3)
    row = "k_lde_rowinv + k_lde_expand" if v & 3 == 2 else "k_lde_step2"
    if v & 16:
        row = "k_lde_rowinv<13> + k_lde_expand<13>"
    return f"LDE passes (k_step1<{tile}> + {row} + k_lde_step3<{tile}>)"


def workload_name(args, n, w, E) -> str:
    lg, lm = args.log_n, args.log_n + E.bit_length() - 1
    if args.mode == "aux":
        return (f"aux-table commit 2^{lg} x {w // 3} XFE columns ({w} limbs), coset LDE x{E} "
                f"(offset 7) to 2^{lm} rows, Tip5 row hashing, level-order Merkle tree")
    if args.mode == "frugal":
        return (f"frugal main-table commit 2^{lg} x {w} streamed in {args.chunk}-column chunks "
                f"(LDE x{E} to 2^{lm} rows, resident per-row Tip5 sponge state), Merkle tree")
    return (f"main-table commit 2^{lg} x {w}, coset LDE x{E} (offset 7) to 2^{lm} rows, Tip5 "
            f"row hashing, level-order Merkle tree")


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--log-n", type=int, default=20)
    ap.add_argument("--width", type=int, default=379)
    ap.add_argument("--expansion", type=int, default=4)
    ap.add_argument("--mode", choices=["main", "aux", "frugal", "batch", "serve"], default="main",
                    help="main: 2^20 x 379 main table (default); aux: 2^20 x 88 XFE columns "
                         "(264 limbs); frugal: 2^21 x 379 streamed in 32-column chunks; "
                         "batch: 96 tables of mixed heights/widths, batched vs per-table; "
                         "serve: Poisson job stream through the commit server")
    ap.add_argument("--tables", type=int, default=96, help="tables in batch mode")
    ap.add_argument("--wave-gib", type=float, default=200.0,
                    help="HBM budget per wave of tables in batch mode")
    ap.add_argument("--chunk", type=int, default=32, help="columns per chunk in frugal mode")
    ap.add_argument("--jobs", type=int, default=2000, help="jobs in serve mode")
    ap.add_argument("--rate", type=float, default=40.0, help="mean arrivals/s in serve mode")
    ap.add_argument("--queries", type=int, default=80, help="leaf queries per served job")
    ap.add_argument("--serve-min-log", type=int, default=14)
    ap.add_argument("--serve-max-log", type=int, default=21)
    ap.add_argument("--inflight", type=int, default=4, help="server stream slots")
    ap.add_argument("--hbm-gib", type=float, default=0.0,
                    help="server HBM budget (0: free memory - 16 GiB)")
    ap.add_argument("--saturated-jobs", type=int, default=400,
                    help="jobs of the capacity run (all submitted at once)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-host-baseline", action="store_true",
                    help="batch/serve modes: skip timing the host-buffer drop-in path")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default LDE kernel forms instead of timing each first")
    ap.add_argument("--lde-variant", type=int, default=0,
                    help="force this LDE variant code (form | tile << 2) for the workload's shape "
                         "instead of autotuning (profiling runs)")
    ap.add_argument("--calibrate-copy-gib", type=float, default=0.0,
                    help="run one device-to-device copy of this size first (HBM counter "
                         "calibration: its FETCH_SIZE/WRITE_SIZE are known bytes)")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}, WORLD_SIZE "
              f"{world}); the commit path is single-GPU in this build.", file=sys.stderr)
        return 2

    mod = _tvh()
    hip_rt = mod.hip_rt
    lib = mod.load()
    rc = lib.tvh_init(0)
    if rc:
        raise RuntimeError(f"tvh_init failed: {lib.status_string(rc)}: {lib.last_error()}")
    if args.mode == "batch":
        print(json.dumps(run_batch(args, mod, lib)))
        return 0
    if args.mode == "serve":
        print(json.dumps(run_serve(args, mod, lib)))
        return 0

    if args.mode == "aux":
        args.width = 3 * 88
    if args.mode == "frugal" and args.log_n == 20:
        args.log_n = 21
    n, w, E = 1 << args.log_n, args.width, args.expansion
    m = E * n
    host_trace = random_field(n * w, seed=1)
    trace = hip_rt.to_device(host_trace)
    del host_trace
    frugal = args.mode == "frugal"
    if frugal:
        cw = hip_rt.DeviceBuffer(lib.frugal_workspace_elems(w, n, m, args.chunk))
    else:
        cw = hip_rt.DeviceBuffer(m * w)
    tree = hip_rt.DeviceBuffer(5 * (2 * m - 1))
    stream = hip_rt.Stream()
    if args.calibrate_copy_gib > 0:
        nb = int(args.calibrate_copy_gib * (1 << 30)) // 8
        a, b = hip_rt.DeviceBuffer(nb), hip_rt.DeviceBuffer(nb)
        hip_rt.check(hip_rt.hip().hipMemcpy(b.ptr, a.ptr, 8 * nb, hip_rt.D2D), "calibration copy")
        hip_rt.synchronize()
        del a, b

    def step(ev=None):
        if frugal:  # LDE chunks and sponge absorption interleave: one phase
            if ev:
                ev[0].record(stream)
            lib.commit_frugal(trace, w, n, None, 0, 1, 7, m, args.chunk, cw, tree,
                              mod.TREE_LEVEL_ORDER, stream)
            if ev:
                for e in ev[1:]:
                    e.record(stream)
            return
        if ev:
            ev[0].record(stream)
        if args.mode == "aux":  # limb-major XFE columns, BFE-limb LDE with the XFE rule
            lib.lde_xfe_limbs(trace, w // 3, n, None, 0, 1, 7, m, cw, 0, stream)
        else:
            lib.lde_batch(trace, w, n, None, 0, 1, 7, m, cw, stream)
        if ev:
            ev[1].record(stream)
        lib.hash_bfield_rows(cw, m, w, tree, stream)  # leaves = level 0 of the tree
        if ev:
            ev[2].record(stream)
        lib.merkle_tree(tree, tree, m, mod.TREE_LEVEL_ORDER, stream)
        if ev:
            ev[3].record(stream)

    tuned = autotune(args, lib, [(n, w, m)])
    for _ in range(args.warmup):
        step()
    stream.synchronize()

    evs = [[hip_rt.Event() for _ in range(4)] for _ in range(args.steps)]
    hip_rt.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step(evs[i])
    stream.synchronize()
    hip_rt.synchronize()
    wall = time.perf_counter() - t0

    lde_ms = sum(e[0].elapsed_ms(e[1]) for e in evs) / args.steps
    hash_ms = sum(e[1].elapsed_ms(e[2]) for e in evs) / args.steps
    merkle_ms = sum(e[2].elapsed_ms(e[3]) for e in evs) / args.steps
    step_ms = wall * 1000 / args.steps

    cells = m * w
    lde_bytes = 8 * n * w + 8 * m * w
    perms = m * ((w + 1 + 9) // 10) + (m - 1)
    achieved = lde_bytes / (lde_ms / 1000)
    out = {
        "metric": METRIC,
        "value": cells / (step_ms / 1000),
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": step_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (uniform Goldilocks from numpy PCG64, uploaded before timing)",
        "config": {"workload": workload_name(args, n, w, E),
                   "trace_rows": n, "columns": w, "codeword_rows": m},
        "phases_ms": {"lde": lde_ms, "hash_rows": hash_ms, "merkle": merkle_ms},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK,
                     "traffic": measured_traffic(args.log_n, w, E),
                     "kernel": lde_kernels(lib, n, m),
                     "algorithmic_bytes": lde_bytes},
        "lde_forms": tuned,
        "tip5": {"permutations": perms,
                 "permutations_per_s": perms / ((hash_ms + merkle_ms) / 1000)},
    }
    if frugal:
        out["phases_ms"] = {"commit_streamed": lde_ms}
        out["roofline"] = None
        out["tip5"]["permutations_per_s"] = perms / (step_ms / 1000)
    if not args.no_cpu_baseline and args.mode == "main":
        out["cpu_baseline"] = cpu_baseline(13, w, E)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
