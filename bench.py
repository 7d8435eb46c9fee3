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


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--log-n", type=int, default=20)
    ap.add_argument("--width", type=int, default=379)
    ap.add_argument("--expansion", type=int, default=4)
    ap.add_argument("--no-cpu-baseline", action="store_true")
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

    n, w, E = 1 << args.log_n, args.width, args.expansion
    m = E * n
    host_trace = random_field(n * w, seed=1)
    trace = hip_rt.to_device(host_trace)
    del host_trace
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
        if ev:
            ev[0].record(stream)
        lib.lde_batch(trace, w, n, None, 0, 1, 7, m, cw, stream)
        if ev:
            ev[1].record(stream)
        lib.hash_bfield_rows(cw, m, w, tree, stream)  # leaves = level 0 of the tree
        if ev:
            ev[2].record(stream)
        lib.merkle_tree(tree, tree, m, mod.TREE_LEVEL_ORDER, stream)
        if ev:
            ev[3].record(stream)

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
        "config": {"workload": f"main-table commit 2^{args.log_n} x {w}, coset LDE x{E} "
                               f"(offset 7) to 2^{args.log_n + E.bit_length() - 1} rows, Tip5 "
                               f"row hashing, level-order Merkle tree",
                   "trace_rows": n, "columns": w, "codeword_rows": m},
        "phases_ms": {"lde": lde_ms, "hash_rows": hash_ms, "merkle": merkle_ms},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK,
                     "traffic": measured_traffic(args.log_n, w, E),
                     "kernel": "LDE passes (k_step1 + k_lde_step2 + k_lde_step3)",
                     "algorithmic_bytes": lde_bytes},
        "tip5": {"permutations": perms,
                 "permutations_per_s": perms / ((hash_ms + merkle_ms) / 1000)},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(13, w, E)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
