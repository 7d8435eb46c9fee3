"""Benchmark of the table-commitment hot path on one MI355X.

Headline workload: the main-table commit at 2^20 rows x 379 Goldilocks
columns -- coset LDE x4 (offset 7) to 2^22 rows, Tip5 hash of every row,
Merkle tree over the 2^22 row digests.  One step = LDE + row hashing + Merkle
build with the trace resident in HBM and every output left on the device.  The
trace is the seeded table of the workload definitions (SplitMix64 from seed 1,
row-major draws, draws >= p rejected), drawn on the device by
tvh_fill_table_dev before timing, so the timed table is the one the oracle's
seeded draw reproduces.

Prints one JSON line.  `value` is output codeword cells per second (rows_out *
columns / step time).  `roofline` is the LDE (the NTT passes): algorithmic
bytes 8*N*W + 8*M*W per launch over the LDE's event-timed duration against
8 TB/s, with the HBM bytes the rocprof counters measure for it (profiles/).
`tip5` is the hashing phase: permutations/s and the VALU issue fraction of the
row-hash kernel.  `batch` and `serve` are the two systems workloads (96 mixed
tables through one batch plan; a Poisson stream of 2000 jobs through the commit
server) with their latencies and their ratios to the plain per-table path.
`cpu_baseline` runs the reference's own CPU code (oracle/_ref, built from the
reference sources) on bounded samples on the host.
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import os
import re
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "triton-vm-prover_amd")
METRIC = "LDE+Tip5 Merkle commit cells/s & HBM GB/s at 2^20x379; p50/p99 job latency"
HBM_PEAK = 8.0e12
# MI355X: 256 CUs x 4 SIMDs at up to 2.4 GHz
SIMDS, CLOCK_HZ = 1024, 2.4e9


def _load(name, path):
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def profile_value(fname: str, key: str):
    """A figure recorded in a committed profile summary (profiles/<fname>), as
    `key: value` on its own line, or None."""
    path = os.path.join(ROOT, "profiles", fname)
    if not os.path.exists(path):
        return None
    m = re.search(rf"^{re.escape(key)}:\s*([0-9.eE+-]+)", open(path).read(), re.M)
    return float(m.group(1)) if m else None


def pct(xs, q):
    import numpy as np

    return float(np.percentile(np.asarray(xs, dtype=float), q)) if len(xs) else None


# ---------------------------------------------------------------- CPU baseline
def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(width: int, expansion: int) -> dict:
    """The reference's own CPU commit (oracle/_ref, compiled from the reference
    sources): the parallel form (OpenMP over columns and rows, Merkle levels) on
    a 2^17-row sample with every host thread, and the single-thread form on a
    2^13-row sample.  Fails when oracle/_ref is not built."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Oracle, Reference

    if not Reference.available():
        raise RuntimeError("cpu baseline: oracle/_ref/libtvm_ref.so is not built "
                           "(make -C oracle ref, with the reference sources)")
    orc, ref = Oracle(), Reference()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1
    out = {}
    for lg, th in ((17, threads), (13, 1)):
        n, m = 1 << lg, expansion << lg
        trace = orc.table_rowmajor_draw(1, n, width)
        ref.set_threads(th)
        t0 = time.perf_counter()
        if th > 1:
            ref.commit_parallel(trace, n, width, m)
        else:
            ref.commit(trace, n, width, m)
        out[th] = (m * width / (time.perf_counter() - t0), lg, time.perf_counter() - t0)
    v, lg, dt = out[threads]
    v1, lg1, dt1 = out[1]
    return {"value": v, "unit": "cells/s", "cores": threads, "kind": "reference",
            "sample": f"reference CPU commit of a 2^{lg} x {width} trace (LDE x{expansion}, Tip5 "
                      f"rows, Merkle), OpenMP over columns and rows on {threads} threads, "
                      f"{dt:.2f} s",
            "single_thread": {"value": v1, "cores": 1,
                              "sample": f"2^{lg1} x {width}, 1 thread, {dt1:.2f} s"},
            "cpu_model": cpu_model(), "nproc": os.cpu_count(),
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


# ---------------------------------------------------------------- main table
def width_class(w):
    """The library's width classes for tuned LDE variants."""
    return "<64" if w < 64 else "64..255" if w < 256 else ">=256"


def autotune(lib, shapes, tvh, force=0, skip=False):
    """The library's LDE autotuner once per (trace_len, target_len, width class),
    timed on the first width of the class; returns {"log2 trace_len/width
    class": {"variant", "ms", ...}}."""
    out = {}
    for n, w, m in sorted(set(shapes)):
        lg = n.bit_length() - 1
        key = f"{lg}/{width_class(w)}"
        if key in out:
            continue
        if force:
            lib.lde_set_variant(lg, (m // n).bit_length() - 1, force)
            out[key] = {"variant": tvh.lde_variant_name(force), "forced": True}
        elif not skip:
            chosen, ms = lib.lde_autotune(w, n, m)
            ran = {k: v for k, v in ms.items() if v >= 0}  # -1: the variant cannot run the shape
            slowest = max(ran, key=ran.get)
            out[key] = {"variant": tvh.lde_variant_name(chosen), "ms": round(ms[chosen], 3),
                       "slowest": tvh.lde_variant_name(slowest), "slowest_ms": round(ran[slowest], 3),
                       "variants_timed": len(ran)}
    return out


def lde_kernels(lib, n, m, w) -> str:
    v = lib.lde_get_variant(n.bit_length() - 1, (m // n).bit_length() - 1, cols=w)
    tile = 14 - ((v >> 2) & 3)
    if v & 3 == 1:
        row = "k_lde_step2"
    else:
        rt = 13 if v & 16 else 12 if v & 64 else 14
        it = 13 if v & 32 else rt
        row = f"k_lde_rowinv<{it}> + k_lde_expand<{rt}>"
    return f"LDE passes (k_step1<{tile}> + {row} + k_lde_step3<{tile}>)"


def run_main(args, tvh, lib, hip_rt):
    n, w, E = 1 << args.log_n, args.width, args.expansion
    m = E * n
    if args.mode == "aux":
        w = 3 * 88
    trace = hip_rt.DeviceBuffer(n * w)
    lib.fill_table(trace, n, w, {"main": 1, "aux": 2, "frugal": 3}[args.mode])
    hip_rt.synchronize()  # filled on the null stream; the steps run on a non-blocking one
    frugal = args.mode == "frugal"
    cw = hip_rt.DeviceBuffer(lib.frugal_workspace_elems(w, n, m, args.chunk) if frugal else m * w)
    tree = hip_rt.DeviceBuffer(5 * (2 * m - 1))
    stream = hip_rt.Stream()
    if args.calibrate_copy_gib > 0:
        nb = int(args.calibrate_copy_gib * (1 << 30)) // 8
        a, b = hip_rt.DeviceBuffer(nb), hip_rt.DeviceBuffer(nb)
        hip_rt.check(hip_rt.hip().hipMemcpy(b.ptr, a.ptr, 8 * nb, hip_rt.D2D), "calibration copy")
        hip_rt.synchronize()
        del a, b
    lib.lde_prepare(n, m, 1, 7)

    def step(ev=None):
        if ev:
            ev[0].record(stream)
        if frugal:  # LDE chunks and sponge absorption interleave: one phase
            lib.commit_frugal(trace, w, n, None, 0, 1, 7, m, args.chunk, cw, tree, 0, stream)
            if ev:
                for e in ev[1:]:
                    e.record(stream)
            return
        if args.mode == "aux":  # limb-major XFE columns, BFE-limb LDE with the XFE rule
            lib.lde_xfe_limbs(trace, w // 3, n, None, 0, 1, 7, m, cw, 0, stream)
        else:
            lib.lde_batch(trace, w, n, None, 0, 1, 7, m, cw, stream)
        if ev:
            ev[1].record(stream)
        lib.hash_bfield_rows(cw, m, w, tree, stream)  # leaves = level 0 of the tree
        if ev:
            ev[2].record(stream)
        lib.merkle_tree(tree, tree, m, 0, stream)
        if ev:
            ev[3].record(stream)

    tuned = autotune(lib, [(n, w, m)], tvh, args.lde_variant, args.no_autotune)
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
    root = [int(v) for v in tree.download(5, 5 * (2 * m - 2))]
    del trace, cw, tree

    cells = m * w
    lde_bytes = 8 * n * w + 8 * m * w
    perms_rows = m * ((w + 1 + 9) // 10)
    perms = perms_rows + (m - 1)
    achieved = lde_bytes / (lde_ms / 1000)
    tag = f"2p{args.log_n}x{w}"
    counters = f"commit_{tag}_hbm_counters.md"
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
        "data": f"synthetic (seeded SplitMix64 table, seed {dict(main=1, aux=2, frugal=3)[args.mode]}, "
                "row-major draws with rejection, drawn on the device before timing)",
        "config": {"workload": workload_name(args, n, w, E), "trace_rows": n, "columns": w,
                   "codeword_rows": m},
        "phases_ms": {"lde": lde_ms, "hash_rows": hash_ms, "merkle": merkle_ms},
        "root": root,
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK,
                     "traffic": profile_value(counters, "traffic_bytes_per_launch"),
                     "kernel": lde_kernels(lib, n, m, w), "algorithmic_bytes": lde_bytes},
        "lde_forms": tuned,
    }
    # VALU issue fractions against the chip's peak (1024 SIMDs x 2.4 GHz): the
    # busy SIMD cycles per unit of work that the rocprof counters measured for
    # these kernels (profiles/), over this run's event-timed phase durations
    lde_cyc = profile_value(counters, "lde_valu_busy_cycles_per_cell")
    if lde_cyc and not frugal:
        out["roofline"]["valu"] = {
            "busy_cycles_per_cell": lde_cyc,
            "frac": lde_cyc * cells / (lde_ms / 1000) / (SIMDS * CLOCK_HZ),
            "note": "the LDE kernels are bound by VALU issue and its dependency latency at 4 "
                    "waves per SIMD, not by HBM: frac is their share of the chip's peak issue, "
                    "busy cycles from counted instructions at 2/4/8 SIMD cycles per 32-bit/"
                    "f64/64-bit-integer wave op (a model; an upper estimate)"}
    cyc = profile_value(counters, "valu_busy_cycles_per_permutation")
    tip5 = {"permutations": perms, "permutations_per_s": perms / ((hash_ms + merkle_ms) / 1000),
            "row_hash_permutations_per_s": perms_rows / (hash_ms / 1000)}
    if cyc:
        busy = cyc * perms_rows / (hash_ms / 1000)  # SIMD cycles/s the kernel keeps busy
        tip5.update({"valu_frac": busy / (SIMDS * CLOCK_HZ),
                     "valu_busy_cycles_per_permutation": cyc,
                     "peak_simd_cycles_per_s": SIMDS * CLOCK_HZ,
                     "valu_note": "busy cycles from counted instructions at 2/4/8 SIMD cycles per "
                                  "32-bit/f64/64-bit-integer wave op (a model; an upper estimate)"})
    out["tip5"] = tip5
    if frugal:
        out["phases_ms"] = {"commit_streamed": lde_ms}
        out["roofline"] = None
        out["tip5"]["permutations_per_s"] = perms / (step_ms / 1000)
    return out


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


# ---------------------------------------------------------------- two-phase cosets
def run_cosets(args, tvh, lib, hip_rt, reps=3):
    """The prover's frugal main-table form (gpu_stark.cu:2356-2420): trace
    coefficients once (phase 1), then per coset of the FRI domain 8N (98
    randomizers per column, prover form): evaluate the coset, hash its rows,
    scatter the digests; then the Merkle tree over all 8N leaves.  The
    reference's source comments quote about 487 ms for its phase 1 and about
    20 ms per coset evaluation (gpu_stark.cu:2368,2382), hardware unstated."""
    P = 0xFFFFFFFF00000001
    n, w, K, k = 1 << args.log_n, args.width, 8, 98
    m = K * n
    tr = hip_rt.DeviceBuffer(n * w)
    lib.fill_table(tr, n, w, 1)
    rand = hip_rt.DeviceBuffer(w * k)
    lib.fill_uniform(rand, w * k, 98)
    hip_rt.synchronize()  # filled on the null stream; the commits run on a non-blocking one
    coeffs = hip_rt.DeviceBuffer(n * w)
    work = hip_rt.DeviceBuffer(n * w)
    dig = hip_rt.DeviceBuffer(5 * n)
    tree = hip_rt.DeviceBuffer(5 * (2 * m - 1))
    st = hip_rt.Stream()
    wm = pow(7, (P - 1) // m, P)
    gs = [7 * pow(wm, kk, P) % P for kk in range(K)]

    def commit(ev=None):
        if ev:
            ev[0].record(st)
        lib.trace_coefficients(tr, w, n, 1, coeffs, st)
        if ev:
            ev[1].record(st)
        for kk in range(K):
            lib.evaluate_coset(coeffs, w, n, rand, k, 1, gs[kk], work, st)
            if ev and kk == 0:
                ev[2].record(st)
            lib.hash_bfield_rows(work, n, w, dig, st)
            lib.scatter_digests(dig, tree, n, kk, K, st)
        if ev:
            ev[3].record(st)
        lib.merkle_tree(tree, tree, m, 0, st)
        if ev:
            ev[4].record(st)

    commit()
    st.synchronize()
    evs = [[hip_rt.Event() for _ in range(5)] for _ in range(reps)]
    for i in range(reps):
        commit(evs[i])
    st.synchronize()
    avg = lambda a, b: sum(e[a].elapsed_ms(e[b]) for e in evs) / reps
    total = avg(0, 4)
    root = [int(v) for v in tree.download(5, 5 * (2 * m - 2))]
    return {"workload": f"frugal two-phase main-table commit 2^{args.log_n} x {w}, {k} randomizers "
                        f"per column, FRI domain {K}N = 2^{(m).bit_length() - 1} rows in {K} cosets",
            "phase1_trace_coefficients_ms": avg(0, 1),
            "coset_evaluate_ms": avg(1, 2),
            "coset_evaluate_hash_scatter_ms": (avg(1, 3)) / K,
            "merkle_ms": avg(3, 4), "total_ms": total,
            "cells_per_s": m * w / (total / 1000), "root": root,
            "reference_comments_ms": {"phase1": 487, "per_coset_evaluation": 20}}


# ---------------------------------------------------------------- batched tables
class Pinned:
    """Page-locked host memory of u64 (hipHostMalloc)."""

    def __init__(self, hip_rt, count):
        import ctypes as C

        import numpy as np

        self.hip_rt = hip_rt
        p = C.c_void_p()
        hip_rt.check(hip_rt.hip().hipHostMalloc(C.byref(p), max(8, 8 * count), 0), "hipHostMalloc")
        self.ptr = p.value
        self.count = count
        self.array = np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_uint64)), shape=(max(1, count),))

    def data_ptr(self) -> int:
        return self.ptr

    def __del__(self):
        try:
            self.hip_rt.hip().hipHostFree(self.ptr)
        except Exception:
            pass


class HostArray:
    """A host trace in ordinary (pageable) memory: the boxes this runs on let a
    process page-lock only a few GiB, so host-resident tables live where a
    prover's own tables do, and uploads go through the runtime's staging."""

    def __init__(self, a):
        self.array = a
        self.ptr = a.ctypes.data

    def data_ptr(self) -> int:
        return self.ptr


class HostTraces:
    """One column-major host trace per log2 height at a width of `width` columns,
    drawn on the device from a fixed seed per height; a table of that height and
    any width <= `width` reads its first columns (a prefix)."""

    def __init__(self, lib, hip_rt, logs, width, stream):
        self.t = {}
        for lg in sorted(logs):
            n = 1 << lg
            tmp = hip_rt.DeviceBuffer(n * width)
            lib.fill_uniform(tmp, n * width, 7000 + lg, stream)
            stream.synchronize()
            self.t[lg] = HostArray(tmp.download())
            del tmp

    def __getitem__(self, lg):
        return self.t[lg]


def memcpy_async(hip_rt, dst, src, nbytes, kind, stream):
    hip_rt.check(hip_rt.hip().hipMemcpyAsync(dst, src, nbytes, kind, stream.ptr), "hipMemcpyAsync")


def run_batch(args, tvh, lib, hip_rt, W, hosts):
    """96 mixed tables: (1) device-resident -- per-table tvh_commit_dev vs one
    batch plan per wave replayed as a hipGraph; (2) host-resident traces -- the
    plain per-table drop-in (upload with the runtime's pageable copy, commit,
    download digests and root, one table at a time; and the same with the
    staged upload) vs the batched form (staged uploads of wave k+1 on a copy
    stream overlapped with wave k's graph, digests downloaded as each wave
    finishes)."""
    import numpy as np

    E = args.expansion
    shapes = W.batch_shapes(args.tables)
    tb = lambda lg, w: 8 * (1 << lg) * w + 8 * E * (1 << lg) * w + 40 * (2 * E * (1 << lg) - 1)
    order = sorted(range(len(shapes)), key=lambda i: -tb(*shapes[i]))
    budget = int(args.wave_gib * (1 << 30))
    waves, cur, used = [], [], 0
    for i in order:
        b = tb(*shapes[i])
        if cur and used + b > budget:
            waves.append(cur)
            cur, used = [], 0
        cur.append(i)
        used += b
    waves.append(cur)
    tuned = autotune(lib, [(1 << lg, w, E << lg) for lg, w in shapes], tvh, args.lde_variant,
                     args.no_autotune)
    for lg, w in set(shapes):
        lib.lde_prepare(1 << lg, E << lg, 1, 7)
    cells = sum((E << lg) * w for lg, w in shapes)
    trace_bytes = sum(8 * (1 << lg) * w for lg, w in shapes)
    comp, copy = hip_rt.Stream(), hip_rt.Stream()

    # two buffer sets (ping-pong) of the largest wave, carved per wave
    def layout(wave):
        offs, o = [], 0
        for i in wave:
            lg, w = shapes[i]
            n, m = 1 << lg, E << lg
            parts = (n * w, m * w, 5 * (2 * m - 1))
            offs.append((o, o + parts[0], o + parts[0] + parts[1]))
            o += sum(parts) + 32
        return offs, o

    lay = [layout(wv) for wv in waves]
    set_elems = max(o for _, o in lay)
    sets = [hip_rt.DeviceBuffer(set_elems), hip_rt.DeviceBuffer(set_elems)]
    # host traces: one per height (tables of a height share it)
    host = {i: hosts[lg] for i, (lg, w) in enumerate(shapes)}
    # digests come back through one page-locked staging buffer of the largest table
    dig_stage = Pinned(hip_rt, 5 * (E << max(lg for lg, _ in shapes)))
    plans, descs = [], []
    for k, wv in enumerate(waves):
        buf = sets[k % 2]
        ds = []
        for i, (ot, oc, otr) in zip(wv, lay[k][0]):
            lg, w = shapes[i]
            n, m = 1 << lg, E << lg
            ds.append((i, n, m, w, buf.ptr + 8 * ot, buf.ptr + 8 * oc, buf.ptr + 8 * otr))
        descs.append(ds)
        plans.append(lib.batch_plan([lib.table_desc(t, w, n, m, c, tr)
                                     for (_, n, m, w, t, c, tr) in ds]))

    def upload(k, stream, staged=False):
        for (i, n, m, w, t, c, tr) in descs[k]:
            if staged:  # the library's page-locked staging ring (tvh_upload_dev)
                lib.upload(t, host[i].ptr, n * w, stream)
            else:  # the runtime's copy from pageable memory, as the prover does
                memcpy_async(hip_rt, t, host[i].ptr, 8 * n * w, hip_rt.H2D, stream)

    def download(k, stream):
        for (i, n, m, w, t, c, tr) in descs[k]:
            memcpy_async(hip_rt, dig_stage.ptr, tr, 40 * m, hip_rt.D2H, stream)

    # warm-up: every plan's graph captured once, every shape committed once
    for k in range(len(waves)):
        upload(k, comp)
        lib.batch_launch_graph(plans[k], comp)
        for (i, n, m, w, t, c, tr) in descs[k]:
            lib.commit(t, w, n, None, 0, 1, 7, m, c, tr, 0, comp)
    comp.synchronize()
    print("bench: batch warmed up", file=sys.stderr, flush=True)

    ev0, ev1 = hip_rt.Event(), hip_rt.Event()
    root_host = Pinned(hip_rt, 5)

    def root_of(tr, m):
        memcpy_async(hip_rt, root_host.ptr, tr + 8 * 5 * (2 * m - 2), 40, hip_rt.D2H, comp)
        comp.synchronize()
        return root_host.array[:5].copy()

    # (1) device-resident: per-table commits vs graph replays (traces already up)
    t_plain = t_graph = 0.0
    mism = 0
    for k in range(len(waves)):
        upload(k, comp)
        comp.synchronize()
        ev0.record(comp)
        for (i, n, m, w, t, c, tr) in descs[k]:
            lib.commit(t, w, n, None, 0, 1, 7, m, c, tr, 0, comp)
        ev1.record(comp)
        comp.synchronize()
        t_plain += ev0.elapsed_ms(ev1)
        roots = {i: root_of(tr, m) for (i, n, m, w, t, c, tr) in descs[k]}
        ev0.record(comp)
        lib.batch_launch_graph(plans[k], comp)
        ev1.record(comp)
        comp.synchronize()
        t_graph += ev0.elapsed_ms(ev1)
        mism += sum(not np.array_equal(root_of(tr, m), roots[i])
                    for (i, n, m, w, t, c, tr) in descs[k])
    # (2) host-resident traces
    # plain drop-in: per table, upload, commit, download digests and root, wait
    hip_rt.synchronize()
    t0 = time.perf_counter()
    want = {}
    for k in range(len(waves)):
        for (i, n, m, w, t, c, tr) in descs[k]:
            memcpy_async(hip_rt, t, host[i].ptr, 8 * n * w, hip_rt.H2D, comp)
            lib.commit(t, w, n, None, 0, 1, 7, m, c, tr, 0, comp)
            memcpy_async(hip_rt, dig_stage.ptr, tr, 40 * m, hip_rt.D2H, comp)
            want[i] = root_of(tr, m)  # waits for the table
    t_host_plain = time.perf_counter() - t0
    # the same per-table sequence with the staged upload
    hip_rt.synchronize()
    t0 = time.perf_counter()
    for k in range(len(waves)):
        for (i, n, m, w, t, c, tr) in descs[k]:
            lib.upload(t, host[i].ptr, n * w, comp)
            lib.commit(t, w, n, None, 0, 1, 7, m, c, tr, 0, comp)
            memcpy_async(hip_rt, dig_stage.ptr, tr, 40 * m, hip_rt.D2H, comp)
            root_of(tr, m)
    t_host_plain_staged = time.perf_counter() - t0
    # batched: a copy thread uploads wave k+1 on the copy stream while the
    # compute stream runs wave k's graph (uploads from pageable memory return
    # once the runtime has taken the data, so they run on a thread of their own)
    done = [hip_rt.Event() for _ in waves]
    up = [hip_rt.Event() for _ in waves]
    up_rec = [threading.Event() for _ in waves]
    done_rec = [threading.Event() for _ in waves]
    wait_ev = lambda st, ev: hip_rt.check(hip_rt.hip().hipStreamWaitEvent(st.ptr, ev.ptr, 0), "wait")

    def uploader():
        for k in range(len(waves)):
            if k >= 2:  # wave k reuses the buffer set of wave k-2
                done_rec[k - 2].wait()
                wait_ev(copy, done[k - 2])
            upload(k, copy, staged=True)
            up[k].record(copy)
            up_rec[k].set()

    hip_rt.synchronize()
    t0 = time.perf_counter()
    th = threading.Thread(target=uploader)
    th.start()
    for k in range(len(waves)):
        up_rec[k].wait()
        wait_ev(comp, up[k])
        lib.batch_launch_graph(plans[k], comp)
        download(k, comp)
        done[k].record(comp)
        done_rec[k].set()
    th.join()
    comp.synchronize()
    copy.synchronize()
    t_host_batch = time.perf_counter() - t0
    # the last wave's roots against the plain drop-in's (earlier waves' buffers
    # were reused by later ones)
    dmism = sum(not np.array_equal(root_of(tr, m), want[i])
                for (i, n, m, w, t, c, tr) in descs[-1])
    for p in plans:
        lib.batch_destroy(p)
    return {
        "workload": f"{len(shapes)} tables, log2 height 12..20, widths {{264,156,379}} "
                    f"(SplitMix64 seed 7), traces from seeds 1000+t, LDE x{E}, Tip5 rows, Merkle",
        "tables": len(shapes), "waves": len(waves), "total_cells": cells,
        "trace_bytes": trace_bytes,
        "device_resident": {"batch_graph_cells_per_s": cells / (t_graph / 1000),
                            "per_table_cells_per_s": cells / (t_plain / 1000),
                            "speedup_vs_per_table": t_plain / t_graph,
                            "ms": {"per_table": t_plain, "batch_graph": t_graph}},
        "host_ingest": {"batched_cells_per_s": cells / t_host_batch,
                        "plain_dropin_cells_per_s": cells / t_host_plain,
                        "plain_dropin_staged_upload_cells_per_s": cells / t_host_plain_staged,
                        "speedup_vs_plain_dropin": t_host_plain / t_host_batch,
                        "speedup_vs_plain_dropin_staged_upload": t_host_plain_staged / t_host_batch,
                        "upload_GB_per_s_batched": trace_bytes / t_host_batch,
                        "s": {"plain_dropin": t_host_plain, "plain_dropin_staged_upload":
                              t_host_plain_staged, "batched": t_host_batch},
                        "root_mismatches_last_wave": dmism},
        "root_mismatches_batch_vs_per_table": mism,
        "lde_forms": tuned,
    }


# ---------------------------------------------------------------- serving
def serve_run(lib, jobs, host_traces=None, inflight=4, hbm_bytes=0, keep=()):
    """Submits `jobs` on their arrival schedule (job["t"], seconds) to one server; a
    second thread waits for each job in submission order, answers its queries
    (opened rows + authentication paths) and releases it.  host_traces: {n: pinned
    trace} for host-ingest jobs, else traces are drawn on the device from the job
    seeds.  Returns per-job (submit, admit, done, answered, root) and kept answers."""
    srv = lib.server_create(hbm_budget_bytes=hbm_bytes, max_inflight=inflight)
    ids = [None] * len(jobs)
    ready = [threading.Event() for _ in jobs]
    res = [None] * len(jobs)
    kept, err = {}, []

    def consumer():
        try:
            for k, j in enumerate(jobs):
                ready[k].wait()
                if ids[k] is None:
                    return
                root, (ts, ta, td) = lib.server_wait(srv, ids[k])
                rows, _ = lib.server_open(srv, ids[k], j["leaves"], j["w"], want_auth=False)
                paths = lib.server_paths(srv, ids[k], j["leaves"], j["m"])
                tq = time.monotonic()
                lib.server_release(srv, ids[k])
                res[k] = (ts, ta, td, tq, root)
                if j["i"] in keep:
                    kept[j["i"]] = (root, rows, paths)
        except Exception as e:
            err.append(e)

    th = threading.Thread(target=consumer)
    th.start()
    t0 = time.monotonic()
    try:
        for k, j in enumerate(jobs):
            delay = t0 + j["t"] - time.monotonic()
            if delay > 0:
                time.sleep(delay)
            tr = host_traces[j["n"]] if host_traces else None
            ids[k] = lib.server_submit(srv, j["w"], j["n"], j["m"], seed=j["seed"], trace=tr)
            ready[k].set()
    finally:
        for e in ready:
            e.set()
        th.join()
        lib.server_destroy(srv)
    if err:
        raise err[0]
    return res, kept


def latency_summary(res):
    lat = [r[2] - r[0] for r in res]
    ans = [r[3] - r[0] for r in res]
    wait = [r[1] - r[0] for r in res]
    return {"commit_p50": pct(lat, 50), "commit_p99": pct(lat, 99),
            "answered_p50": pct(ans, 50), "answered_p99": pct(ans, 99),
            "admission_wait_p50": pct(wait, 50), "admission_wait_p99": pct(wait, 99)}


def run_serve(args, tvh, lib, hip_rt, W, hosts):
    """The served job stream (2000 jobs, Poisson 40 jobs/s, log2 height 14..21 x
    379, 80 queries each) through the commit server: latency at the configured
    rate; capacity with every job submitted at once, device-drawn and
    host-ingest, against the plain per-table path on the same inputs; latency at
    a carried host-ingest load."""
    import numpy as np

    E = args.expansion
    jobs = W.serve_jobs(args.jobs, rate=args.rate, queries=args.queries, width=args.width,
                        expansion=E)
    heights = sorted({j["n"] for j in jobs})
    tuned = autotune(lib, [(n, args.width, E * n) for n in heights], tvh, args.lde_variant,
                     args.no_autotune)
    for n in heights:
        lib.lde_prepare(n, E * n, 1, 7)
    stream = hip_rt.Stream()
    host = {n: hosts[n.bit_length() - 1] for n in heights}  # host trace per height
    # plain per-table paths, timed per height (best of 2): device-resident
    # (draw, commit, root, path gather) and host-ingest (upload, commit,
    # download digests and root)
    nmax = max(heights)
    w = args.width
    tr, cw = hip_rt.DeviceBuffer(nmax * w), hip_rt.DeviceBuffer(E * nmax * w)
    tree = hip_rt.DeviceBuffer(5 * (2 * E * nmax - 1))
    dig = Pinned(hip_rt, 5 * E * nmax)
    plain_dev, plain_host = {}, {}
    for n in heights:
        m = E * n
        leaves = jobs[0]["leaves"] % np.uint64(m)
        for rep in range(3):
            t0 = time.perf_counter()
            lib.fill_uniform(tr, n * w, 5, stream)
            lib.commit(tr, w, n, None, 0, 1, 7, m, cw, tree, 0, stream)
            stream.synchronize()
            tree.download(5, 5 * (2 * m - 2))
            idx = np.concatenate([lib.auth_path_indices(m, int(x)) for x in leaves]).astype(np.uint64)
            d_idx = hip_rt.to_device(idx)
            d_out = hip_rt.DeviceBuffer(5 * idx.size)
            lib.merkle_gather(tree, m, 0, d_idx, idx.size, d_out, stream)
            stream.synchronize()
            d_out.download()
            dt = time.perf_counter() - t0
            if rep:
                plain_dev[n] = min(plain_dev.get(n, dt), dt)
            t0 = time.perf_counter()
            memcpy_async(hip_rt, tr.ptr, host[n].ptr, 8 * n * w, hip_rt.H2D, stream)
            lib.commit(tr, w, n, None, 0, 1, 7, m, cw, tree, 0, stream)
            memcpy_async(hip_rt, dig.ptr, tree.ptr, 40 * m, hip_rt.D2H, stream)
            stream.synchronize()
            tree.download(5, 5 * (2 * m - 2))
            dt = time.perf_counter() - t0
            if rep:
                plain_host[n] = min(plain_host.get(n, dt), dt)
    del tr, cw, tree, dig
    print(f"serve: plain per-height s device {plain_dev} host {plain_host}", file=sys.stderr,
          flush=True)

    cells = sum(j["m"] * j["w"] for j in jobs)
    # (1) the configured stream: device-drawn per-job seeds, Poisson arrivals
    sample = sorted({next(j["i"] for j in jobs if j["n"] == n) for n in heights} |
                    {j["i"] for j in jobs[:4]})
    t0 = time.monotonic()
    res, kept = serve_run(lib, jobs, inflight=args.inflight, keep=set(sample))
    wall = time.monotonic() - t0
    span = max(r[3] for r in res) - min(r[0] for r in res)
    # bit-exactness sample: root, rows and paths against the plain path
    mism = 0
    tr = hip_rt.DeviceBuffer(nmax * w)
    cw = hip_rt.DeviceBuffer(E * nmax * w)
    tree = hip_rt.DeviceBuffer(5 * (2 * E * nmax - 1))
    for i in sample:
        j = jobs[i]
        n, m = j["n"], j["m"]
        lib.fill_uniform(tr, n * w, j["seed"], stream)
        lib.commit(tr, w, n, None, 0, 1, 7, m, cw, tree, 0, stream)
        stream.synchronize()
        root = [int(v) for v in tree.download(5, 5 * (2 * m - 2))]
        sroot, srows, spaths = kept[i]
        rows = np.stack([cw.gather_strided(int(x), m, w) for x in j["leaves"]])
        idx = np.concatenate([lib.auth_path_indices(m, int(x)) for x in j["leaves"]]).astype(np.uint64)
        d_out = hip_rt.DeviceBuffer(5 * idx.size)
        lib.merkle_gather(tree, m, 0, hip_rt.to_device(idx), idx.size, d_out, stream)
        stream.synchronize()
        paths = d_out.download().reshape(len(j["leaves"]), m.bit_length() - 1, 5)
        if sroot != root or not np.array_equal(srows, rows) or not np.array_equal(spaths, paths):
            mism += 1
    del tr, cw, tree
    # (2) capacity: the first jobs all submitted at once, device-drawn and host-ingest
    sat = [dict(j, t=0.0) for j in jobs[:args.saturated_jobs]]
    sat_cells = sum(j["m"] * j["w"] for j in sat)
    sres, _ = serve_run(lib, sat, inflight=args.inflight)
    sat_span = max(r[2] for r in sres) - min(r[0] for r in sres)
    hres, _ = serve_run(lib, sat, host_traces=host, inflight=args.inflight)
    host_span = max(r[2] for r in hres) - min(r[0] for r in hres)
    sat_plain_dev = sum(plain_dev[j["n"]] for j in sat)
    sat_plain_host = sum(plain_host[j["n"]] for j in sat)
    host_cap = sat_cells / host_span
    # (3) latency at a carried load: host-ingest arrivals at 70 % of that capacity
    per_job = cells / len(jobs)
    rate = 0.7 * host_cap / per_job
    carried = W.serve_jobs(args.carried_jobs, rate=rate, queries=args.queries, width=w,
                           expansion=E)
    cres, _ = serve_run(lib, carried, host_traces=host, inflight=args.inflight)
    return {
        "workload": f"{len(jobs)} LDE+commit jobs, Poisson {args.rate} jobs/s, log2 trace height "
                    f"{min(heights).bit_length() - 1}..{max(heights).bit_length() - 1}, width {w}, "
                    f"LDE x{E}, Tip5 rows, Merkle, {args.queries} queries (rows + authentication "
                    f"paths) per job",
        "jobs": len(jobs), "inflight_slots": args.inflight, "total_cells": cells,
        "configured_stream": {"cells_per_s": cells / span, "offered_cells_per_s":
                              cells / max(1e-9, jobs[-1]["t"]), "latency_s": latency_summary(res),
                              "span_s": span, "wall_s": wall, "checked_jobs": len(sample),
                              "mismatches_vs_plain": mism},
        "capacity": {"jobs": len(sat),
                     "device_drawn_cells_per_s": sat_cells / sat_span,
                     "plain_device_per_table_cells_per_s": sat_cells / sat_plain_dev,
                     "speedup_vs_plain_device": sat_plain_dev / sat_span,
                     "host_ingest_cells_per_s": host_cap,
                     "plain_host_dropin_cells_per_s": sat_cells / sat_plain_host,
                     "speedup_vs_plain_host_dropin": sat_plain_host / host_span},
        "carried_load": {"jobs": len(carried), "rate_jobs_per_s": rate,
                         "offered_cells_per_s": rate * per_job, "host_ingest": True,
                         "latency_s": latency_summary(cres)},
        "plain_s_per_job_by_log_height": {
            "device": {str(n.bit_length() - 1): v for n, v in plain_dev.items()},
            "host_dropin": {str(n.bit_length() - 1): v for n, v in plain_host.items()}},
        "lde_forms": tuned,
    }


# ---------------------------------------------------------------- entry
def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--log-n", type=int, default=20)
    ap.add_argument("--width", type=int, default=379)
    ap.add_argument("--expansion", type=int, default=4)
    ap.add_argument("--mode", choices=["main", "aux", "frugal"], default="main",
                    help="headline table: main 2^20 x 379 (default); aux 2^20 x 88 XFE columns "
                         "(264 limbs); frugal 2^21 x 379 streamed in 32-column chunks")
    ap.add_argument("--sections", default="main,cosets,batch,serve,cpu",
                    help="comma list of main, cosets, batch, serve, cpu")
    ap.add_argument("--tables", type=int, default=96, help="batched tables")
    ap.add_argument("--wave-gib", type=float, default=24.0,
                    help="HBM per wave of batched tables (smaller waves overlap host uploads better)")
    ap.add_argument("--chunk", type=int, default=32, help="columns per chunk in frugal mode")
    ap.add_argument("--jobs", type=int, default=2000, help="served jobs")
    ap.add_argument("--rate", type=float, default=40.0, help="mean arrivals/s of the served stream")
    ap.add_argument("--queries", type=int, default=80, help="leaf queries per served job")
    ap.add_argument("--inflight", type=int, default=8,
                    help="server stream slots (8: host uploads keep running while jobs compute)")
    ap.add_argument("--saturated-jobs", type=int, default=400, help="jobs of the capacity runs")
    ap.add_argument("--carried-jobs", type=int, default=400, help="jobs of the carried-load run")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--lde-variant", type=int, default=0)
    ap.add_argument("--calibrate-copy-gib", type=float, default=0.0)
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}, WORLD_SIZE "
              f"{world}); this build has no multi-GPU path.", file=sys.stderr)
        return 2
    sections = set(args.sections.split(","))
    if args.mode == "frugal" and args.log_n == 20:
        args.log_n = 21
    tvh = _load("tvh", os.path.join(PKG, "tvh.py"))
    W = _load("workloads", os.path.join(PKG, "workloads.py"))
    hip_rt = tvh.hip_rt
    lib = tvh.load()
    rc = lib.tvh_init(0)
    if rc:
        raise RuntimeError(f"tvh_init failed: {lib.status_string(rc)}: {lib.last_error()}")
    out = run_main(args, tvh, lib, hip_rt)
    print(f"bench: main {out['value']:.4g} cells/s", file=sys.stderr, flush=True)
    if "cosets" in sections and args.mode == "main":
        out["cosets"] = run_cosets(args, tvh, lib, hip_rt)
        print(f"bench: cosets {out['cosets']['total_ms']:.1f} ms", file=sys.stderr, flush=True)
    hosts = None
    if ("batch" in sections or "serve" in sections) and args.mode == "main":
        # host traces for the host-ingest measurements: heights 2^12..2^21 at
        # width 379 (12.7 GB, pageable), shared by both sections
        hosts = HostTraces(lib, hip_rt, range(12, 22), max(args.width, 379), hip_rt.Stream())
    if "batch" in sections and args.mode == "main":
        out["batch"] = run_batch(args, tvh, lib, hip_rt, W, hosts)
        print("bench: batch done", file=sys.stderr, flush=True)
    if "serve" in sections and args.mode == "main":
        hip_rt.synchronize()  # the batch section's buffers are released before the servers size their arenas
        out["serve"] = run_serve(args, tvh, lib, hip_rt, W, hosts)
        print("bench: serve done", file=sys.stderr, flush=True)
    del hosts
    if "cpu" in sections and args.mode == "main":
        out["cpu_baseline"] = cpu_baseline(args.width, args.expansion)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
