"""The two systems workloads at their real sizes.

Batched tables: 96 tables, log2 height 12..20 and width {264, 156, 379} drawn
from SplitMix64(7), traces drawn on the device from seed 1000 + t, LDE x4 and
commit.  Every table's root from the hipGraph-replayed batch plan must equal
its own tvh_commit_dev, and the smallest tables' roots the oracle's.  Tables
go through the plan in waves that bound HBM use.

Served stream: 2000 LDE + commit jobs, log2 height 14..21 x 379, Poisson
arrivals at 40 jobs/s, per-job seeds, 80 queries each, through the commit
server.  Every job's root must equal the plain per-table path's; for the first
job of each height the opened rows, authentication structure and paths must
equal the plain path's too, and the paths must verify against the root with
the oracle's hash of the opened rows.
"""
import importlib.util
import os
import threading
import time

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

spec = importlib.util.spec_from_file_location(
    "workloads", os.path.join(ROOT, "triton-vm-prover_amd", "workloads.py"))
W = importlib.util.module_from_spec(spec)
spec.loader.exec_module(W)

E = 4
WAVE_BYTES = 96 << 30


def _table_bytes(lg, w):
    n = 1 << lg
    return 8 * n * w + 8 * E * n * w + 40 * (2 * E * n - 1)


def test_batch_96_tables(T, tvhmod, oracle):
    hip_rt = tvhmod.hip_rt
    lib = T.lib
    shapes = W.batch_shapes(96)
    order = sorted(range(96), key=lambda i: -_table_bytes(*shapes[i]))
    waves, cur, used = [], [], 0
    for i in order:
        b = _table_bytes(*shapes[i])
        if cur and used + b > WAVE_BYTES:
            waves.append(cur)
            cur, used = [], 0
        cur.append(i)
        used += b
    waves.append(cur)
    smallest = sorted(range(96), key=lambda i: _table_bytes(*shapes[i]))[:2]
    stream = hip_rt.Stream()
    checked = 0
    for wi, wave in enumerate(waves):
        print(f"batch wave {wi + 1}/{len(waves)}: {len(wave)} tables", flush=True)
        tabs = []
        for i in wave:
            lg, w = shapes[i]
            n, m = 1 << lg, E << lg
            tr = hip_rt.DeviceBuffer(n * w)
            lib.fill_uniform(tr, n * w, W.batch_seed(i), stream)
            tabs.append(dict(i=i, n=n, m=m, w=w, trace=tr, cw=hip_rt.DeviceBuffer(m * w),
                             tree=hip_rt.DeviceBuffer(5 * (2 * m - 1))))
        # per-table plain commits first: their roots are the reference
        for t in tabs:
            lib.commit(t["trace"], t["w"], t["n"], None, 0, 1, 7, t["m"], t["cw"], t["tree"], 0,
                       stream)
        stream.synchronize()
        want = {t["i"]: t["tree"].download(5, 5 * (2 * t["m"] - 2)) for t in tabs}
        for t in tabs:  # wipe, so the plan must rebuild every node
            t["tree"].zero()
        plan = lib.batch_plan([lib.table_desc(t["trace"], t["w"], t["n"], t["m"], t["cw"],
                                              t["tree"]) for t in tabs])
        try:
            for rep in range(2):  # capture + launch, then a replay of the graph
                lib.batch_launch_graph(plan, stream)
                stream.synchronize()
                for t in tabs:
                    got = t["tree"].download(5, 5 * (2 * t["m"] - 2))
                    assert np.array_equal(got, want[t["i"]]), (rep, t["i"], t["n"], t["w"])
                    checked += rep
                for t in tabs:
                    t["tree"].zero()
        finally:
            lib.batch_destroy(plan)
        for t in tabs:
            if t["i"] in smallest:
                tr = t["trace"].download()
                root, _, _ = oracle.commit(tr, t["n"], t["w"], t["m"])
                assert np.array_equal(want[t["i"]], root), t["i"]
        del tabs
    assert checked == 96


def _plain_commit(lib, hip_rt, j, bufs, stream):
    """The plain per-table path on preallocated buffers: fill, commit, root."""
    n, m, w = j["n"], j["m"], j["w"]
    tr, cw, tree = bufs
    lib.fill_uniform(tr, n * w, j["seed"], stream)
    lib.commit(tr, w, n, None, 0, 1, 7, m, cw, tree, 0, stream)
    stream.synchronize()
    return [int(v) for v in tree.download(5, 5 * (2 * m - 2))]


def test_serve_2000_jobs(T, tvhmod, oracle):
    hip_rt = tvhmod.hip_rt
    lib = T.lib
    jobs = W.serve_jobs(2000)
    keep = {}
    for j in jobs:  # the first job of each height keeps its buffers for the opening checks
        keep.setdefault(j["n"], j["i"])
    keep_ids = set(keep.values())
    # the plain path's buffers (largest shape) before the server takes its budget
    nmax = max(j["n"] for j in jobs)
    w = jobs[0]["w"]
    bufs = (hip_rt.DeviceBuffer(nmax * w), hip_rt.DeviceBuffer(E * nmax * w),
            hip_rt.DeviceBuffer(5 * (2 * E * nmax - 1)))
    free0, _ = hip_rt.mem_info()
    srv = lib.server_create(max_inflight=4)
    ids = [None] * len(jobs)
    ready = [threading.Event() for _ in jobs]
    res = [None] * len(jobs)
    opened = {}
    err = []

    def consumer():
        try:
            for k, j in enumerate(jobs):
                ready[k].wait()
                if ids[k] is None:
                    return
                root, (ts, ta, td) = lib.server_wait(srv, ids[k])
                paths = lib.server_paths(srv, ids[k], j["leaves"], j["m"])
                if j["i"] in keep_ids:
                    rows, auth = lib.server_open(srv, ids[k], j["leaves"], j["w"])
                    opened[j["i"]] = (root, rows, auth, paths)
                lib.server_release(srv, ids[k])
                res[k] = (ts, ta, td, root)
                if (k + 1) % 250 == 0:
                    print(f"serve: {k + 1}/{len(jobs)} jobs done", flush=True)
        except Exception as e:  # re-raised below
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
    free1, _ = hip_rt.mem_info()
    print(f"serve: free HBM before create {free0 / 2**30:.1f} GiB, after destroy "
          f"{free1 / 2**30:.1f} GiB", flush=True)
    if err:
        raise err[0]
    assert free1 > free0 - (4 << 30), "the server's arena was not returned on destroy"
    lat = np.array([r[2] - r[0] for r in res])
    print(f"serve: commit latency p50 {np.percentile(lat, 50):.3f} s, "
          f"p99 {np.percentile(lat, 99):.3f} s", flush=True)

    # the plain path for every job, on the buffers of the largest shape
    stream = hip_rt.Stream()
    for k, j in enumerate(jobs):
        root = _plain_commit(lib, hip_rt, j, bufs, stream)
        assert res[k][3] == root, (k, j["n"])
        if j["i"] in keep_ids:
            _, rows, auth, paths = opened[j["i"]]
            m = j["m"]
            tr, cw, tree = bufs
            for q, leaf in enumerate(j["leaves"]):
                assert np.array_equal(rows[q], cw.gather_strided(int(leaf), m, w)), (k, leaf)
                digest = oracle.hash_varlen(rows[q])
                assert np.array_equal(tree.download(5, 5 * int(leaf)), digest), (k, leaf)
                assert oracle.verify_path(root, int(leaf), digest, paths[q]), (k, leaf)
            hidx = lib.auth_structure_indices(m, j["leaves"])
            d_out = hip_rt.DeviceBuffer(5 * max(1, hidx.size))
            lib.merkle_gather(tree, m, 0, T.dev(hidx), hidx.size, d_out)
            T.sync()
            assert np.array_equal(auth.reshape(-1), d_out.download(5 * hidx.size)), k
        if (k + 1) % 500 == 0:
            print(f"plain path: {k + 1}/{len(jobs)} roots equal", flush=True)
    assert len(opened) == len(keep) == 8
