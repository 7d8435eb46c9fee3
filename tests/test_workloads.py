"""The synthetic workload definitions (triton-vm-prover_amd/workloads.py) that
bench.py and the GPU tests share: the field-table draw equals the oracle's
SplitMix64 draw, and the batched/served streams follow their seed rules."""
import importlib.util
import os

import numpy as np

from conftest import ROOT

spec = importlib.util.spec_from_file_location(
    "workloads", os.path.join(ROOT, "triton-vm-prover_amd", "workloads.py"))
W = importlib.util.module_from_spec(spec)
spec.loader.exec_module(W)


def test_rowmajor_table_is_the_oracle_draw(oracle):
    for seed, rows, cols in [(1, 8, 7), (1, 1024, 379), (2, 16, 264), (12, 16, 9)]:
        assert np.array_equal(W.rowmajor_table(seed, rows, cols),
                              oracle.table_rowmajor_draw(seed, rows, cols))


def test_stream_and_rejection():
    g = W.splitmix_stream(0)
    first = [next(g) for _ in range(3)]
    # SplitMix64 from state 0: published first outputs
    assert first[0] == 0xE220A8397B1DCDAF
    assert first[1] == 0x6E789E6AA1B965F4
    # a stream whose draws include values >= p: the rejected ones are skipped
    z = W._splitmix_block(5, 0, 1 << 16)
    keep = z[z < np.uint64(W.P)]
    assert np.array_equal(W.splitmix_field(5, keep.size), keep)


def test_device_fill_rule():
    z = W._splitmix_block(9, 0, 1000)
    d = W.device_fill(9, 1000)
    assert np.array_equal(d[z < np.uint64(W.P)], z[z < np.uint64(W.P)])
    assert int(d.max()) < W.P
    assert np.array_equal(W.device_fill(9, 10, first=990), d[990:])


def test_batch_shapes():
    s = W.batch_shapes(96)
    assert len(s) == 96
    assert all(12 <= lg <= 20 and w in (264, 156, 379) for lg, w in s)
    assert len({lg for lg, _ in s}) >= 7 and len({w for _, w in s}) == 3
    assert s == W.batch_shapes(96)


def test_serve_jobs():
    jobs = W.serve_jobs(2000)
    assert len(jobs) == 2000
    assert all(14 <= j["n"].bit_length() - 1 <= 21 and j["m"] == 4 * j["n"] for j in jobs)
    ts = np.array([j["t"] for j in jobs])
    assert ts[0] == 0 and np.all(np.diff(ts) >= 0)
    assert 40.0 / 1.2 < 1999 / ts[-1] < 40.0 * 1.2  # mean rate 40 jobs/s
    assert all(len(j["leaves"]) == 80 and int(j["leaves"].max()) < j["m"] for j in jobs)
    assert [j["seed"] for j in jobs[:3]] == [100000, 100001, 100002]
