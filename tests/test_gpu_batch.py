"""Batched commits (tvh_batch_plan_*, tvh_commit_batch_dev): every table of a
mixed batch must come out exactly as its own tvh_commit_dev, and small ones
as the oracle; hipGraph replays are checked after eager work in between."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [  # (log n, width, expansion, layout, randomizers)
    (0, 3, 4, 0, 0), (2, 11, 4, 1, 0), (5, 37, 2, 0, 3), (6, 156, 4, 0, 0), (8, 264, 4, 1, 5),
    (10, 379, 4, 0, 0), (11, 9, 8, 0, 0), (12, 20, 4, 0, 0), (13, 7, 4, 1, 0), (14, 33, 4, 0, 0),
    (9, 1, 16, 0, 2), (12, 0, 4, 0, 0),
]


def make_tables(T, seed0, hip_rt):
    tabs = []
    for i, (lg, w, E, layout, k) in enumerate(SHAPES):
        n, m = 1 << lg, E << lg
        tr = T.empty(n * w)
        if w:
            T.lib.fill_uniform(tr, n * w, seed0 + i)
        r = None
        if k:
            r = T.empty(w * k)
            T.lib.fill_uniform(r, w * k, seed0 + 100 + i)
        cw = T.empty(m * w)
        nodes = 2 * m if layout == 1 else 2 * m - 1
        tree = T.empty(5 * nodes)
        tabs.append(dict(n=n, w=w, m=m, k=k, layout=layout, trace=tr, rand=r, cw=cw, tree=tree,
                         nodes=nodes))
    hip_rt.synchronize()
    return tabs


def descs(T, tabs, key="tree"):
    return [T.lib.table_desc(t["trace"], t["w"], t["n"], t["m"], t["cw"], t[key],
                             rand=t["rand"], k=t["k"], layout=t["layout"]) for t in tabs]


def reference_trees(T, tabs):
    out = []
    for t in tabs:
        cw = T.empty(t["m"] * t["w"])
        tree = T.empty(5 * t["nodes"])
        T.lib.commit(t["trace"], t["w"], t["n"], t["rand"], t["k"], 1, 7, t["m"], cw, tree,
                     t["layout"])
        out.append((T.host(tree), T.host(cw)))
    return out


def root_of(t, tree):
    nodes = tree.reshape(-1, 5)
    return nodes[1] if t["layout"] == 1 else nodes[-1]


def test_commit_batch_matches_single(T, oracle, tvhmod):
    hip_rt = tvhmod.hip_rt
    tabs = make_tables(T, 1000, hip_rt)
    want = reference_trees(T, tabs)
    T.lib.commit_batch(descs(T, tabs))
    for t, (wt, wcw) in zip(tabs, want):
        got = T.host(t["tree"])
        if t["layout"] == 1:
            got, wt = got[5:], wt[5:]  # node 0 is unused in the heap layout
        assert np.array_equal(got, wt), (t["n"], t["w"])
        if t["w"]:  # a zero-width table has no codewords (its buffer is a 1-element stub)
            assert np.array_equal(T.host(t["cw"]), wcw)
    for t in tabs:  # small tables against the oracle as well
        if t["n"] * t["w"] <= 4096:
            tr = T.host(t["trace"], t["n"] * t["w"])
            r = T.host(t["rand"], t["w"] * t["k"]).reshape(t["w"], t["k"]) if t["k"] else None
            root, _, _ = oracle.commit(tr, t["n"], t["w"], t["m"], rand=r)
            assert np.array_equal(root_of(t, T.host(t["tree"])), root)


def test_batch_graph_replays(T, tvhmod):
    hip_rt = tvhmod.hip_rt
    tabs = make_tables(T, 2000, hip_rt)
    stream = hip_rt.Stream()
    plan = T.lib.batch_plan(descs(T, tabs))
    try:
        for rep in range(3):
            # new inputs (eager work between replays), then the captured graph
            for i, t in enumerate(tabs):
                if t["w"]:
                    T.lib.fill_uniform(t["trace"], t["n"] * t["w"], 3000 + 10 * rep + i, stream)
            stream.synchronize()
            want = reference_trees(T, tabs)
            hip_rt.synchronize()
            T.lib.batch_launch_graph(plan, stream)
            stream.synchronize()
            for t, (wt, _) in zip(tabs, want):
                assert np.array_equal(root_of(t, T.host(t["tree"])), root_of(t, wt)), (rep, t["n"])
    finally:
        T.lib.batch_destroy(plan)


def test_fill_uniform_is_canonical_and_seeded(T):
    a, b = T.empty(100000), T.empty(100000)
    T.lib.fill_uniform(a, 100000, 5)
    T.lib.fill_uniform(b, 100000, 5)
    ha, hb = T.host(a), T.host(b)
    assert np.array_equal(ha, hb)
    assert int(ha.max()) < 0xFFFFFFFF00000001
    assert len(np.unique(ha)) > 99990
