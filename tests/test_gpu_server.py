"""The device-resident commit server (tvh_server_*): concurrent jobs admitted by HBM
budget must give the same roots, opened rows, authentication structures and paths as
the oracle, and each job's result must be the same alone and in a mix of others."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GAMMA = np.uint64(0x9E3779B97F4A7C15)
P = 0xFFFFFFFF00000001


def fill_uniform(seed, count):
    """numpy statement of tvh_fill_uniform_dev (SplitMix64 at seed + (i+1)*gamma, then
    values >= p lowered by p)."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (np.arange(1, count + 1, dtype=np.uint64) * GAMMA)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return np.where(z >= np.uint64(P), z - np.uint64(P), z)


def job_inputs(j):
    """Column-major trace and randomizers of a job spec (host-given or seed-drawn)."""
    if j.get("trace") is not None:
        return j["trace"], j.get("rand")
    tr = fill_uniform(j["seed"], j["n"] * j["w"])
    rd = fill_uniform(j["seed"] ^ 0x5BD1E995, j["w"] * j["k"]) if j["k"] else None
    return tr, rd


def submit(lib, srv, j):
    return lib.server_submit(srv, j["w"], j["n"], j["m"], seed=j.get("seed", 0),
                             trace=j.get("trace"), rand=j.get("rand"), k=j["k"])


def check_job(lib, srv, jid, j, oracle, rng):
    tr, rd = job_inputs(j)
    root, times = lib.server_wait(srv, jid)
    want_root, cw, dg = oracle.commit(tr, j["n"], j["w"], j["m"], rand=rd, want_codewords=True,
                                      want_digests=True)
    assert root == [int(v) for v in want_root], j
    assert times[0] <= times[1] <= times[2]
    leaves = rng.choice(j["m"], size=min(j["m"], 6), replace=False)
    rows, auth = lib.server_open(srv, jid, leaves, j["w"])
    cwm = cw.reshape(j["w"], j["m"]) if j["w"] else np.zeros((0, j["m"]), np.uint64)
    for q, leaf in enumerate(leaves):
        assert np.array_equal(rows[q], cwm[:, leaf]), (j, leaf)
    heap = oracle.merkle_heap(dg)
    want_idx = oracle.auth_structure_indices(j["m"], leaves)
    assert np.array_equal(auth, heap[want_idx.astype(np.int64)]), j
    paths = lib.server_paths(srv, jid, leaves, j["m"])
    for q, leaf in enumerate(leaves):
        assert oracle.verify_path(want_root, int(leaf), dg[leaf], paths[q]), (j, leaf)
    return root, rows, auth, paths


def test_server_jobs_match_oracle(T, oracle):
    lib = T.lib
    rng = np.random.default_rng(71)
    specs = [dict(n=16, w=7, m=64, k=0, seed=1), dict(n=64, w=37, m=256, k=3, seed=2),
             dict(n=1, w=3, m=4, k=0, seed=3), dict(n=256, w=379, m=1024, k=0, seed=4),
             dict(n=32, w=1, m=512, k=2, seed=5), dict(n=128, w=156, m=512, k=0, seed=6),
             dict(n=8, w=0, m=32, k=0, seed=7), dict(n=512, w=23, m=2048, k=4, seed=8)]
    host = oracle.table_rowmajor_draw(99, 64, 11)
    specs.append(dict(n=64, w=11, m=256, k=2, trace=host, rand=oracle.splitmix(98, 22)))
    srv = lib.server_create(hbm_budget_bytes=1 << 30, max_inflight=3)
    try:
        ids = [submit(lib, srv, j) for j in specs]
        for jid, j in zip(ids, specs):
            check_job(lib, srv, jid, j, oracle, rng)
        for jid in ids:
            lib.server_release(srv, jid)
        st = lib.server_stats(srv)
        assert st["arena_used"] == 0 and st["queued"] == 0 and st["running"] == 0
    finally:
        lib.server_destroy(srv)


def test_same_result_alone_and_in_mix(T, oracle, tvhmod):
    """One job run alone on a one-slot server, then the same job in a mix of 24
    concurrent jobs of other sizes on a four-slot server: identical root, opened rows,
    authentication structure and paths; also identical to a plain tvh_commit_dev."""
    lib = T.lib
    probe = dict(n=1 << 12, w=379, m=1 << 14, k=0, seed=424242)
    leaves = np.random.default_rng(5).choice(probe["m"], size=80, replace=False)

    def run(srv):
        jid = submit(lib, srv, probe)
        return jid

    def collect(srv, jid):
        root, _ = lib.server_wait(srv, jid)
        rows, auth = lib.server_open(srv, jid, leaves, probe["w"])
        paths = lib.server_paths(srv, jid, leaves, probe["m"])
        return root, rows, auth, paths

    srv = lib.server_create(hbm_budget_bytes=1 << 30, max_inflight=1)
    try:
        alone = collect(srv, run(srv))
    finally:
        lib.server_destroy(srv)

    rng = np.random.default_rng(8)
    srv = lib.server_create(hbm_budget_bytes=4 << 30, max_inflight=4)
    try:
        others = []
        for i in range(24):
            lg = int(rng.integers(6, 14))
            j = dict(n=1 << lg, w=int(rng.choice([7, 156, 379])), m=4 << lg, k=int(rng.integers(0, 3)),
                     seed=1000 + i)
            others.append((submit(lib, srv, j), j))
            if i == 11:
                pid = run(srv)
        mixed = collect(srv, pid)
        for jid, j in others:
            lib.server_wait(srv, jid)
        # spot-check two of the others against the oracle
        for jid, j in others[:2]:
            check_job(lib, srv, jid, j, oracle, rng)
    finally:
        lib.server_destroy(srv)

    assert alone[0] == mixed[0]
    for a, b in zip(alone[1:], mixed[1:]):
        assert np.array_equal(a, b)

    # plain drop-in path on the same input
    tr = T.empty(probe["n"] * probe["w"])
    lib.fill_uniform(tr, probe["n"] * probe["w"], probe["seed"])
    cw = T.empty(probe["m"] * probe["w"])
    tree = T.empty(5 * (2 * probe["m"] - 1))
    lib.commit(tr, probe["w"], probe["n"], None, 0, 1, 7, probe["m"], cw, tree)
    T.sync()
    assert [int(v) for v in T.host(tree).reshape(-1, 5)[-1]] == alone[0]


def test_admission_waits_for_memory(T, oracle, tvhmod):
    """Budget for one job: the second waits (queued) while the first is held, a third
    is refused by the queue bound, a job larger than the budget fails with OOM, and
    releasing the first admits the second."""
    lib, mod = T.lib, tvhmod
    j = dict(n=1 << 10, w=64, m=1 << 12, k=0)
    need = 8 * (j["n"] * j["w"] + j["m"] * j["w"]) + 40 * (2 * j["m"] - 1)
    srv = lib.server_create(hbm_budget_bytes=need + (need // 2), max_inflight=2, max_queue=1)
    try:
        a = submit(lib, srv, dict(j, seed=11))
        lib.server_wait(srv, a)
        b = submit(lib, srv, dict(j, seed=12))
        with pytest.raises(mod.TVHError) as e:
            submit(lib, srv, dict(j, seed=13))
        assert e.value.status == 8  # TVH_ERR_RANGE: queue full
        assert lib.server_poll(srv, b) == mod.JOB_QUEUED
        st = lib.server_stats(srv)
        assert st["queued"] == 1 and st["running"] == 0 and st["arena_used"] <= st["arena_size"]
        lib.server_release(srv, a)
        check_job(lib, srv, b, dict(j, seed=12), oracle, np.random.default_rng(2))
        big = submit(lib, srv, dict(n=1 << 12, w=64, m=1 << 14, k=0, seed=14))
        with pytest.raises(mod.TVHError) as e:
            lib.server_wait(srv, big)
        assert e.value.status == 3  # TVH_ERR_OOM
        assert lib.server_poll(srv, big) == mod.JOB_FAILED
        lib.server_release(srv, big)
        lib.server_release(srv, b)
        assert lib.server_stats(srv)["arena_used"] == 0
    finally:
        lib.server_destroy(srv)


def test_server_argument_errors(T, tvhmod):
    lib, mod = T.lib, tvhmod
    srv = lib.server_create(hbm_budget_bytes=64 << 20, max_inflight=1)
    try:
        with pytest.raises(mod.TVHError) as e:
            lib.server_submit(srv, 3, 12, 48)
        assert e.value.status == 2  # not a power of two
        with pytest.raises(mod.TVHError):
            lib.server_submit(srv, 3, 16, 8)  # target shorter than trace
        with pytest.raises(mod.TVHError) as e:
            lib.server_open(srv, 12345, [0], 3)
        assert e.value.status == 8
        jid = lib.server_submit(srv, 3, 16, 64, seed=1)
        lib.server_wait(srv, jid)
        with pytest.raises(mod.TVHError) as e:
            lib.server_open(srv, jid, [64], 3)
        assert e.value.status == 8  # leaf out of range
        lib.server_release(srv, jid)
        with pytest.raises(mod.TVHError):
            lib.server_wait(srv, jid)
    finally:
        lib.server_destroy(srv)
