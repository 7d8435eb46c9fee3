"""Pins the CPU oracle (oracle/tvm_oracle.c) before anything is compared to it.

Sources of truth, none of which retypes the oracle's own formulas:
  * Rust twenty-first known-answer tests (tip5/mod.rs hash10/hash_varlen,
    math/ntt/mod.rs bfield_basic_test_of_chu_ntt), values quoted below;
  * golden fixtures written from the reference's compiled CPU code
    (tests/golden/golden.json, tests/golden/make_golden.py);
  * mathematics: O(n^2) DFT, Horner evaluation of the interpolant on the coset,
    S-box closed form, MDS = SHA-256("Tip5"), Merkle verify round trips;
  * the compiled reference itself on fresh random inputs, when built here.
"""
import hashlib

import numpy as np
import pytest

from conftest import P, u64


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<u8").tobytes()).hexdigest()


# ---------------------------------------------------------------- field
def test_field_edges(oracle):
    o = oracle
    cases = [0, 1, 2, P - 1, P - 2, 2**32 - 1, 2**32, 2**63, P - 2**32, 12345678901234567]
    for a in cases:
        for b in cases:
            assert o.add(a, b) == (a + b) % P
            assert o.sub(a, b) == (a - b) % P
            assert o.mul(a, b) == (a * b) % P
    rng = np.random.default_rng(1)
    for _ in range(2000):
        a, b = (int(v) % P for v in rng.integers(0, 2**64, size=2, dtype=np.uint64))
        assert o.mul(a, b) == a * b % P
    assert o.inv(7) * 7 % P == 1
    assert o.pow(3, P - 1) == 1


def test_montyred_is_canonical_multiplication_by_r_inverse(oracle):
    rinv = pow(2**64, -1, P)
    rng = np.random.default_rng(2)
    r2 = 0xFFFFFFFE00000001
    for _ in range(3000):
        v = int(rng.integers(0, 2**64, dtype=np.uint64))
        x = v * r2
        got = oracle.montyred(x >> 64, x & (2**64 - 1))
        assert got == x * rinv % P and got < P


def test_roots_of_unity(oracle):
    for k in range(0, 33):
        w = oracle.root_of_unity(k)
        assert pow(w, 2**k, P) == 1
        if k:
            assert pow(w, 2 ** (k - 1), P) == P - 1
        assert w == pow(7, (P - 1) >> k, P)


# ---------------------------------------------------------------- NTT
def test_ntt_rust_kat(oracle):
    out = oracle.ntt([1, 4, 0, 0])
    assert [int(x) for x in out] == [5, 1125899906842625, 18446744069414584318,
                                     18445618169507741698]
    assert [int(x) for x in oracle.intt(out)] == [1, 4, 0, 0]


@pytest.mark.parametrize("lg", [0, 1, 2, 3, 5, 7])
def test_ntt_matches_dft(oracle, lg):
    n = 1 << lg
    x = [int(v) for v in oracle.splitmix(77 + lg, n)]
    w = pow(7, (P - 1) >> lg, P)
    dft = [sum(x[i] * pow(w, i * j, P) for i in range(n)) % P for j in range(n)]
    assert [int(v) for v in oracle.ntt(x)] == dft
    winv = pow(w, P - 2, P)
    ninv = pow(n, P - 2, P)
    idft = [sum(x[i] * pow(winv, i * j, P) for i in range(n)) * ninv % P for j in range(n)]
    assert [int(v) for v in oracle.intt(x)] == idft


def test_ntt_golden(oracle, golden):
    for c in golden["ntt"]:
        x = oracle.splitmix(c["seed"], c["n"])
        assert sha(oracle.ntt(x)) == c["fwd_sha256"]
        assert sha(oracle.intt(x)) == c["inv_sha256"]


def test_ntt_rejects_non_power_of_two(oracle):
    with pytest.raises(ValueError):
        oracle.ntt([1, 2, 3])


# ---------------------------------------------------------------- LDE
def horner_lde(trace, m, toff, goff, rand):
    """Interpolant by Lagrange-free route: solve via the O(n^2) inverse DFT,
    then evaluate the randomized polynomial at goff * w_m^j with Horner."""
    n = len(trace)
    lg = n.bit_length() - 1
    w = pow(7, (P - 1) >> lg, P)
    winv, ninv, oinv = pow(w, P - 2, P), pow(n, P - 2, P), pow(toff, P - 2, P)
    c = [sum(trace[q] * pow(winv, i * q, P) for q in range(n)) * ninv % P for i in range(n)]
    c = [ci * pow(oinv, i, P) % P for i, ci in enumerate(c)]
    poly = c + [0] * len(rand)
    on = pow(toff, n, P)
    for i, r in enumerate(rand):
        poly[i] = (poly[i] - r * on) % P
        poly[n + i] = (poly[n + i] + r) % P
    wm = pow(7, (P - 1) >> (m.bit_length() - 1), P)
    out = []
    for j in range(m):
        xj = goff * pow(wm, j, P) % P
        acc = 0
        for a in reversed(poly):
            acc = (acc * xj + a) % P
        out.append(acc)
    return out


@pytest.mark.parametrize("n,m,k,toff", [(4, 16, 0, 1), (8, 32, 3, 1), (8, 8, 3, 3),
                                        (4, 4, 6, 5), (16, 64, 2, 1)])
def test_lde_matches_horner(oracle, n, m, k, toff):
    x = [int(v) for v in oracle.splitmix(10 + n + k, n)]
    r = [int(v) for v in oracle.splitmix(20 + k, k)] if k else []
    got = oracle.lde_column(x, m, toff, 7, r if k else None)
    assert [int(v) for v in got] == horner_lde(x, m, toff, 7, r)


@pytest.mark.parametrize("n,m,k,toff", [(16, 64, 0, 1), (32, 128, 5, 3), (8, 8, 8, 1),
                                         (64, 1024, 2, 11), (4, 64, 0, 7)])
def test_lde_sample_matches_full_lde(oracle, n, m, k, toff):
    """The sampled-row evaluation equals rows of the full (chunked) LDE and, with
    truncate, of the XFE-rule LDE."""
    cols = 3
    x = oracle.splitmix(300 + n, n * cols)
    r = oracle.splitmix(301 + k, k * cols).reshape(cols, k) if k else None
    rows = [0, 1, m - 1, m // 2, 5 % m]
    got = oracle.lde_sample(x, n, cols, m, rows, toff, 7, r)
    full = oracle.lde_table(x, n, cols, m, toff, 7, r).reshape(cols, m)
    assert np.array_equal(got, full[:, rows])
    got_t = oracle.lde_sample(x, n, cols, m, rows, toff, 7, r, truncate=True)
    for c in range(cols):
        want = oracle.lde_limb_truncating(x[c * n:(c + 1) * n], m, toff, 7,
                                          None if r is None else r[c])
        assert np.array_equal(got_t[c], want[rows])


def test_lde_contains_trace_at_stride(oracle):
    # offset-1 target: the LDE on <w_m> restricted to stride m/n is the trace
    x = oracle.splitmix(31, 64)
    out = oracle.lde_column(x, 256, 1, 1)
    assert np.array_equal(out[::4], x)


def test_lde_golden(oracle, golden):
    for c in golden["lde_column"]:
        x = oracle.splitmix(c["seed_trace"], c["n"])
        r = oracle.splitmix(c["seed_rand"], c["k"]) if c["k"] else None
        got = oracle.lde_column(x, c["m"], c["trace_off"], c["target_off"], r)
        assert [int(v) for v in got] == c["out"], c


def test_lde_xfe_golden(oracle, golden):
    for c in golden["lde_xfe_column"]:
        n, m, k = c["n"], c["m"], c["k"]
        x = oracle.splitmix(c["seed_trace"], 3 * n).reshape(n, 3)
        if k:
            r = oracle.splitmix(c["seed_rand"], 3 * k if c["xfe_rand"] else k)
            rl = [r.reshape(k, 3)[:, j] for j in range(3)] if c["xfe_rand"] else [r, None, None]
        else:
            rl = [None, None, None]
        got = np.stack([oracle.lde_limb_truncating(x[:, j], m, c["trace_off"], c["target_off"],
                                                   rl[j]) for j in range(3)], axis=1)
        assert [int(v) for v in got.reshape(-1)] == c["out"], c


# ---------------------------------------------------------------- Tip5
def test_tip5_constants(oracle):
    lut = oracle.lookup_table()
    assert [int(v) for v in lut] == [((x + 1) ** 3 % 257 - 1) % 256 for x in range(256)]
    assert sorted(int(v) for v in lut) == list(range(256))
    h = hashlib.sha256(b"Tip5").digest()
    assert oracle.mds_first_column() == [int.from_bytes(h[2 * i:2 * i + 2], "little")
                                         for i in range(16)]
    assert all(int(c) < P for c in oracle.round_constants())


def test_tip5_rust_hash10_kat(oracle):
    pre = np.zeros(10, np.uint64)
    for i in range(6):
        pre[i:i + 5] = oracle.hash_10(pre)
    assert [int(v) for v in oracle.hash_10(pre)] == [
        10869784347448351760, 1853783032222938415, 6856460589287344822,
        17178399545409290325, 7650660984651717733]


def test_tip5_rust_hash_varlen_kat(oracle):
    s = [0] * 5
    for i in range(20):
        d = oracle.hash_varlen(np.arange(i, dtype=np.uint64))
        s = [(a + int(b)) % P for a, b in zip(s, d)]
    assert s == [7610004073009036015, 5725198067541094245, 4721320565792709122,
                 1732504843634706218, 259800783350288362]


def test_tip5_golden(oracle, golden):
    for c in golden["tip5_permutation"]:
        assert [int(v) for v in oracle.permutation(c["in"])] == c["out"]
    for c in golden["hash_varlen"]:
        assert [int(v) for v in oracle.hash_varlen(oracle.splitmix(c["seed"], c["len"]))] == c["out"]
    for c in golden["hash_pair"]:
        got = oracle.hash_pair(oracle.splitmix(c["seed_l"], 5), oracle.splitmix(c["seed_r"], 5))
        assert [int(v) for v in got] == c["out"]


def test_tip5_rounds(oracle, golden):
    # the zero state is fixed by the S-box and the MDS, so round r maps it to
    # the round-r constants (pins the constant index of each round)
    rc = oracle.round_constants()
    for r in range(5):
        assert np.array_equal(oracle.tip5_round(np.zeros(16, np.uint64), r), rc[16 * r:16 * r + 16])
    # rounds 0..4 in order compose to the golden permutations
    for c in golden["tip5_permutation"]:
        st = np.array(c["in"], np.uint64)
        for r in range(5):
            st = oracle.tip5_round(st, r)
        assert [int(v) for v in st] == c["out"]


def test_hash_pair_is_hash10_with_capacity_ones(oracle):
    l, r = oracle.splitmix(1, 5), oracle.splitmix(2, 5)
    st = np.concatenate([l, r, np.ones(6, np.uint64)])
    assert np.array_equal(oracle.permutation(st)[:5], oracle.hash_pair(l, r))
    # order matters
    assert not np.array_equal(oracle.hash_pair(l, r), oracle.hash_pair(r, l))


# ---------------------------------------------------------------- Merkle
def test_merkle_golden(oracle, golden):
    g = golden["merkle"]
    leaves = oracle.splitmix(g["seed"], g["n"] * 5).reshape(g["n"], 5)
    nodes = oracle.merkle_heap(leaves)
    assert [int(v) for v in nodes[1]] == g["root"]
    assert sha(nodes) == g["nodes_sha256"]
    tree = oracle.merkle_level_order(leaves)
    assert np.array_equal(tree[-1], nodes[1])
    for c in golden["auth_structure"]:
        idx = oracle.auth_structure_indices(g["n"], c["indices"])
        assert [int(v) for v in nodes[idx].reshape(-1)] == c["digests"], c


def test_merkle_paths_verify(oracle):
    leaves = oracle.splitmix(5, 32 * 5).reshape(32, 5)
    nodes = oracle.merkle_heap(leaves)
    for i in [0, 1, 17, 31]:
        v, path = 32 + i, []
        while v > 1:
            path.append(nodes[v ^ 1])
            v //= 2
        assert oracle.verify_path(nodes[1], i, leaves[i], np.array(path))
        assert not oracle.verify_path(nodes[1], i ^ 1, leaves[i], np.array(path))


def test_merkle_rejects_non_power_of_two(oracle):
    with pytest.raises(ValueError):
        oracle.merkle_heap(np.zeros((3, 5), np.uint64))


# ---------------------------------------------------------------- commit
def test_commit_golden(oracle, golden):
    for c in golden["commit"]:
        if c["n"] * c["w"] > 300_000:
            continue  # the 2^10 x 379 case runs in test_commit_golden_config1
        t = oracle.table_rowmajor_draw(c["seed"], c["n"], c["w"])
        root, cw, dg = oracle.commit(t, c["n"], c["w"], c["m"], want_codewords=True,
                                     want_digests=True)
        assert [int(v) for v in root] == c["root"], (c["n"], c["w"])
        assert sha(cw) == c["codewords_sha256"]
        assert sha(dg) == c["digests_sha256"]


def test_commit_golden_config1(oracle, golden):
    c = next(c for c in golden["commit"] if c["n"] == 1024 and c["w"] == 379)
    t = oracle.table_rowmajor_draw(c["seed"], 1024, 379)
    root, cw, dg = oracle.commit(t, 1024, 379, 4096, want_codewords=True, want_digests=True)
    assert [int(v) for v in root] == c["root"]
    assert int(cw[0]) == c["cw_00"] and int(cw[1]) == c["cw_10"] and int(cw[-1]) == c["cw_last"]
    assert [int(v) for v in dg[0]] == c["leaf0"]


# ---------------------------------------------------------------- planted mistakes
def test_planted_mistakes_are_caught(oracle, golden):
    """Plausible slips in the oracle's data path must break a pin above."""
    c = next(c for c in golden["commit"] if c["n"] == 16 and c["w"] == 9)
    t = oracle.table_rowmajor_draw(c["seed"], 16, 9)
    ok, _, _ = oracle.commit(t, 16, 9, 64)
    assert [int(v) for v in ok] == c["root"]
    # transposed table (row-major read as column-major)
    bad, _, _ = oracle.commit(oracle.splitmix(c["seed"], 16 * 9), 16, 9, 64)
    assert [int(v) for v in bad] != c["root"]
    # wrong coset offset
    bad, _, _ = oracle.commit(t, 16, 9, 64, target_off=3)
    assert [int(v) for v in bad] != c["root"]
    # padding one position off: hash of row without the trailing element
    x = oracle.splitmix(1009, 9)
    assert [int(v) for v in oracle.hash_varlen(x[:8])] != golden["hash_varlen"][9]["out"]


# ---------------------------------------------------------------- against the compiled reference
def test_against_reference_random(oracle, reference):
    rng = np.random.default_rng(9)
    for lg in range(0, 9):
        x = oracle.splitmix(100 + lg, 1 << lg)
        assert np.array_equal(oracle.ntt(x), reference.ntt(x))
        assert np.array_equal(oracle.intt(x), reference.intt(x))
    for _ in range(20):
        st = oracle.splitmix(int(rng.integers(1, 1 << 30)), 16)
        assert np.array_equal(oracle.permutation(st), reference.permutation(st))
        tr = reference.tip5_trace(st)  # Tip5::trace: state before and after each round
        cur = st
        for r in range(5):
            cur = oracle.tip5_round(cur, r)
            assert np.array_equal(cur, tr[r + 1])
    for n, m, k, toff in [(32, 128, 4, 1), (64, 64, 9, 3), (16, 256, 0, 11)]:
        x = oracle.splitmix(n + m, n)
        r = oracle.splitmix(k + 1, k) if k else None
        assert np.array_equal(oracle.lde_column(x, m, toff, 7, r),
                              reference.lde_column(x, m, toff, 7, r))
    t = oracle.table_rowmajor_draw(3, 32, 23)
    a = oracle.commit(t, 32, 23, 128, want_digests=True)
    b = reference.commit(t, 32, 23, 128, want_digests=True)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[2], b[2])


def test_reference_parallel_commit(oracle, reference):
    """The parallel reference commit (the bench's CPU baseline) gives the same root
    as the reference's serial commit and the oracle."""
    t = oracle.table_rowmajor_draw(4, 64, 37)
    reference.set_threads(4)
    par = reference.commit_parallel(t, 64, 37, 256)
    reference.set_threads(1)
    ser = reference.commit(t, 64, 37, 256)[0]
    assert np.array_equal(par, ser)
    assert np.array_equal(par, oracle.commit(t, 64, 37, 256)[0])
