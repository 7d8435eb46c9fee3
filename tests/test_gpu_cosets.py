"""Two-phase coset evaluation (trace coefficients once, then one coset of the
FRI domain at a time), the coset digest scatter, the row-major <-> column-major
transposes, and the prover's real LDE form (98 randomizers, FRI domain 8 N) at
2^20 x 379 -- all against the oracle and against the plain commit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

P = 0xFFFFFFFF00000001


@pytest.mark.parametrize("n,cols,toff,inplace", [(1, 2, 1, False), (16, 3, 1, True),
                                                (64, 5, 3, False), (1024, 4, 11, True),
                                                (1 << 12, 3, 1, False), (1 << 13, 2, 7, False)])
def test_trace_coefficients(T, oracle, n, cols, toff, inplace):
    x = oracle.splitmix(70 + n, n * cols)
    d = T.dev(x)
    out = d if inplace else T.empty(n * cols)
    T.lib.trace_coefficients(d, cols, n, toff, out)
    got = T.host(out)
    if not inplace:  # out of place, the trace is only read
        assert np.array_equal(T.host(d), np.asarray(x, dtype=np.uint64))
    oinv = oracle.inv(toff)
    for c in range(cols):
        coef = oracle.intt(x[c * n:(c + 1) * n])
        s = 1
        for i in range(n):
            assert int(got[c * n + i]) == oracle.mul(int(coef[i]), s), (c, i)
            s = oracle.mul(s, oinv)


@pytest.mark.parametrize("n,cols,K,k,toff", [(16, 3, 4, 0, 1), (64, 2, 8, 5, 1), (32, 4, 2, 32, 3),
                                             (256, 3, 8, 98, 1), (8, 2, 16, 3, 7),
                                             (1, 2, 4, 1, 1), (2, 3, 8, 2, 3),
                                             (1 << 12, 2, 4, 1 << 12, 3), (1 << 18, 2, 4, 98, 1)])
def test_evaluate_cosets_equal_the_lde(T, oracle, n, cols, K, k, toff):
    """Coset kk of the M = K n point domain 7 <w_M> is 7 w_M^kk <w_n>: its row j is
    row j K + kk of the oracle's randomized LDE."""
    m = K * n
    x = oracle.splitmix(80 + n, n * cols)
    r = oracle.splitmix(81 + k, k * cols) if k else None
    coeffs = T.empty(n * cols)
    T.lib.trace_coefficients(T.dev(x), cols, n, toff, coeffs)
    want = oracle.lde_table(x, n, cols, m, toff, 7, None if r is None else r.reshape(cols, k))
    want = want.reshape(cols, m)
    wm = oracle.root_of_unity(m.bit_length() - 1)
    out = T.empty(n * cols)
    dr = T.dev(r) if k else None
    for kk in range(K):
        g = oracle.mul(7, oracle.pow(wm, kk))
        T.lib.evaluate_coset(coeffs, cols, n, dr, k, toff, g, out)
        got = T.host(out).reshape(cols, n)
        assert np.array_equal(got, want[:, kk::K]), kk


@pytest.mark.parametrize("n,cols", [(1 << 10, 3), (1 << 20, 300)])
def test_evaluate_coset_in_place(T, oracle, n, cols):
    """d_output == d_coefficients gives the out-of-place values.  At 2^20 x 300
    the NTT runs in two column groups (256 + 44 columns), each read before it
    is written; the buffers are compared through their row digests."""
    k, toff = 7, 1
    tr, r = T.empty(n * cols), T.empty(k * cols)
    T.lib.fill_uniform(tr, n * cols, 500 + n)
    T.lib.fill_uniform(r, k * cols, 501)
    coeffs = T.empty(n * cols)
    T.lib.trace_coefficients(tr, cols, n, toff, coeffs)
    del tr
    g = oracle.mul(7, oracle.root_of_unity(3))

    def digests(buf):
        d = T.empty(5 * n)
        T.lib.hash_bfield_rows(buf, n, cols, d)
        return T.host(d)

    before = digests(coeffs)
    out = T.empty(n * cols)
    T.lib.evaluate_coset(coeffs, cols, n, r, k, toff, g, out)
    T.lib.evaluate_coset(coeffs, cols, n, r, k, toff, g, coeffs)
    want = digests(out)
    assert np.array_equal(digests(coeffs), want)
    assert not np.array_equal(want, before)


@pytest.mark.parametrize("lg", [22, 23])
def test_evaluate_cosets_equal_the_lde_large(T, oracle, lg):
    """At 2^22 and 2^23 rows (the batched NTT's large splits: 2^13 column tiles at
    2^22, 2^14 tiles and capped rows at 2^23) every coset of the 4N-point domain,
    98 randomizers, equals the matching rows of the LDE path, whose kernels are
    separate and oracle-checked."""
    n, cols, K, k, toff = 1 << lg, 2, 4, 98, 1
    m = K * n
    tr, r = T.empty(n * cols), T.empty(k * cols)
    T.lib.fill_uniform(tr, n * cols, 700 + lg)
    T.lib.fill_uniform(r, k * cols, 701)
    cw = T.empty(m * cols)
    T.lib.lde_batch(tr, cols, n, r, k, toff, 7, m, cw)
    want = T.host(cw).reshape(cols, m)
    del cw
    coeffs, out = T.empty(n * cols), T.empty(n * cols)
    T.lib.trace_coefficients(tr, cols, n, toff, coeffs)
    wm = oracle.root_of_unity(lg + 2)
    for kk in range(K):
        g = oracle.mul(7, oracle.pow(wm, kk))
        T.lib.evaluate_coset(coeffs, cols, n, r, k, toff, g, out)
        assert np.array_equal(T.host(out).reshape(cols, n), want[:, kk::K]), kk


def _two_phase_commit(T, dt, w, n, K, dr, k, toff, goff):
    """The frugal caller's sequence (gpu_stark.cu:2356-2420): coefficients once,
    then per coset: evaluate, hash rows, scatter the digests at stride K; then
    the Merkle tree over the scattered leaves."""
    lib = T.lib
    m = K * n
    coeffs = T.empty(n * w)
    lib.trace_coefficients(dt, w, n, toff, coeffs)
    work, dig = T.empty(n * w), T.empty(5 * n)
    tree = T.empty(5 * (2 * m - 1))
    wm = pow_root(m)
    for kk in range(K):
        g = mulmod(goff, pow(wm, kk, P))
        lib.evaluate_coset(coeffs, w, n, dr, k, toff, g, work)
        lib.hash_bfield_rows(work, n, w, dig)
        lib.scatter_digests(dig, tree, n, kk, K)
    lib.merkle_tree(tree, tree, m, 0)
    T.sync()
    return tree


def mulmod(a, b):
    return a * b % P


def pow_root(m):
    return pow(7, (P - 1) // m, P)


def test_two_phase_commit_matches_plain_small(T, oracle):
    n, w, K, k = 64, 23, 8, 5
    m = K * n
    t = oracle.table_rowmajor_draw(61, n, w)
    r = oracle.splitmix(62, w * k)
    dt, dr = T.dev(t), T.dev(r)
    tree = _two_phase_commit(T, dt, w, n, K, dr, k, 1, 7)
    root, _, dg = oracle.commit(t, n, w, m, rand=r.reshape(w, k), want_digests=True)
    got = T.host(tree).reshape(-1, 5)
    assert np.array_equal(got[:m], dg)
    assert np.array_equal(got[-1], root)


def test_scatter_digests(T, oracle):
    n, K = 1000, 8
    d = [oracle.splitmix(90 + kk, 5 * n) for kk in range(K)]
    leaves = T.empty(5 * n * K)
    for kk in range(K):
        T.lib.scatter_digests(T.dev(d[kk]), leaves, n, kk, K)
    got = T.host(leaves).reshape(n, K, 5)
    for kk in range(K):
        assert np.array_equal(got[:, kk, :], d[kk].reshape(n, 5))


@pytest.mark.parametrize("n,width", [(1, 1), (1, 379), (37, 5), (64, 64), (1000, 379),
                                     (4099, 130), (1 << 16, 3)])
def test_transposes(T, oracle, n, width):
    x = oracle.splitmix(n + width, n * width)
    src = T.dev(x)
    cm = T.empty(n * width)
    T.lib.transpose(src, cm, n, width, to_colmajor=True)
    assert np.array_equal(T.host(cm), x.reshape(n, width).T.reshape(-1))
    back = T.empty(n * width)
    T.lib.transpose(cm, back, n, width, to_colmajor=False)
    assert np.array_equal(T.host(back), x)
    if width % 3 == 0:  # XFE forms: xwidth XFE columns of 3 interleaved limbs
        xcm = T.empty(n * width)
        T.lib.transpose(src, xcm, n, width // 3, to_colmajor=True, xfe=True)
        assert np.array_equal(T.host(xcm), T.host(cm))


def test_prover_lde_form_full_size(T, oracle):
    """The prover's main-table LDE: 2^20 x 379, 98 randomizers per column, FRI
    domain 8 N (2^23 rows).  Every column at sampled rows against the oracle's
    randomized interpolant; sampled row digests against the oracle's hash; the
    two-phase 8-coset commit's whole tree equals the plain commit's."""
    n, w, E, k = 1 << 20, 379, 8, 98
    m = E * n
    t = oracle.table_rowmajor_draw(1, n, w)
    r = oracle.splitmix(98, w * k)
    dt, dr = T.dev(t), T.dev(r)
    cw = T.empty(m * w)
    tree = T.empty(5 * (2 * m - 1))
    T.lib.commit(dt, w, n, dr, k, 1, 7, m, cw, tree, 0)
    T.sync()
    rng = np.random.default_rng(8)
    rows = [0, 1, 7, 8, m // 2 + 3, m - 1] + [int(v) for v in rng.integers(0, m, 2)]
    want = oracle.lde_sample(t, n, w, m, rows, rand=r.reshape(w, k))
    for q, row in enumerate(rows):
        vals = cw.gather_strided(row, m, w)
        assert np.array_equal(vals, want[:, q]), row
        assert np.array_equal(tree.download(5, 5 * row), oracle.hash_varlen(vals)), row
    plain = T.host(tree)
    del cw, tree
    tree2 = _two_phase_commit(T, dt, w, n, E, dr, k, 1, 7)
    assert np.array_equal(T.host(tree2), plain)
