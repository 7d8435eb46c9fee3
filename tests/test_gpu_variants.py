"""Every LDE kernel variant (row-step form TVH_LDE_WIDE / TVH_LDE_SPLIT times the
column-step tile 2^14 / 2^13 / 2^12, and the split form with 2^13-element tiles
for both row kernels or the inverse rows only, or 2^12-element row tiles) against
the oracle on every shape class (rows
shorter than, equal to and split from 4096; expansions 1..16; randomizers; trace
offsets), in single-table calls and inside batch plans, and the autotuner's contract."""
import numpy as np
import pytest

from test_gpu_parity import LDE_CASES

pytestmark = pytest.mark.gpu

WIDE, SPLIT, HALF, IHALF, QUARTER = 1, 2, 16, 32, 64
# form | tile code << 2 | smaller row tiles (2^13 for both row kernels or the inverse
# rows only, 2^12 for both)
VARIANTS = (1, 2, 5, 6, 9, 10, 18, 22, 26, 34, 38, 42, 66, 70, 74)
EXTRA = [(20, 4, 0, 1), (18, 2, 5, 3), (17, 16, 0, 1), (19, 1, 0, 1)]


@pytest.fixture
def variant(T):
    """Sets the form for a shape; restores the default afterwards."""
    touched = []

    def set_(logn, E, v):
        T.lib.lde_set_variant(logn, E.bit_length() - 1, v)
        touched.append((logn, E.bit_length() - 1))

    yield set_
    for ln, le in touched:
        T.lib.lde_set_variant(ln, le, 0)


@pytest.mark.parametrize("form", VARIANTS)
@pytest.mark.parametrize("logn,E,k,toff", LDE_CASES)
def test_lde_forms_match_oracle(T, oracle, variant, form, logn, E, k, toff):
    variant(logn, E, form)
    n, cols = 1 << logn, 3
    m = E * n
    x = oracle.splitmix(5000 + logn, n * cols)
    r = oracle.splitmix(6000 + k, k * cols) if k else None
    out = T.empty(m * cols)
    T.lib.lde_batch(T.dev(x), cols, n, T.dev(r) if k else None, k, toff, 7, m, out)
    want = oracle.lde_table(x, n, cols, m, toff, 7, None if r is None else r.reshape(cols, k))
    assert np.array_equal(T.host(out), want)


@pytest.mark.parametrize("logn,E,k,toff", EXTRA)
def test_lde_forms_agree_large(T, oracle, variant, logn, E, k, toff):
    """Large shapes: the two forms give identical codewords, and sampled codewords
    match the oracle's direct evaluation."""
    n, cols = 1 << logn, 2
    m = E * n
    x = T.empty(n * cols)
    T.lib.fill_uniform(x, n * cols, 40 + logn)
    r = None
    if k:
        r = T.empty(k * cols)
        T.lib.fill_uniform(r, k * cols, 50 + logn)
    outs = []
    for form in VARIANTS:
        variant(logn, E, form)
        out = T.empty(m * cols)
        T.lib.lde_batch(x, cols, n, r, k, toff, 7, m, out)
        outs.append(T.host(out))
    for o in outs[1:]:
        assert np.array_equal(outs[0], o)
    # sampled check of column 1 against the oracle's coset evaluation of the
    # interpolant (the full column LDE on the CPU is a few seconds at these sizes)
    col = T.host(x)[n:2 * n]
    rr = None if r is None else T.host(r)[k:2 * k]
    want = oracle.lde_column(col, m, toff, 7, rr)
    assert np.array_equal(outs[0][m:2 * m], want)


@pytest.mark.parametrize("logn", [17, 18, 19, 20, 21])
def test_every_variant_full_size_against_oracle(T, oracle, tvhmod, variant, logn):
    """Every variant the autotuner can pick, at the heights it picks them for
    (2^17..2^21 rows, expansion 4): all columns at sampled rows against the
    oracle's interpolant evaluated at 7 w_M^j."""
    n, cols, E = 1 << logn, 4, 4
    m = E * n
    x = oracle.splitmix(900 + logn, n * cols)
    dx = T.dev(x)
    rows = [0, 1, 4097, m // 3, m - 2, m - 1]
    want = oracle.lde_sample(x, n, cols, m, rows)
    out = T.empty(m * cols)
    ran = 0
    for form in VARIANTS:
        variant(logn, E, form)
        try:
            T.lib.lde_batch(dx, cols, n, None, 0, 1, 7, m, out)
        except tvhmod.TVHError as e:
            assert e.status == 7, e  # TVH_ERR_UNSUPPORTED: this variant cannot run the shape
            continue
        T.sync()
        for q, r in enumerate(rows):
            assert np.array_equal(out.gather_strided(r, m, cols), want[:, q]), (form, r)
        ran += 1
    assert ran >= 12


def test_batch_plan_with_both_forms(T, oracle, variant):
    """A batch whose tables use different forms equals per-table commits."""
    shapes = [(10, 5, 4, WIDE), (12, 3, 4, SPLIT), (13, 7, 2, SPLIT | 4), (8, 11, 8, WIDE | 8),
              (14, 2, 4, SPLIT | 8), (11, 4, 4, SPLIT), (15, 3, 2, WIDE | 4),
              (16, 2, 4, SPLIT | HALF), (9, 3, 16, SPLIT | 4 | HALF), (17, 2, 4, SPLIT | IHALF),
              (18, 2, 4, SPLIT | QUARTER), (6, 5, 16, SPLIT | 8 | QUARTER)]
    tabs = []
    for i, (lg, w, E, form) in enumerate(shapes):
        variant(lg, E, form)
        n, m = 1 << lg, E << lg
        tr = T.empty(n * w)
        T.lib.fill_uniform(tr, n * w, 700 + i)
        tabs.append(dict(n=n, m=m, w=w, tr=tr, cw=T.empty(m * w), tree=T.empty(5 * (2 * m - 1))))
    descs = [T.lib.table_desc(t["tr"], t["w"], t["n"], t["m"], t["cw"], t["tree"]) for t in tabs]
    T.lib.commit_batch(descs)
    T.sync()
    for t, (lg, w, E, form) in zip(tabs, shapes):
        x = T.host(t["tr"])
        root, cw, _ = oracle.commit(x, t["n"], w, t["m"], want_codewords=True)
        assert np.array_equal(T.host(t["cw"]), cw), (lg, w, E, form)
        assert np.array_equal(T.host(t["tree"]).reshape(-1, 5)[-1], root)


def test_variant_codes(T, tvhmod):
    """The library accepts exactly the listed codes (and 0, the default)."""
    assert T.lib.lib.tvh_lde_variant_count() == len(VARIANTS)
    for v in range(160):
        ok = v == 0 or v in VARIANTS
        if ok:
            T.lib.lde_set_variant(13, 2, v)
        else:
            with pytest.raises(tvhmod.TVHError):
                T.lib.lde_set_variant(13, 2, v)
    T.lib.lde_set_variant(13, 2, 0)


def test_autotune_contract(T, variant):
    chosen, ms = T.lib.lde_autotune(8, 1 << 14, 1 << 16)
    variant(14, 4, chosen)  # restore afterwards
    assert chosen in VARIANTS
    assert T.lib.lde_get_variant(14, 2, cols=8) == chosen
    assert all(v > 0 for v in ms.values())
    assert ms[chosen] == min(ms.values())


def test_autotune_keyed_by_width(T, variant):
    """A choice is kept for the width class it was tuned on (< 64, 64..255,
    >= 256 columns); setting a variant sets every class."""
    variant(15, 4, 0)  # restored to the default afterwards
    T.lib.lde_set_variant(15, 2, VARIANTS[0])
    assert [T.lib.lde_get_variant(15, 2, cols=c) for c in (1, 100, 379)] == [VARIANTS[0]] * 3
    chosen, _ = T.lib.lde_autotune(300, 1 << 15, 1 << 17)
    assert T.lib.lde_get_variant(15, 2, cols=379) == chosen
    assert T.lib.lde_get_variant(15, 2) == chosen  # the >= 256 class
    assert T.lib.lde_get_variant(15, 2, cols=1) == VARIANTS[0]
    assert T.lib.lde_get_variant(15, 2, cols=100) == VARIANTS[0]
    chosen8, _ = T.lib.lde_autotune(8, 1 << 15, 1 << 17)
    assert T.lib.lde_get_variant(15, 2, cols=8) == chosen8
    assert T.lib.lde_get_variant(15, 2, cols=379) == chosen


def test_autotune_skips_unsupported_variants(T, variant):
    """At 2^21 rows and expansion 16 the 2^12-element row tiles cannot hold a
    row's cosets: those variants are skipped (time -1), the rest are timed."""
    chosen, ms = T.lib.lde_autotune(2, 1 << 21, 1 << 25)
    variant(21, 16, chosen)
    assert chosen in VARIANTS and ms[chosen] > 0
    skipped = [v for v, t in ms.items() if t < 0]
    assert skipped and all(v & QUARTER for v in skipped)
    assert ms[chosen] == min(t for t in ms.values() if t > 0)
