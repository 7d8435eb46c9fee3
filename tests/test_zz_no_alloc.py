"""tvh_ntt_batched_noalloc_dev: the batched NTT through the library's static
device scratch, with no allocation.  Checked against the allocating entry
points (tvh_ntt_{forward,inverse}_batched_dev) bit for bit, on two calls in a
row (the second on a smaller input, over the scratch the first one left), and
against the oracle on a small case.  Collected last, so that a failure here
hides no other test."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_N = 1 << 22  # the static scratch holds 2^22 u64


def _run(T, fn, x, n, cols, inverse):
    d = T.dev(x)
    fn(d, n, cols, inverse)
    T.sync()
    return T.host(d)


@pytest.mark.parametrize("inverse", [False, True])
def test_noalloc_matches_allocating_entry_point(T, oracle, inverse):
    lib = T.lib
    ref = (lambda d, n, c, inv: lib.ntt_inverse_batched(d, n, c)) if inverse else \
          (lambda d, n, c, inv: lib.ntt_forward_batched(d, n, c))
    # first call: 2^20 x 7 columns (several groups through the 2^22 scratch),
    # second call: a smaller input (2^10 x 3) over what the first left behind
    for n, cols, seed in [(1 << 20, 7, 11), (1 << 10, 3, 12)]:
        print(f"noalloc inverse={inverse} n={n} cols={cols}", flush=True)
        x = oracle.splitmix(seed, n * cols)
        got = _run(T, lib.ntt_batched_noalloc, x, n, cols, inverse)
        want = _run(T, ref, x, n, cols, inverse)
        assert np.array_equal(got, want), (n, cols)
        if n <= 1 << 10:
            f = oracle.intt if inverse else oracle.ntt
            for c in range(cols):
                assert np.array_equal(got[c * n:(c + 1) * n], f(x[c * n:(c + 1) * n])), c


def test_noalloc_largest_and_rejects(T, tvhmod, oracle):
    lib = T.lib
    x = oracle.splitmix(21, MAX_N)
    got = _run(T, lib.ntt_batched_noalloc, x, MAX_N, 1, False)
    want = _run(T, lambda d, n, c, inv: lib.ntt_forward_batched(d, n, c), x, MAX_N, 1, False)
    assert np.array_equal(got, want)
    d = T.empty(2 * MAX_N)
    with pytest.raises(tvhmod.TVHError) as e:
        lib.ntt_batched_noalloc(d, 2 * MAX_N, 1, False)
    assert e.value.status == 7  # TVH_ERR_UNSUPPORTED
    with pytest.raises(tvhmod.TVHError) as e:
        lib.ntt_batched_noalloc(d, 3, 1, False)
    assert e.value.status == 2  # TVH_ERR_NOT_POW2
