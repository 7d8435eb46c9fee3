"""FRI codeword commitment (XFE leaves (c0, c1, c2, 0, 0), gpu_stark.cu:863-875) and the
device Fiat-Shamir sponge (fiat_shamir_kernel.cu: pad [data, 1, 0...] to a multiple of
10, overwrite the rate, permute; squeeze = rate then permute), against the oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RATE = 10


def oracle_absorb(oracle, state, data):
    """pad_and_absorb_all on a canonical 16-word state, composed from the oracle's
    permutation (tip5.cpp:92-203)."""
    st = [int(v) for v in state]
    padded = [int(v) for v in data] + [1]
    while len(padded) % RATE:
        padded.append(0)
    for b in range(0, len(padded), RATE):
        st[:RATE] = padded[b:b + RATE]
        st = [int(v) for v in oracle.permutation(st)]
    return st


@pytest.mark.parametrize("length", [0, 1, 6, 9, 10, 11, 25, 100])
def test_sponge_absorb_matches_hash_varlen(T, oracle, length):
    """From the zero state, absorbing x leaves hash_varlen(x) in state[0..5)."""
    x = oracle.splitmix(800 + length, length)
    state = T.empty(16)
    T.lib.sponge_init(state)
    T.lib.sponge_absorb(state, T.dev(x) if length else None, length)
    got = T.host(state, 16)
    assert [int(v) for v in got[:5]] == [int(v) for v in oracle.hash_varlen(x)]
    assert [int(v) for v in got] == oracle_absorb(oracle, [0] * 16, x)


def test_sponge_transcript_sequence(T, oracle):
    """absorb, absorb, squeeze, absorb, squeeze on one device state equals the same
    sequence composed on the CPU; the Merkle-root item [0, root] included."""
    state = T.empty(16)
    T.lib.sponge_init(state)
    want = [0] * 16
    out = T.empty(RATE)
    root = oracle.splitmix(77, 5)
    item = np.concatenate([[0], root]).astype(np.uint64)
    seq = [("a", item), ("a", oracle.splitmix(78, 23)), ("s", None), ("a", oracle.splitmix(79, 4)),
           ("s", None)]
    for op, data in seq:
        if op == "a":
            T.lib.sponge_absorb(state, T.dev(data), data.size)
            want = oracle_absorb(oracle, want, data)
        else:
            T.lib.sponge_squeeze(state, out)
            got_out = [int(v) for v in T.host(out, RATE)]
            assert got_out == want[:RATE]
            want = [int(v) for v in oracle.permutation(want)]
        assert [int(v) for v in T.host(state, 16)] == want


@pytest.mark.parametrize("logn,layout", [(0, 0), (1, 1), (5, 0), (10, 1), (14, 0)])
def test_fri_commit_matches_oracle(T, oracle, logn, layout):
    n = 1 << logn
    xfe = oracle.splitmix(900 + logn, 3 * n)
    leaves = np.zeros((n, 5), np.uint64)
    leaves[:, :3] = xfe.reshape(n, 3)
    d_leaves = T.empty(5 * n)
    T.lib.xfe_leaves(T.dev(xfe), n, d_leaves)
    assert np.array_equal(T.host(d_leaves).reshape(n, 5), leaves)
    nodes = 2 * n if layout == 1 else 2 * n - 1
    tree = T.empty(5 * nodes)
    T.lib.fri_commit(T.dev(xfe), n, tree, layout)
    got = T.host(tree).reshape(-1, 5)
    if layout == 1:
        assert np.array_equal(got[1:], oracle.merkle_heap(leaves)[1:])
    else:
        assert np.array_equal(got, oracle.merkle_level_order(leaves))


def test_fri_commit_rejects_bad_length(T, tvhmod):
    with pytest.raises(tvhmod.TVHError) as e:
        T.lib.fri_commit(T.empty(9), 3, T.empty(25))
    assert e.value.status == 2
