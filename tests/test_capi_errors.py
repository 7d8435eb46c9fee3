"""Argument errors of the device entry points, on a host without a GPU.

Every `_dev` entry point checks its arguments before it touches the device, so
on a CPU-only host an invalid call returns its own status (and message) while
the same call with valid arguments gets as far as the device and returns
TVH_ERR_DEVICE.  The messages of the Merkle and NTT checks are the reference's
own (src/merkle/merkle_tree.cpp:14,19; src/ntt/ntt.cpp:62).  Buffers are
placeholder addresses: no call here reaches a kernel or dereferences them."""
import os
import subprocess

import pytest

from conftest import ROOT, tvh_module

OK, INVALID, NOT_POW2, DEVICE, LAYOUT, UNSUPPORTED, RANGE = 0, 1, 2, 4, 5, 7, 8
P = 0xFFFFFFFF00000001
BUF = 0x10000  # placeholder device address
BUF2 = 0x20000


@pytest.fixture(scope="module")
def lib():
    mod = tvh_module()
    if mod.hip_rt.device_count() > 0:
        pytest.skip("a GPU is present: these calls are for a host without one")
    if not os.path.exists(mod.LIB_PATH):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "triton-vm-prover_amd"), "-j8"],
                       check=True)
    return mod.load()


def check(lib, want, name, *args, msg=None):
    rc = lib.raw(name, *args)
    assert rc == want, (name, args, rc, lib.last_error())
    if msg is not None:
        assert msg in lib.last_error()


def test_ntt(lib):
    check(lib, NOT_POW2, "tvh_ntt_forward_dev", BUF, 3, None, msg="NTT size must be a power of 2")
    check(lib, INVALID, "tvh_ntt_forward_dev", None, 8, None)
    check(lib, NOT_POW2, "tvh_ntt_inverse_batched_dev", BUF, 12, 5, None)
    check(lib, OK, "tvh_ntt_forward_batched_dev", BUF, 3, 0, None)  # no columns: nothing to do
    check(lib, NOT_POW2, "tvh_ntt_batched_noalloc_dev", BUF, 6, 2, 0, None)
    check(lib, UNSUPPORTED, "tvh_ntt_batched_noalloc_dev", BUF, 1 << 40, 1, 1, None)
    check(lib, DEVICE, "tvh_ntt_forward_dev", BUF, 8, None)


def lde(lib, want, cols=2, n=8, rand=None, k=0, toff=1, goff=7, m=32, out=BUF2):
    check(lib, want, "tvh_lde_batch_dev", BUF, cols, n, rand, k, toff, goff, m, out, None)


def test_lde_batch(lib):
    lde(lib, NOT_POW2, m=24)
    lde(lib, NOT_POW2, n=6, m=32)
    lde(lib, INVALID, n=8, m=4)            # target shorter than the trace
    lde(lib, UNSUPPORTED, n=8, m=8 * 32)   # expansion above 16
    lde(lib, UNSUPPORTED, rand=BUF, k=9)   # more randomizers than trace rows
    lde(lib, INVALID, rand=None, k=3)      # randomizers announced, none given
    lde(lib, INVALID, toff=0)
    lde(lib, INVALID, toff=P)
    lde(lib, INVALID, goff=P)
    lde(lib, UNSUPPORTED, n=1 << 24, m=1 << 25)
    lde(lib, INVALID, out=None)
    lde(lib, OK, cols=0, m=24)             # no columns: nothing to check or do
    lde(lib, DEVICE)
    lde(lib, DEVICE, rand=BUF, k=8, m=8 * 16)


def test_commit(lib):
    args = dict(cols=3, n=8, rand=None, k=0, toff=1, goff=7, m=32)

    def commit(want, layout=0, cw=BUF2, tree=BUF2 + 0x1000, **kw):
        a = dict(args, **kw)
        check(lib, want, "tvh_commit_dev", BUF, a["cols"], a["n"], a["rand"], a["k"], a["toff"],
              a["goff"], a["m"], cw, tree, layout, None)

    commit(LAYOUT, layout=2)
    commit(INVALID, tree=None)
    commit(NOT_POW2, m=48)
    commit(INVALID, toff=0)
    commit(DEVICE)
    commit(DEVICE, layout=1)


def test_two_phase_coset(lib):
    check(lib, NOT_POW2, "tvh_trace_coefficients_dev", BUF, 2, 6, 1, BUF2, None)
    check(lib, INVALID, "tvh_trace_coefficients_dev", BUF, 2, 8, 0, BUF2, None)
    check(lib, INVALID, "tvh_trace_coefficients_dev", BUF, 2, 8, P, BUF2, None)
    check(lib, INVALID, "tvh_trace_coefficients_dev", None, 2, 8, 1, BUF2, None)
    check(lib, NOT_POW2, "tvh_evaluate_coset_dev", BUF, 2, 12, None, 0, 1, 7, BUF2, None, None)
    check(lib, UNSUPPORTED, "tvh_evaluate_coset_dev", BUF, 2, 8, BUF, 9, 1, 7, BUF2, None, None)
    check(lib, INVALID, "tvh_evaluate_coset_dev", BUF, 2, 8, None, 3, 1, 7, BUF2, None, None)
    check(lib, INVALID, "tvh_evaluate_coset_dev", BUF, 2, 8, None, 0, 1, P, BUF2, None, None)
    check(lib, RANGE, "tvh_scatter_digests_strided_dev", BUF, BUF2, 8, 4, 4, None)
    check(lib, INVALID, "tvh_scatter_digests_strided_dev", None, BUF2, 8, 0, 4, None)


def test_transposes_are_not_in_place(lib):
    for name in ("tvh_rowmajor_to_colmajor_bfe_dev", "tvh_colmajor_to_rowmajor_bfe_dev"):
        check(lib, INVALID, name, BUF, BUF, 4, 3, None, msg="not in place")
        check(lib, INVALID, name, None, BUF2, 4, 3, None)


def test_merkle(lib):
    check(lib, INVALID, "tvh_merkle_tree_dev", BUF, BUF2, 0, 0, None,
          msg="Cannot create Merkle tree with no leaves")
    check(lib, NOT_POW2, "tvh_merkle_tree_dev", BUF, BUF2, 6, 0, None,
          msg="Number of leaves must be a power of 2")
    check(lib, LAYOUT, "tvh_merkle_tree_dev", BUF, BUF2, 8, 7, None)
    check(lib, INVALID, "tvh_merkle_tree_dev", None, BUF2, 8, 0, None)
    check(lib, DEVICE, "tvh_merkle_tree_dev", BUF, BUF2, 8, 1, None)
    check(lib, INVALID, "tvh_merkle_root_dev", BUF, BUF2, 0, None)
    check(lib, NOT_POW2, "tvh_merkle_root_dev", BUF, BUF2, 12, None)
    check(lib, DEVICE, "tvh_merkle_root_dev", BUF, BUF2, 16, None)
    check(lib, LAYOUT, "tvh_merkle_gather_dev", BUF, 8, 3, BUF, 1, BUF2, None)
    check(lib, NOT_POW2, "tvh_merkle_gather_dev", BUF, 6, 0, BUF, 1, BUF2, None)
    check(lib, INVALID, "tvh_merkle_gather_dev", BUF, 8, 0, None, 1, BUF2, None)


def test_tip5(lib):
    check(lib, RANGE, "tvh_tip5_rounds_dev", BUF, 4, 3, 3, None, None)
    check(lib, RANGE, "tvh_tip5_rounds_dev", BUF, 4, 6, 0, None, None)
    check(lib, INVALID, "tvh_tip5_rounds_dev", None, 4, 0, 5, None, None)
    check(lib, DEVICE, "tvh_tip5_rounds_dev", BUF, 4, 0, 5, None, None)
    check(lib, INVALID, "tvh_tip5_permutation_batch_dev", None, 4, None)
    check(lib, INVALID, "tvh_hash_bfield_rows_dev", None, 4, 3, BUF2, None)
    check(lib, INVALID, "tvh_hash_bfield_rows_dev", BUF, 4, 3, None, None)
    check(lib, INVALID, "tvh_hash_pairs_dev", BUF, None, BUF2, 4, None)


def test_autotune_shape(lib):
    check(lib, NOT_POW2, "tvh_lde_autotune", 4, 6, 24, None, None)
    check(lib, INVALID, "tvh_lde_autotune", 0, 8, 32, None, None)
