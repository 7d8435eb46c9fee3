"""C-ABI checks that need no GPU: the library loads, exports every entry point
include/tvh.h declares, and its host-only helpers match the oracle."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, tvh_module


@pytest.fixture(scope="module")
def lib():
    mod = tvh_module()
    if not os.path.exists(mod.LIB_PATH):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "triton-vm-prover_amd"), "-j8"],
                       check=True)
    return mod.load()


def test_exports_every_declared_symbol(lib):
    names = tvh_module().declared_symbols()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib.lib, n)]
    assert missing == []


def test_header_declares_extern_c_only(lib):
    text = open(os.path.join(ROOT, "include", "tvh.h")).read()
    assert 'extern "C"' in text
    for banned in ("std::", "torch", "at::", "class "):
        assert banned not in text


def test_status_strings(lib):
    assert lib.status_string(0) == "ok"
    assert "power of two" in lib.status_string(2)
    assert lib.status_string(999) == "unknown status"


def test_auth_path_indices(lib):
    assert list(lib.auth_path_indices(8, 0)) == [9, 5, 3]
    assert list(lib.auth_path_indices(8, 7)) == [14, 6, 2]
    assert list(lib.auth_path_indices(1, 0)) == []
    rc = lib.raw("tvh_auth_path_indices", 8, 8, None, 0, None)
    assert rc == 8  # TVH_ERR_RANGE


def test_auth_structure_indices_match_oracle_and_golden(lib, oracle, golden):
    n = golden["merkle"]["n"]
    for c in golden["auth_structure"]:
        got = lib.auth_structure_indices(n, c["indices"])
        want = oracle.auth_structure_indices(n, c["indices"])
        assert [int(x) for x in got] == [int(x) for x in want]
        assert len(got) * 5 == len(c["digests"])
    rng = np.random.default_rng(4)
    for n in [1, 2, 64, 1 << 12]:
        for cnt in [0, 1, 5, 80]:
            idx = rng.integers(0, n, cnt)
            assert [int(x) for x in lib.auth_structure_indices(n, idx)] == \
                   [int(x) for x in oracle.auth_structure_indices(n, idx)]


def test_auth_structure_errors(lib):
    idx = np.array([3], dtype=np.uintp)
    assert lib.raw("tvh_auth_structure_indices", 6, idx.ctypes.data, 1, None, 0, None) == 2
    assert lib.raw("tvh_auth_structure_indices", 0, idx.ctypes.data, 1, None, 0, None) == 1
    idx = np.array([9], dtype=np.uintp)
    assert lib.raw("tvh_auth_structure_indices", 8, idx.ctypes.data, 1, None, 0, None) == 8


def test_init_fails_loudly_without_a_gpu(lib):
    if tvh_module().hip_rt.device_count() > 0:
        pytest.skip("a GPU is present")
    assert lib.tvh_init(0) == 4  # TVH_ERR_DEVICE
    rc = lib.raw("tvh_ntt_forward_host", np.zeros(4, np.uint64).ctypes.data, 4)
    assert rc == 4
