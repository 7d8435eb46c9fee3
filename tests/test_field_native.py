"""Runs the native host check of the Goldilocks primitives (tests/native/gl64_check.cpp,
built by triton-vm-prover_amd/Makefile): the carry-chain add/sub/canon/fold/products the
kernels use, against unsigned __int128 arithmetic."""
import os
import subprocess

import pytest

from conftest import ROOT

BIN = os.path.join(ROOT, "triton-vm-prover_amd", "_build", "gl64_check")


def test_gl64_primitives_match_u128():
    if not os.path.exists(BIN):
        pytest.skip("gl64_check not built (run __graft_entry__.build())")
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "bad=0" in r.stdout
