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


ARENA = os.path.join(ROOT, "triton-vm-prover_amd", "_build", "arena_check")


def test_server_arena_first_fit_matches_model():
    """The commit server's HBM admission allocator (csrc/arena.hpp) against a brute-force
    granule-bitmap model: same offsets, used bytes, hole count and largest hole over
    random alloc/release sequences; everything coalesces back to one hole."""
    if not os.path.exists(ARENA):
        pytest.skip("arena_check not built (run __graft_entry__.build())")
    r = subprocess.run([ARENA], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert '"ok": true' in r.stdout
