"""Runs the device build of the Goldilocks primitives (tests/native/gl64_device_check.hip,
built by triton-vm-prover_amd/Makefile) on the GPU: add/sub/canon, the 64x64 products,
Montgomery reduction and products, to_mont, the MDS recombination, the general product and
every shift multiply 2^0..2^95, each also composed with canon(), against unsigned __int128
arithmetic on the host.  The host check runs the same source through the host compiler; this
one covers what the gfx950 compiler makes of the carry chains."""
import os
import subprocess

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

BIN = os.path.join(ROOT, "triton-vm-prover_amd", "_build", "gl64_device_check")


def test_gl64_device_primitives_match_u128():
    assert os.path.exists(BIN), "gl64_device_check not built (run __graft_entry__.build())"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "device bad=0" in r.stdout
