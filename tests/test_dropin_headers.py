"""The drop-in C++ headers at the reference's include paths compile where the
reference's own headers would.

The headers under include/ (types/*.hpp, ntt/ntt.hpp, lde/*.hpp,
hash/tip5.hpp, merkle/merkle_tree.hpp, table/table_commitment.hpp) replace the
reference's headers of the same names.  A prover build puts include/ first on
its include path and keeps its own include directory for everything else.
These tests compile (syntax only) translation units of the reference that call
the hot path -- FRI (MerkleTree, Tip5, NTT, ArithmeticDomain), the Fiat-Shamir
challenges and the program hash (Tip5 sponge), the proof stream and the table
padding -- against an include directory that holds the reference's other
headers but none of the replaced ones, so every hot-path declaration they see
comes from this repository.  Skipped when the reference tree is absent.
"""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

REF = "/root/reference"
INCLUDE = os.path.join(ROOT, "include")
# headers this repository replaces (relative to the reference's include/)
REPLACED = {
    "types/b_field_element.hpp", "types/x_field_element.hpp", "types/digest.hpp",
    "ntt/ntt.hpp", "lde/lde_randomized.hpp", "lde/lde_table.hpp", "hash/tip5.hpp",
    "merkle/merkle_tree.hpp", "table/table_commitment.hpp",
}
# reference translation units that call the hot path and compile with g++ alone
CALLERS = ["src/fri/fri.cpp", "src/stark/challenges.cpp", "src/vm/program.cpp",
           "src/proof_stream/proof_stream.cpp", "src/table/table_padding.cpp"]

GXX = shutil.which("g++")


def _compile(src, incs, extra=()):
    cmd = [GXX, "-std=c++20", "-fsyntax-only", "-fopenmp", *extra] + [f"-I{i}" for i in incs] + [src]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    return r.returncode, r.stderr


@pytest.mark.skipif(GXX is None, reason="g++ not found")
def test_headers_compile_standalone(tmp_path):
    """Every drop-in header compiles on its own (no prover headers on the path)."""
    for h in sorted(REPLACED):
        src = tmp_path / "one.cpp"
        src.write_text(f'#include "{h}"\n#include "{h}"\nint main() {{ return 0; }}\n')
        rc, err = _compile(str(src), [INCLUDE])
        assert rc == 0, f"{h}:\n{err[-3000:]}"


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "include")), reason="reference tree absent")
@pytest.mark.skipif(GXX is None, reason="g++ not found")
def test_reference_callers_compile(tmp_path):
    # the reference's include/ minus the replaced headers, as symlinks
    other = tmp_path / "ref_other_include"
    other.mkdir()
    ref_inc = os.path.join(REF, "include")
    for entry in os.listdir(ref_inc):
        p = os.path.join(ref_inc, entry)
        if os.path.isdir(p):
            (other / entry).mkdir()
            for f in os.listdir(p):
                if f"{entry}/{f}" not in REPLACED:
                    os.symlink(os.path.join(p, f), other / entry / f)
        else:
            os.symlink(p, other / entry)
    assert not (other / "hash" / "tip5.hpp").exists()
    failures = []
    for tu in CALLERS:
        rc, err = _compile(os.path.join(REF, tu), [INCLUDE, str(other)])
        if rc:
            failures.append(f"{tu}:\n{err[-2000:]}")
    assert not failures, "\n\n".join(failures)
