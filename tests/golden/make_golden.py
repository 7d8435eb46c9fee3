"""Write tests/golden/golden.json from the reference's own CPU code.

Runs in the build container only (it needs oracle/_ref/libtvm_ref.so, compiled
from the reference sources by `make -C oracle ref`).  The fixture holds inputs
(as SplitMix64 seeds or explicit values) and the reference's outputs, so that
tests on any machine can pin the oracle and the HIP path without the reference.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
from oracle import Oracle, Reference  # noqa: E402

P = 0xFFFFFFFF00000001


def ints(a):
    return [int(x) for x in np.asarray(a).reshape(-1)]


def digest_checksum(d):
    """SHA-256 over the little-endian bytes of a uint64 array."""
    return hashlib.sha256(np.ascontiguousarray(d, dtype="<u8").tobytes()).hexdigest()


def main() -> None:
    if not Reference.available():
        sys.exit("oracle/_ref/libtvm_ref.so is not built (make -C oracle ref)")
    ref, orc = Reference(), Oracle()
    ref.set_threads(1)
    g: dict = {"p": P, "source": "reference CPU sources compiled by oracle/Makefile"}

    # Tip5 permutation on edge and random states.
    states = [np.zeros(16, np.uint64), np.full(16, P - 1, np.uint64),
              np.arange(16, dtype=np.uint64)]
    states += [orc.splitmix(500 + i, 16) for i in range(5)]
    g["tip5_permutation"] = [{"in": ints(s), "out": ints(ref.permutation(s))} for s in states]

    # hash_varlen over lengths that hit every padding position.
    g["hash_varlen"] = []
    for ln in list(range(0, 26)) + [156, 264, 379]:
        x = orc.splitmix(1000 + ln, ln)
        g["hash_varlen"].append({"seed": 1000 + ln, "len": ln, "out": ints(ref.hash_varlen(x))})

    # hash_pair
    g["hash_pair"] = []
    for i in range(4):
        l, r = orc.splitmix(2000 + i, 5), orc.splitmix(3000 + i, 5)
        g["hash_pair"].append({"seed_l": 2000 + i, "seed_r": 3000 + i,
                               "out": ints(ref.hash_pair(l, r))})

    # NTT forward / inverse, n = 1 .. 2^10.
    g["ntt"] = []
    for lg in range(0, 11):
        n = 1 << lg
        x = orc.splitmix(4000 + lg, n)
        g["ntt"].append({"seed": 4000 + lg, "n": n, "fwd_sha256": digest_checksum(ref.ntt(x)),
                         "inv_sha256": digest_checksum(ref.intt(x)),
                         "fwd_head": ints(ref.ntt(x)[:4]), "inv_head": ints(ref.intt(x)[:4])})

    # Randomized LDE columns (BFE), including the chunked case n+k > m.
    g["lde_column"] = []
    for (n, m, k, toff, goff, seed) in [(16, 64, 0, 1, 7, 1), (16, 64, 3, 1, 7, 2),
                                        (16, 64, 5, 3, 7, 3), (32, 32, 4, 1, 7, 4),
                                        (8, 16, 12, 5, 7, 5), (256, 1024, 7, 1, 7, 6),
                                        (1, 4, 0, 1, 7, 7), (1, 1, 0, 1, 7, 8)]:
        x = orc.splitmix(5000 + seed, n)
        r = orc.splitmix(6000 + seed, k) if k else None
        out = ref.lde_column(x, m, toff, goff, r)
        g["lde_column"].append({"n": n, "m": m, "k": k, "trace_off": toff, "target_off": goff,
                                "seed_trace": 5000 + seed, "seed_rand": 6000 + seed,
                                "out": ints(out)})

    # XFE columns (truncating evaluation), BFE-lifted and XFE randomizers.
    g["lde_xfe_column"] = []
    for (n, m, k, xr, toff, seed) in [(16, 64, 0, 0, 1, 1), (16, 64, 3, 0, 1, 2),
                                      (16, 64, 3, 1, 3, 3), (16, 16, 5, 1, 1, 4),
                                      (16, 16, 5, 0, 1, 5)]:
        x = orc.splitmix(7000 + seed, 3 * n).reshape(n, 3)
        r = None
        if k:
            r = orc.splitmix(8000 + seed, 3 * k if xr else k)
            r = r.reshape(k, 3) if xr else r
        out = ref.lde_xfe_column(x, m, toff, 7, r, bool(xr))
        g["lde_xfe_column"].append({"n": n, "m": m, "k": k, "xfe_rand": xr, "trace_off": toff,
                                    "target_off": 7, "seed_trace": 7000 + seed,
                                    "seed_rand": 8000 + seed, "out": ints(out)})

    # Whole commits: table drawn row-major from SplitMix64(seed), expansion 4, offset 7.
    g["commit"] = []
    for (n, w, seed) in [(8, 7, 1), (1024, 379, 1), (1024, 264, 1), (16, 1, 11), (16, 9, 12),
                         (16, 10, 13), (16, 11, 14), (64, 156, 15), (256, 379, 16), (4, 3, 17)]:
        m = 4 * n
        t = orc.table_rowmajor_draw(seed, n, w)
        root, cw, dg = ref.commit(t, n, w, m, want_codewords=True, want_digests=True)
        g["commit"].append({
            "n": n, "w": w, "m": m, "seed": seed, "root": ints(root),
            "cw_00": int(cw[0]), "cw_10": int(cw[1]), "cw_last": int(cw[-1]),
            "leaf0": ints(dg[0]), "leaf_last": ints(dg[-1]),
            "codewords_sha256": digest_checksum(cw), "digests_sha256": digest_checksum(dg)})

    # Merkle heap and authentication structures.
    leaves = orc.splitmix(9000, 64 * 5).reshape(64, 5)
    root, nodes = ref.merkle(leaves, want_nodes=True)
    g["merkle"] = {"seed": 9000, "n": 64, "root": ints(root), "nodes_sha256": digest_checksum(nodes)}
    g["auth_structure"] = []
    for idx in [[0], [63], [0, 1], [5, 9, 40], [3, 3], list(range(0, 64, 7))]:
        s = ref.auth_structure(leaves, idx)
        g["auth_structure"].append({"indices": idx, "digests": ints(s)})

    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden.json")
    with open(path, "w") as f:
        json.dump(g, f, indent=0, separators=(",", ":"))
    print(f"wrote {path} ({os.path.getsize(path)} bytes)")


if __name__ == "__main__":
    main()
