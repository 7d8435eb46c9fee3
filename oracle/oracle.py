"""ctypes bindings for the CPU checkers.

TEST INFRASTRUCTURE ONLY: the oracle (liboracle.so, a plain-C restatement of the
reference CPU path) and, where it was built, oracle/_ref/libtvm_ref.so (the
reference's own CPU sources).  Only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py import this module; the library under
triton-vm-prover_amd/ never does.

All arrays are numpy uint64 with canonical Goldilocks values.  Tables are
column-major: element (row r, column c) of an n-row table is at c*n + r.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
P = 0xFFFFFFFF00000001

_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_szp = np.ctypeslib.ndpointer(dtype=np.uintp, flags="C_CONTIGUOUS")


def build(with_ref: bool = True) -> None:
    """Compile liboracle.so (and _ref/libtvm_ref.so when the reference is present)."""
    target = [] if with_ref else ["oracle"]
    subprocess.run(["make", "-s", "-C", HERE] + target, check=True)


def _load(path: str) -> C.CDLL:
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is not built; run make -C {HERE}")
    return C.CDLL(path)


class Oracle:
    """The C restatement (oracle/tvm_oracle.c)."""

    def __init__(self, path: str | None = None):
        lib = self.lib = _load(path or os.path.join(HERE, "liboracle.so"))
        u64, sz, i32 = C.c_uint64, C.c_size_t, C.c_int
        sig = {
            "orc_add": (u64, [u64, u64]), "orc_sub": (u64, [u64, u64]),
            "orc_mul": (u64, [u64, u64]), "orc_pow": (u64, [u64, u64]),
            "orc_inv": (u64, [u64]), "orc_root_of_unity": (u64, [C.c_uint32]),
            "orc_montyred": (u64, [u64, u64]),
            "orc_ntt_forward": (i32, [_u64p, sz]), "orc_ntt_inverse": (i32, [_u64p, sz]),
            "orc_evaluate_on_coset": (i32, [_u64p, sz, sz, u64, _u64p]),
            "orc_lde_column": (i32, [_u64p, sz, u64, C.c_void_p, sz, u64, sz, _u64p]),
            "orc_lde_limb_truncating": (i32, [_u64p, sz, u64, C.c_void_p, sz, u64, sz, _u64p]),
            "orc_lde_table_colmajor": (i32, [_u64p, sz, sz, u64, C.c_void_p, sz, u64, sz, _u64p]),
            "orc_lde_sample_colmajor": (i32, [_u64p, sz, sz, u64, C.c_void_p, sz, u64, sz,
                                              _u64p, sz, i32, _u64p]),
            "orc_tip5_permutation": (None, [_u64p]),
            "orc_tip5_round": (None, [_u64p, i32]),
            "orc_hash_10": (None, [_u64p, _u64p]),
            "orc_hash_varlen": (None, [_u64p, sz, _u64p]),
            "orc_hash_pair": (None, [_u64p, _u64p, _u64p]),
            "orc_hash_rows_colmajor": (None, [_u64p, sz, sz, _u64p]),
            "orc_tip5_lookup_table": (C.POINTER(C.c_uint8), []),
            "orc_tip5_round_constants": (C.POINTER(C.c_uint64), []),
            "orc_tip5_mds_first_column": (C.POINTER(C.c_int64), []),
            "orc_merkle_heap": (i32, [_u64p, sz, _u64p]),
            "orc_merkle_level_order": (i32, [_u64p, sz, _u64p]),
            "orc_auth_structure_indices": (C.c_long, [sz, _szp, sz, _szp, sz]),
            "orc_merkle_verify_path": (i32, [_u64p, sz, _u64p, _u64p, sz]),
            "orc_commit": (i32, [_u64p, sz, sz, u64, C.c_void_p, sz, u64, sz,
                                 C.c_void_p, C.c_void_p, _u64p]),
            "orc_splitmix_fill": (None, [u64, _u64p, sz]),
        }
        for name, (res, args) in sig.items():
            f = getattr(lib, name)
            f.restype, f.argtypes = res, args

    # field -----------------------------------------------------------------
    def add(self, a, b): return self.lib.orc_add(a, b)
    def sub(self, a, b): return self.lib.orc_sub(a, b)
    def mul(self, a, b): return self.lib.orc_mul(a, b)
    def pow(self, a, e): return self.lib.orc_pow(a, e)
    def inv(self, a): return self.lib.orc_inv(a)
    def root_of_unity(self, k): return self.lib.orc_root_of_unity(k)
    def montyred(self, hi, lo): return self.lib.orc_montyred(hi, lo)

    # NTT / LDE ---------------------------------------------------------------
    def ntt(self, a):
        a = np.ascontiguousarray(a, dtype=np.uint64).copy()
        if self.lib.orc_ntt_forward(a, a.size):
            raise ValueError("NTT size must be a power of 2")
        return a

    def intt(self, a):
        a = np.ascontiguousarray(a, dtype=np.uint64).copy()
        if self.lib.orc_ntt_inverse(a, a.size):
            raise ValueError("NTT size must be a power of 2")
        return a

    def evaluate_on_coset(self, coeffs, m, offset):
        c = np.ascontiguousarray(coeffs, dtype=np.uint64)
        out = np.zeros(m, np.uint64)
        if self.lib.orc_evaluate_on_coset(c, c.size, m, offset, out):
            raise ValueError("bad size")
        return out

    @staticmethod
    def _rand_ptr(rand):
        if rand is None or len(rand) == 0:
            return None, 0, None
        r = np.ascontiguousarray(rand, dtype=np.uint64)
        return r.ctypes.data, r.shape[-1], r

    def lde_column(self, trace, m, trace_off=1, target_off=7, rand=None):
        t = np.ascontiguousarray(trace, dtype=np.uint64)
        rp, k, _keep = self._rand_ptr(rand)
        out = np.zeros(m, np.uint64)
        if self.lib.orc_lde_column(t, t.size, trace_off, rp, k, target_off, m, out):
            raise ValueError("bad lde arguments")
        return out

    def lde_limb_truncating(self, trace, m, trace_off=1, target_off=7, rand=None):
        t = np.ascontiguousarray(trace, dtype=np.uint64)
        rp, k, _keep = self._rand_ptr(rand)
        out = np.zeros(m, np.uint64)
        if self.lib.orc_lde_limb_truncating(t, t.size, trace_off, rp, k, target_off, m, out):
            raise ValueError("bad lde arguments")
        return out

    def lde_table(self, trace_cm, n, cols, m, trace_off=1, target_off=7, rand=None):
        """trace_cm: column-major flat (cols*n); rand: (cols, k) or None."""
        t = np.ascontiguousarray(trace_cm, dtype=np.uint64).reshape(-1)
        rp, k, _keep = self._rand_ptr(None if rand is None else np.asarray(rand).reshape(cols, -1))
        out = np.zeros(m * cols, np.uint64)
        if self.lib.orc_lde_table_colmajor(t, n, cols, trace_off, rp, k, target_off, m, out):
            raise ValueError("bad lde arguments")
        return out

    def lde_sample(self, trace_cm, n, cols, m, rows, trace_off=1, target_off=7, rand=None,
                   truncate=False):
        """LDE values of every column at the given rows: (cols, len(rows))."""
        t = np.ascontiguousarray(trace_cm, dtype=np.uint64).reshape(-1)
        rp, k, _keep = self._rand_ptr(None if rand is None else np.asarray(rand).reshape(cols, -1))
        r = np.ascontiguousarray(rows, dtype=np.uint64)
        out = np.zeros(cols * r.size, np.uint64)
        if self.lib.orc_lde_sample_colmajor(t, n, cols, trace_off, rp, k, target_off, m,
                                            r if r.size else np.zeros(1, np.uint64), r.size,
                                            int(truncate), out):
            raise ValueError("bad lde arguments")
        return out.reshape(cols, r.size)

    # Tip5 ---------------------------------------------------------------------
    def permutation(self, st):
        s = np.ascontiguousarray(st, dtype=np.uint64).copy()
        self.lib.orc_tip5_permutation(s)
        return s

    def tip5_round(self, st, r):
        s = np.ascontiguousarray(st, dtype=np.uint64).copy()
        self.lib.orc_tip5_round(s, r)
        return s

    def hash_10(self, x):
        out = np.zeros(5, np.uint64)
        self.lib.orc_hash_10(np.ascontiguousarray(x, dtype=np.uint64), out)
        return out

    def hash_varlen(self, x):
        x = np.ascontiguousarray(x, dtype=np.uint64)
        out = np.zeros(5, np.uint64)
        self.lib.orc_hash_varlen(x if x.size else np.zeros(1, np.uint64), x.size, out)
        return out

    def hash_pair(self, l, r):
        out = np.zeros(5, np.uint64)
        self.lib.orc_hash_pair(np.ascontiguousarray(l, np.uint64),
                               np.ascontiguousarray(r, np.uint64), out)
        return out

    def hash_rows(self, table_cm, rows, cols):
        t = np.ascontiguousarray(table_cm, dtype=np.uint64).reshape(-1)
        out = np.zeros(rows * 5, np.uint64)
        self.lib.orc_hash_rows_colmajor(t if t.size else np.zeros(1, np.uint64), rows, cols, out)
        return out.reshape(rows, 5)

    def lookup_table(self):
        p = self.lib.orc_tip5_lookup_table()
        return np.array([p[i] for i in range(256)], dtype=np.uint8)

    def round_constants(self):
        p = self.lib.orc_tip5_round_constants()
        return np.array([p[i] for i in range(80)], dtype=np.uint64)

    def mds_first_column(self):
        p = self.lib.orc_tip5_mds_first_column()
        return [int(p[i]) for i in range(16)]

    # Merkle -------------------------------------------------------------------
    def merkle_heap(self, leaves):
        lv = np.ascontiguousarray(leaves, dtype=np.uint64).reshape(-1)
        n = lv.size // 5
        nodes = np.zeros(2 * n * 5, np.uint64)
        if self.lib.orc_merkle_heap(lv, n, nodes):
            raise ValueError("Number of leaves must be a power of 2")
        return nodes.reshape(2 * n, 5)

    def merkle_level_order(self, leaves):
        lv = np.ascontiguousarray(leaves, dtype=np.uint64).reshape(-1)
        n = lv.size // 5
        tree = np.zeros((2 * n - 1) * 5, np.uint64)
        if self.lib.orc_merkle_level_order(lv, n, tree):
            raise ValueError("Number of leaves must be a power of 2")
        return tree.reshape(2 * n - 1, 5)

    def auth_structure_indices(self, n, leaf_indices):
        li = np.ascontiguousarray(leaf_indices, dtype=np.uintp)
        cap = max(1, li.size * 64)
        out = np.zeros(cap, np.uintp)
        c = self.lib.orc_auth_structure_indices(n, li if li.size else np.zeros(1, np.uintp),
                                                li.size, out, cap)
        if c < 0:
            raise IndexError("Leaf index out of range")
        return out[:c].copy()

    def verify_path(self, root, idx, leaf, path):
        p = np.ascontiguousarray(path, dtype=np.uint64).reshape(-1)
        return bool(self.lib.orc_merkle_verify_path(
            np.ascontiguousarray(root, np.uint64), idx, np.ascontiguousarray(leaf, np.uint64),
            p if p.size else np.zeros(5, np.uint64), p.size // 5))

    # commit / inputs ----------------------------------------------------------
    def commit(self, trace_cm, n, cols, m, trace_off=1, target_off=7, rand=None,
               want_codewords=False, want_digests=False):
        t = np.ascontiguousarray(trace_cm, dtype=np.uint64).reshape(-1)
        rp, k, _keep = self._rand_ptr(None if rand is None else np.asarray(rand).reshape(cols, -1))
        cw = np.zeros(m * cols, np.uint64) if want_codewords else None
        dg = np.zeros(m * 5, np.uint64) if want_digests else None
        root = np.zeros(5, np.uint64)
        rc = self.lib.orc_commit(t, n, cols, trace_off, rp, k, target_off, m,
                                 None if cw is None else cw.ctypes.data,
                                 None if dg is None else dg.ctypes.data, root)
        if rc:
            raise ValueError("bad commit arguments")
        return root, cw, (None if dg is None else dg.reshape(m, 5))

    def splitmix(self, seed, count):
        out = np.zeros(count, np.uint64)
        self.lib.orc_splitmix_fill(seed, out, count)
        return out

    def table_rowmajor_draw(self, seed, rows, cols):
        """Uniform table drawn row-major from SplitMix64(seed); returned column-major."""
        return np.ascontiguousarray(self.splitmix(seed, rows * cols).reshape(rows, cols).T).reshape(-1)


class Reference:
    """The reference's own CPU classes (oracle/_ref/libtvm_ref.so), when built."""

    PATH = os.path.join(HERE, "_ref", "libtvm_ref.so")

    @classmethod
    def available(cls) -> bool:
        return os.path.exists(cls.PATH)

    def __init__(self):
        lib = self.lib = _load(self.PATH)
        u64, sz, i32 = C.c_uint64, C.c_size_t, C.c_int
        sig = {
            "ref_set_threads": (None, [i32]),
            "ref_ntt_forward": (i32, [_u64p, sz]), "ref_ntt_inverse": (i32, [_u64p, sz]),
            "ref_lde_column": (i32, [_u64p, sz, u64, C.c_void_p, sz, u64, sz, _u64p]),
            "ref_lde_xfe_column": (i32, [_u64p, sz, u64, C.c_void_p, sz, i32, u64, sz, _u64p]),
            "ref_tip5_permutation": (None, [_u64p]),
            "ref_tip5_trace": (None, [_u64p, _u64p]),
            "ref_hash_varlen": (None, [_u64p, sz, _u64p]),
            "ref_hash_pair": (None, [_u64p, _u64p, _u64p]),
            "ref_hash_rows_colmajor": (None, [_u64p, sz, sz, _u64p]),
            "ref_merkle": (i32, [_u64p, sz, _u64p, C.c_void_p]),
            "ref_auth_structure": (C.c_long, [_u64p, sz, _szp, sz, _u64p, sz]),
            "ref_commit": (i32, [_u64p, sz, sz, u64, C.c_void_p, sz, u64, sz,
                                 C.c_void_p, C.c_void_p, _u64p]),
            "ref_commit_parallel": (i32, [_u64p, sz, sz, u64, sz, u64, _u64p]),
        }
        for name, (res, args) in sig.items():
            f = getattr(lib, name)
            f.restype, f.argtypes = res, args

    def set_threads(self, t): self.lib.ref_set_threads(t)

    def ntt(self, a):
        a = np.ascontiguousarray(a, dtype=np.uint64).copy()
        assert self.lib.ref_ntt_forward(a, a.size) == 0
        return a

    def intt(self, a):
        a = np.ascontiguousarray(a, dtype=np.uint64).copy()
        assert self.lib.ref_ntt_inverse(a, a.size) == 0
        return a

    def lde_column(self, trace, m, trace_off=1, target_off=7, rand=None):
        t = np.ascontiguousarray(trace, dtype=np.uint64)
        r = None if rand is None or len(rand) == 0 else np.ascontiguousarray(rand, np.uint64)
        out = np.zeros(m, np.uint64)
        rc = self.lib.ref_lde_column(t, t.size, trace_off, None if r is None else r.ctypes.data,
                                     0 if r is None else r.size, target_off, m, out)
        assert rc == 0
        return out

    def lde_xfe_column(self, trace3, m, trace_off=1, target_off=7, rand=None, xfe_rand=False):
        """trace3: (n, 3) limbs; rand: (k,) BFE or (k, 3) XFE.  Returns (m, 3)."""
        t = np.ascontiguousarray(trace3, dtype=np.uint64).reshape(-1)
        n = t.size // 3
        r = None if rand is None or len(rand) == 0 else np.ascontiguousarray(rand, np.uint64)
        k = 0 if r is None else (r.size // 3 if xfe_rand else r.size)
        out = np.zeros(m * 3, np.uint64)
        rc = self.lib.ref_lde_xfe_column(t, n, trace_off, None if r is None else r.ctypes.data,
                                         k, int(xfe_rand), target_off, m, out)
        assert rc == 0
        return out.reshape(m, 3)

    def tip5_trace(self, st):
        s = np.ascontiguousarray(st, dtype=np.uint64)
        out = np.zeros(6 * 16, np.uint64)
        self.lib.ref_tip5_trace(s, out)
        return out.reshape(6, 16)

    def permutation(self, st):
        s = np.ascontiguousarray(st, dtype=np.uint64).copy()
        self.lib.ref_tip5_permutation(s)
        return s

    def hash_varlen(self, x):
        x = np.ascontiguousarray(x, dtype=np.uint64)
        out = np.zeros(5, np.uint64)
        self.lib.ref_hash_varlen(x if x.size else np.zeros(1, np.uint64), x.size, out)
        return out

    def hash_pair(self, l, r):
        out = np.zeros(5, np.uint64)
        self.lib.ref_hash_pair(np.ascontiguousarray(l, np.uint64),
                               np.ascontiguousarray(r, np.uint64), out)
        return out

    def hash_rows(self, table_cm, rows, cols):
        t = np.ascontiguousarray(table_cm, dtype=np.uint64).reshape(-1)
        out = np.zeros(rows * 5, np.uint64)
        self.lib.ref_hash_rows_colmajor(t if t.size else np.zeros(1, np.uint64), rows, cols, out)
        return out.reshape(rows, 5)

    def merkle(self, leaves, want_nodes=False):
        lv = np.ascontiguousarray(leaves, dtype=np.uint64).reshape(-1)
        n = lv.size // 5
        root = np.zeros(5, np.uint64)
        nodes = np.zeros(2 * n * 5, np.uint64) if want_nodes else None
        assert self.lib.ref_merkle(lv, n, root, None if nodes is None else nodes.ctypes.data) == 0
        return root, (None if nodes is None else nodes.reshape(2 * n, 5))

    def auth_structure(self, leaves, leaf_indices):
        lv = np.ascontiguousarray(leaves, dtype=np.uint64).reshape(-1)
        li = np.ascontiguousarray(leaf_indices, dtype=np.uintp)
        cap = max(1, li.size * 64)
        out = np.zeros(cap * 5, np.uint64)
        c = self.lib.ref_auth_structure(lv, lv.size // 5, li, li.size, out, cap)
        assert c >= 0
        return out[: 5 * c].reshape(c, 5)

    def commit(self, trace_cm, n, cols, m, trace_off=1, target_off=7, rand=None,
               want_codewords=False, want_digests=False):
        t = np.ascontiguousarray(trace_cm, dtype=np.uint64).reshape(-1)
        r = None if rand is None else np.ascontiguousarray(np.asarray(rand).reshape(cols, -1), np.uint64)
        k = 0 if r is None else r.shape[1]
        cw = np.zeros(m * cols, np.uint64) if want_codewords else None
        dg = np.zeros(m * 5, np.uint64) if want_digests else None
        root = np.zeros(5, np.uint64)
        rc = self.lib.ref_commit(t, n, cols, trace_off, None if r is None else r.ctypes.data, k,
                                 target_off, m, None if cw is None else cw.ctypes.data,
                                 None if dg is None else dg.ctypes.data, root)
        assert rc == 0
        return root, cw, (None if dg is None else dg.reshape(m, 5))

    def commit_parallel(self, trace_cm, n, cols, m, trace_off=1, target_off=7):
        """Root of the reference commit with OpenMP over columns and rows (no randomizers)."""
        t = np.ascontiguousarray(trace_cm, dtype=np.uint64).reshape(-1)
        root = np.zeros(5, np.uint64)
        assert self.lib.ref_commit_parallel(t, n, cols, trace_off, m, target_off, root) == 0
        return root
