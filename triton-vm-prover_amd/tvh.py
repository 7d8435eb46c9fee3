"""ctypes binding of libtvh (include/tvh.h) for Python callers and tests.

Device buffers are hip_rt.DeviceBuffer objects (u64 elements allocated by the
same HIP runtime libtvh links); streams are hip_rt.Stream objects or None (the
null stream).  Every call goes through the C ABI and raises TVHError on a
non-zero status.  There is no CPU fallback: without the built library or a GPU
the calls fail.
"""
from __future__ import annotations

import ctypes as C
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)
import hip_rt  # noqa: E402
ROOT = os.path.dirname(HERE)
LIB_PATH = os.path.join(HERE, "libtvh.so")
HEADER = os.path.join(ROOT, "include", "tvh.h")
P = 0xFFFFFFFF00000001

TREE_LEVEL_ORDER = 0
TREE_HEAP = 1


class TVHError(RuntimeError):
    def __init__(self, status: int, name: str, message: str):
        super().__init__(f"{name} (status {status}): {message}")
        self.status = status


def declared_symbols(header: str = HEADER) -> list[str]:
    """Names of every entry point declared in include/tvh.h."""
    with open(header) as f:
        return re.findall(r"^\s*(?:int|const char\*)\s+(tvh_\w+)\(", f.read(), re.M)


_vp, _sz, _u64, _i32 = C.c_void_p, C.c_size_t, C.c_uint64, C.c_int

_SIGS = {
    "tvh_init": [_i32],
    "tvh_status_string": [_i32],
    "tvh_last_error": [],
    "tvh_ntt_forward_batched_dev": [_vp, _sz, _sz, _vp],
    "tvh_ntt_inverse_batched_dev": [_vp, _sz, _sz, _vp],
    "tvh_ntt_forward_dev": [_vp, _sz, _vp],
    "tvh_ntt_inverse_dev": [_vp, _sz, _vp],
    "tvh_lde_batch_dev": [_vp, _sz, _sz, _vp, _sz, _u64, _u64, _sz, _vp, _vp],
    "tvh_lde_column_dev": [_vp, _sz, _vp, _sz, _u64, _u64, _sz, _vp, _vp],
    "tvh_lde_xfe_batch_dev": [_vp, _sz, _sz, _vp, _sz, _u64, _u64, _sz, _vp, _i32, _vp],
    "tvh_lde_xfe_limbs_dev": [_vp, _sz, _sz, _vp, _sz, _u64, _u64, _sz, _vp, _i32, _vp],
    "tvh_hash_bfield_rows_dev": [_vp, _sz, _sz, _vp, _vp],
    "tvh_hash_xfield_rows_dev": [_vp, _sz, _sz, _vp, _vp],
    "tvh_tip5_permutation_batch_dev": [_vp, _sz, _vp],
    "tvh_hash_pairs_dev": [_vp, _vp, _vp, _sz, _vp],
    "tvh_merkle_tree_dev": [_vp, _vp, _sz, _i32, _vp],
    "tvh_merkle_root_dev": [_vp, _vp, _sz, _vp],
    "tvh_merkle_gather_dev": [_vp, _sz, _i32, _vp, _sz, _vp, _vp],
    "tvh_auth_path_indices": [_sz, _sz, _vp, _sz, _vp],
    "tvh_auth_structure_indices": [_sz, _vp, _sz, _vp, _sz, _vp],
    "tvh_commit_dev": [_vp, _sz, _sz, _vp, _sz, _u64, _u64, _sz, _vp, _vp, _i32, _vp],
    "tvh_ntt_forward_host": [_vp, _sz],
    "tvh_ntt_inverse_host": [_vp, _sz],
    "tvh_lde_batch_host": [_vp, _sz, _sz, _vp, _sz, _u64, _u64, _sz, _vp],
    "tvh_lde_xfe_batch_host": [_vp, _sz, _sz, _vp, _sz, _u64, _u64, _sz, _vp, _i32],
    "tvh_tip5_permutation_batch_host": [_vp, _sz],
    "tvh_hash_pairs_host": [_vp, _vp, _vp, _sz],
    "tvh_hash_bfield_rows_host": [_vp, _sz, _sz, _vp],
    "tvh_merkle_tree_host": [_vp, _sz, _i32, _vp, _vp],
    "tvh_commit_host": [_vp, _sz, _sz, _vp, _sz, _u64, _u64, _sz, _vp, _vp, _vp],
}


def _ptr(x):
    """Address of a DeviceBuffer / numpy array / None."""
    if x is None:
        return None
    if hasattr(x, "data_ptr"):
        return x.data_ptr()
    return x.ctypes.data


class TVH:
    """The loaded library.  Methods mirror the C entry points one to one."""

    def __init__(self, path: str = LIB_PATH):
        hip_rt.hip()  # the runtime libtvh links, loaded first and globally
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built (python -c 'import __graft_entry__ as g; g.build()')")
        self.path = path
        self.lib = C.CDLL(path)
        for name, args in _SIGS.items():
            f = getattr(self.lib, name)
            f.argtypes = args
            f.restype = C.c_char_p if name in ("tvh_status_string", "tvh_last_error") else C.c_int

    # -- plumbing -------------------------------------------------------------
    def status_string(self, rc: int) -> str:
        return self.lib.tvh_status_string(rc).decode()

    def last_error(self) -> str:
        return self.lib.tvh_last_error().decode()

    def _call(self, name, *args):
        rc = getattr(self.lib, name)(*args)
        if rc:
            raise TVHError(rc, self.status_string(rc), self.last_error())
        return rc

    def raw(self, name, *args) -> int:
        """Call without raising; returns the status code."""
        return getattr(self.lib, name)(*args)

    def tvh_init(self, device: int = 0) -> int:
        return self.lib.tvh_init(device)

    @staticmethod
    def stream_ptr(stream=None):
        return None if stream is None else stream.ptr

    # -- NTT --------------------------------------------------------------------
    def ntt_forward_batched(self, d, n, cols, stream=None):
        self._call("tvh_ntt_forward_batched_dev", _ptr(d), n, cols, self.stream_ptr(stream))

    def ntt_inverse_batched(self, d, n, cols, stream=None):
        self._call("tvh_ntt_inverse_batched_dev", _ptr(d), n, cols, self.stream_ptr(stream))

    # -- LDE --------------------------------------------------------------------
    def lde_batch(self, trace, cols, n, rand, k, toff, goff, m, out, stream=None):
        self._call("tvh_lde_batch_dev", _ptr(trace), cols, n, _ptr(rand), k, toff, goff, m,
                   _ptr(out), self.stream_ptr(stream))

    def lde_xfe_batch(self, xt, xcols, n, rand, k, toff, goff, m, xout, xfe_rand, stream=None):
        self._call("tvh_lde_xfe_batch_dev", _ptr(xt), xcols, n, _ptr(rand), k, toff, goff, m,
                   _ptr(xout), int(xfe_rand), self.stream_ptr(stream))

    def lde_xfe_limbs(self, limbs, xcols, n, rand, k, toff, goff, m, out, xfe_rand, stream=None):
        self._call("tvh_lde_xfe_limbs_dev", _ptr(limbs), xcols, n, _ptr(rand), k, toff, goff, m,
                   _ptr(out), int(xfe_rand), self.stream_ptr(stream))

    # -- Tip5 -------------------------------------------------------------------
    def hash_bfield_rows(self, table, rows, cols, digests, stream=None):
        self._call("tvh_hash_bfield_rows_dev", _ptr(table), rows, cols, _ptr(digests),
                   self.stream_ptr(stream))

    def hash_xfield_rows(self, table, rows, xcols, digests, stream=None):
        self._call("tvh_hash_xfield_rows_dev", _ptr(table), rows, xcols, _ptr(digests),
                   self.stream_ptr(stream))

    def permutation_batch(self, states, num, stream=None):
        self._call("tvh_tip5_permutation_batch_dev", _ptr(states), num, self.stream_ptr(stream))

    def hash_pairs(self, left, right, out, count, stream=None):
        self._call("tvh_hash_pairs_dev", _ptr(left), _ptr(right), _ptr(out), count,
                   self.stream_ptr(stream))

    # -- Merkle -----------------------------------------------------------------
    def merkle_tree(self, leaves, tree, n, layout=TREE_LEVEL_ORDER, stream=None):
        self._call("tvh_merkle_tree_dev", _ptr(leaves), _ptr(tree), n, layout,
                   self.stream_ptr(stream))

    def merkle_root(self, leaves, root, n, stream=None):
        self._call("tvh_merkle_root_dev", _ptr(leaves), _ptr(root), n, self.stream_ptr(stream))

    def merkle_gather(self, tree, n, layout, idx, count, out, stream=None):
        self._call("tvh_merkle_gather_dev", _ptr(tree), n, layout, _ptr(idx), count, _ptr(out),
                   self.stream_ptr(stream))

    def auth_path_indices(self, n, leaf):
        import numpy as np

        cnt = C.c_size_t(0)
        self._call("tvh_auth_path_indices", n, leaf, None, 0, C.byref(cnt))
        out = np.zeros(max(1, cnt.value), np.uint64)
        self._call("tvh_auth_path_indices", n, leaf, out.ctypes.data, out.size, C.byref(cnt))
        return out[: cnt.value]

    def auth_structure_indices(self, n, leaves):
        import numpy as np

        li = np.ascontiguousarray(leaves, dtype=np.uintp)
        cnt = C.c_size_t(0)
        lp = li.ctypes.data if li.size else None
        self._call("tvh_auth_structure_indices", n, lp, li.size, None, 0, C.byref(cnt))
        out = np.zeros(max(1, cnt.value), np.uint64)
        self._call("tvh_auth_structure_indices", n, lp, li.size, out.ctypes.data, out.size,
                   C.byref(cnt))
        return out[: cnt.value]

    # -- commit -----------------------------------------------------------------
    def commit(self, trace, cols, n, rand, k, toff, goff, m, codewords, tree,
               layout=TREE_LEVEL_ORDER, stream=None):
        self._call("tvh_commit_dev", _ptr(trace), cols, n, _ptr(rand), k, toff, goff, m,
                   _ptr(codewords), _ptr(tree), layout, self.stream_ptr(stream))

    # -- host forms (numpy uint64 arrays) ----------------------------------------
    def ntt_forward_host(self, a):
        self._call("tvh_ntt_forward_host", _ptr(a), a.size)

    def ntt_inverse_host(self, a):
        self._call("tvh_ntt_inverse_host", _ptr(a), a.size)

    def commit_host(self, trace, cols, n, m, toff=1, goff=7, rand=None, k=0, codewords=None,
                    digests=None, root=None):
        self._call("tvh_commit_host", _ptr(trace), cols, n, _ptr(rand), k, toff, goff, m,
                   _ptr(codewords), _ptr(digests), _ptr(root))


# ---- buffer helpers ------------------------------------------------------------
def to_device(a):
    """numpy uint64 -> DeviceBuffer."""
    return hip_rt.to_device(a)


def to_host(buf, count=None):
    """DeviceBuffer -> numpy uint64."""
    return buf.download(count)


def empty(n):
    return hip_rt.DeviceBuffer(max(1, int(n)))


def load(path: str = LIB_PATH) -> TVH:
    return TVH(path)
