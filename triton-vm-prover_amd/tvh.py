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

import numpy as np

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
        return re.findall(r"^\s*(?:int|size_t|const char\*)\s+(tvh_\w+)\(", f.read(), re.M)


_vp, _sz, _u64, _i32 = C.c_void_p, C.c_size_t, C.c_uint64, C.c_int

class TableDesc(C.Structure):
    """tvh_table_desc (include/tvh.h)."""
    _fields_ = [("trace", C.c_void_p), ("num_cols", C.c_size_t), ("trace_len", C.c_size_t),
                ("rand", C.c_void_p), ("rand_len", C.c_size_t), ("trace_offset", C.c_uint64),
                ("target_offset", C.c_uint64), ("target_len", C.c_size_t),
                ("codewords", C.c_void_p), ("tree", C.c_void_p), ("layout", C.c_int)]


class ServerConfig(C.Structure):
    """tvh_server_config (include/tvh.h)."""
    _fields_ = [("hbm_budget_bytes", C.c_size_t), ("max_inflight", C.c_int),
                ("max_queue", C.c_size_t)]


class JobDesc(C.Structure):
    """tvh_job_desc (include/tvh.h)."""
    _fields_ = [("host_trace", C.c_void_p), ("host_rand", C.c_void_p), ("seed", C.c_uint64),
                ("num_cols", C.c_size_t), ("trace_len", C.c_size_t), ("rand_len", C.c_size_t),
                ("trace_offset", C.c_uint64), ("target_offset", C.c_uint64),
                ("target_len", C.c_size_t)]


class JobTimes(C.Structure):
    """tvh_job_times (include/tvh.h)."""
    _fields_ = [("submit_s", C.c_double), ("admit_s", C.c_double), ("done_s", C.c_double)]


JOB_QUEUED, JOB_RUNNING, JOB_DONE, JOB_FAILED = 0, 1, 2, 3
LDE_AUTO, LDE_WIDE, LDE_SPLIT = 0, 1, 2
LDE_ROW_HALF = 16
LDE_INV_HALF = 32
LDE_ROW_QUARTER = 64
# form | column tile code << 2 | LDE_ROW_HALF, LDE_INV_HALF or LDE_ROW_QUARTER (split form only)
LDE_VARIANTS = (1, 2, 5, 6, 9, 10, 18, 22, 26, 34, 38, 42, 66, 70, 74)


def lde_variant_name(v):
    name = {LDE_WIDE: "wide", LDE_SPLIT: "split"}[v & 3] + f"/t{14 - ((v >> 2) & 3)}"
    return (name + ("/r13" if v & LDE_ROW_HALF else "") + ("/i13" if v & LDE_INV_HALF else "")
            + ("/r12" if v & LDE_ROW_QUARTER else ""))

_SIGS = {
    "tvh_init": [_i32],
    "tvh_status_string": [_i32],
    "tvh_last_error": [],
    "tvh_ntt_forward_batched_dev": [_vp, _sz, _sz, _vp],
    "tvh_ntt_inverse_batched_dev": [_vp, _sz, _sz, _vp],
    "tvh_ntt_batched_noalloc_dev": [_vp, _sz, _sz, _i32, _vp],
    "tvh_ntt_forward_dev": [_vp, _sz, _vp],
    "tvh_ntt_inverse_dev": [_vp, _sz, _vp],
    "tvh_lde_batch_dev": [_vp, _sz, _sz, _vp, _sz, _u64, _u64, _sz, _vp, _vp],
    "tvh_lde_column_dev": [_vp, _sz, _vp, _sz, _u64, _u64, _sz, _vp, _vp],
    "tvh_lde_xfe_batch_dev": [_vp, _sz, _sz, _vp, _sz, _u64, _u64, _sz, _vp, _i32, _vp],
    "tvh_lde_xfe_limbs_dev": [_vp, _sz, _sz, _vp, _sz, _u64, _u64, _sz, _vp, _i32, _vp],
    "tvh_trace_coefficients_dev": [_vp, _sz, _sz, _u64, _vp, _vp],
    "tvh_evaluate_coset_dev": [_vp, _sz, _sz, _vp, _sz, _u64, _u64, _vp, _vp, _vp],
    "tvh_scatter_digests_strided_dev": [_vp, _vp, _sz, _sz, _sz, _vp],
    "tvh_rowmajor_to_colmajor_bfe_dev": [_vp, _vp, _sz, _sz, _vp],
    "tvh_rowmajor_to_colmajor_xfe_dev": [_vp, _vp, _sz, _sz, _vp],
    "tvh_colmajor_to_rowmajor_bfe_dev": [_vp, _vp, _sz, _sz, _vp],
    "tvh_colmajor_to_rowmajor_xfe_dev": [_vp, _vp, _sz, _sz, _vp],
    "tvh_hash_bfield_rows_dev": [_vp, _sz, _sz, _vp, _vp],
    "tvh_hash_xfield_rows_dev": [_vp, _sz, _sz, _vp, _vp],
    "tvh_tip5_permutation_batch_dev": [_vp, _sz, _vp],
    "tvh_tip5_rounds_dev": [_vp, _sz, C.c_uint32, C.c_uint32, _vp, _vp],
    "tvh_tip5_rounds_host": [_vp, _sz, C.c_uint32, C.c_uint32, _vp],
    "tvh_lde_prepare": [_sz, _sz, _u64, _u64],
    "tvh_lde_release_constants": [],
    "tvh_hash_pairs_dev": [_vp, _vp, _vp, _sz, _vp],
    "tvh_merkle_tree_dev": [_vp, _vp, _sz, _i32, _vp],
    "tvh_merkle_root_dev": [_vp, _vp, _sz, _vp],
    "tvh_merkle_gather_dev": [_vp, _sz, _i32, _vp, _sz, _vp, _vp],
    "tvh_auth_path_indices": [_sz, _sz, _vp, _sz, _vp],
    "tvh_auth_structure_indices": [_sz, _vp, _sz, _vp, _sz, _vp],
    "tvh_commit_dev": [_vp, _sz, _sz, _vp, _sz, _u64, _u64, _sz, _vp, _vp, _i32, _vp],
    "tvh_xfe_leaves_dev": [_vp, _sz, _vp, _vp],
    "tvh_fri_commit_dev": [_vp, _sz, _vp, _i32, _vp],
    "tvh_sponge_init_dev": [_vp, _vp],
    "tvh_sponge_absorb_dev": [_vp, _vp, _sz, _vp],
    "tvh_sponge_squeeze_dev": [_vp, _vp, _vp],
    "tvh_batch_plan_create": [_vp, _sz, _vp],
    "tvh_batch_plan_launch": [_vp, _vp],
    "tvh_batch_plan_launch_graph": [_vp, _vp],
    "tvh_batch_plan_destroy": [_vp],
    "tvh_commit_batch_dev": [_vp, _sz, _vp],
    "tvh_fill_uniform_dev": [_vp, _sz, _u64, _vp],
    "tvh_upload_dev": [_vp, _vp, _sz, _vp],
    "tvh_fill_table_dev": [_vp, _sz, _sz, _u64, _vp],
    "tvh_commit_frugal_workspace_elems": [_sz, _sz, _sz, _sz],
    "tvh_commit_frugal_dev": [_vp, _sz, _sz, _vp, _sz, _u64, _u64, _sz, _sz, _vp, _vp, _i32, _vp],
    "tvh_lde_set_variant": [C.c_uint32, C.c_uint32, _i32],
    "tvh_lde_get_variant": [C.c_uint32, C.c_uint32, _vp],
    "tvh_lde_get_variant_for_width": [C.c_uint32, C.c_uint32, _sz, _vp],
    "tvh_lde_autotune": [_sz, _sz, _sz, _vp, _vp],
    "tvh_lde_variant_count": [],
    "tvh_server_create": [_vp, _vp],
    "tvh_server_submit": [_vp, _vp, _vp],
    "tvh_server_poll": [_vp, _u64, _vp],
    "tvh_server_wait": [_vp, _u64, _vp, _vp],
    "tvh_server_open": [_vp, _u64, _vp, _sz, _vp, _vp, _sz, _vp],
    "tvh_server_paths": [_vp, _u64, _vp, _sz, _vp],
    "tvh_server_buffers": [_vp, _u64, _vp, _vp],
    "tvh_server_release": [_vp, _u64],
    "tvh_server_stats": [_vp, _vp, _vp, _vp, _vp],
    "tvh_server_destroy": [_vp],
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
    """Address of a DeviceBuffer / numpy array / raw address (int) / None."""
    if x is None:
        return None
    if isinstance(x, int):
        return x
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
            f.restype = (C.c_char_p if name in ("tvh_status_string", "tvh_last_error") else
                         C.c_size_t if name.endswith("_elems") else C.c_int)

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

    def ntt_batched_noalloc(self, d, n, cols, inverse, stream=None):
        self._call("tvh_ntt_batched_noalloc_dev", _ptr(d), n, cols, int(inverse),
                   self.stream_ptr(stream))

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

    def lde_prepare(self, n, m, toff=1, goff=7):
        self._call("tvh_lde_prepare", n, m, toff, goff)

    # -- two-phase coset evaluation, digest scatter, transposes ---------------------
    def trace_coefficients(self, trace, cols, n, toff, coeffs, stream=None):
        self._call("tvh_trace_coefficients_dev", _ptr(trace), cols, n, toff, _ptr(coeffs),
                   self.stream_ptr(stream))

    def evaluate_coset(self, coeffs, cols, n, rand, k, toff, coset_offset, out, stream=None):
        self._call("tvh_evaluate_coset_dev", _ptr(coeffs), cols, n, _ptr(rand), k, toff,
                   coset_offset, _ptr(out), None, self.stream_ptr(stream))

    def scatter_digests(self, dig, leaves, length, coset, num_cosets, stream=None):
        self._call("tvh_scatter_digests_strided_dev", _ptr(dig), _ptr(leaves), length, coset,
                   num_cosets, self.stream_ptr(stream))

    def transpose(self, src, dst, n, width, to_colmajor=True, xfe=False, stream=None):
        name = ("tvh_rowmajor_to_colmajor_" if to_colmajor else "tvh_colmajor_to_rowmajor_") + \
               ("xfe_dev" if xfe else "bfe_dev")
        self._call(name, _ptr(src), _ptr(dst), n, width, self.stream_ptr(stream))

    # -- Tip5 -------------------------------------------------------------------
    def hash_bfield_rows(self, table, rows, cols, digests, stream=None):
        self._call("tvh_hash_bfield_rows_dev", _ptr(table), rows, cols, _ptr(digests),
                   self.stream_ptr(stream))

    def hash_xfield_rows(self, table, rows, xcols, digests, stream=None):
        self._call("tvh_hash_xfield_rows_dev", _ptr(table), rows, xcols, _ptr(digests),
                   self.stream_ptr(stream))

    def permutation_batch(self, states, num, stream=None):
        self._call("tvh_tip5_permutation_batch_dev", _ptr(states), num, self.stream_ptr(stream))

    def tip5_rounds(self, states, num, first, count, trace=None, stream=None):
        self._call("tvh_tip5_rounds_dev", _ptr(states), num, first, count, _ptr(trace),
                   self.stream_ptr(stream))

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

    # -- FRI leaves and the Fiat-Shamir sponge -------------------------------------
    def xfe_leaves(self, xfe, n, leaves, stream=None):
        self._call("tvh_xfe_leaves_dev", _ptr(xfe), n, _ptr(leaves), self.stream_ptr(stream))

    def fri_commit(self, xfe, n, tree, layout=TREE_LEVEL_ORDER, stream=None):
        self._call("tvh_fri_commit_dev", _ptr(xfe), n, _ptr(tree), layout, self.stream_ptr(stream))

    def sponge_init(self, state, stream=None):
        self._call("tvh_sponge_init_dev", _ptr(state), self.stream_ptr(stream))

    def sponge_absorb(self, state, data, length, stream=None):
        self._call("tvh_sponge_absorb_dev", _ptr(state), _ptr(data), length,
                   self.stream_ptr(stream))

    def sponge_squeeze(self, state, out, stream=None):
        self._call("tvh_sponge_squeeze_dev", _ptr(state), _ptr(out), self.stream_ptr(stream))

    # -- batches ----------------------------------------------------------------
    @staticmethod
    def table_desc(trace, cols, n, m, codewords, tree, rand=None, k=0, toff=1, goff=7,
                   layout=TREE_LEVEL_ORDER):
        return TableDesc(_ptr(trace), cols, n, _ptr(rand), k, toff, goff, m, _ptr(codewords),
                         _ptr(tree), layout)

    def batch_plan(self, descs):
        arr = (TableDesc * max(1, len(descs)))(*descs)
        plan = C.c_void_p()
        self._call("tvh_batch_plan_create", C.cast(arr, C.c_void_p), len(descs), C.byref(plan))
        return plan

    def batch_launch(self, plan, stream=None):
        self._call("tvh_batch_plan_launch", plan, self.stream_ptr(stream))

    def batch_launch_graph(self, plan, stream):
        self._call("tvh_batch_plan_launch_graph", plan, self.stream_ptr(stream))

    def batch_destroy(self, plan):
        self._call("tvh_batch_plan_destroy", plan)

    def commit_batch(self, descs, stream=None):
        arr = (TableDesc * max(1, len(descs)))(*descs)
        self._call("tvh_commit_batch_dev", C.cast(arr, C.c_void_p), len(descs),
                   self.stream_ptr(stream))

    def fill_uniform(self, d, count, seed, stream=None):
        self._call("tvh_fill_uniform_dev", _ptr(d), count, seed, self.stream_ptr(stream))

    def upload(self, d, host, count, stream=None):
        """count u64 from host memory (numpy array or address) to device buffer d."""
        self._call("tvh_upload_dev", _ptr(d), _ptr(host), count, self.stream_ptr(stream))

    def fill_table(self, d, rows, cols, seed, stream=None):
        self._call("tvh_fill_table_dev", _ptr(d), rows, cols, seed, self.stream_ptr(stream))

    def commit_frugal(self, trace, cols, n, rand, k, toff, goff, m, chunk, work, tree,
                      layout=TREE_LEVEL_ORDER, stream=None):
        self._call("tvh_commit_frugal_dev", _ptr(trace), cols, n, _ptr(rand), k, toff, goff, m,
                   chunk, _ptr(work), _ptr(tree), layout, self.stream_ptr(stream))

    def frugal_workspace_elems(self, cols, n, m, chunk):
        return self.lib.tvh_commit_frugal_workspace_elems(cols, n, m, chunk)

    # -- LDE kernel forms -----------------------------------------------------------
    def lde_set_variant(self, log_n, log_e, variant):
        self._call("tvh_lde_set_variant", log_n, log_e, variant)

    def lde_get_variant(self, log_n, log_e, cols=None):
        v = C.c_int()
        if cols is None:
            self._call("tvh_lde_get_variant", log_n, log_e, C.byref(v))
        else:
            self._call("tvh_lde_get_variant_for_width", log_n, log_e, cols, C.byref(v))
        return v.value

    def lde_autotune(self, cols, n, m):
        """-> (chosen variant code, {variant code: ms per LDE})"""
        chosen = C.c_int()
        ms = (C.c_float * self.lib.tvh_lde_variant_count())()
        self._call("tvh_lde_autotune", cols, n, m, C.byref(chosen), ms)
        return chosen.value, dict(zip(LDE_VARIANTS, list(ms)))

    # -- commit server ------------------------------------------------------------
    def server_create(self, hbm_budget_bytes=0, max_inflight=0, max_queue=0):
        cfg = ServerConfig(hbm_budget_bytes, max_inflight, max_queue)
        srv = C.c_void_p()
        self._call("tvh_server_create", C.byref(cfg), C.byref(srv))
        return srv

    def server_submit(self, srv, cols, n, m, seed=0, trace=None, rand=None, k=0, toff=1, goff=7):
        """trace / rand: numpy uint64 (column-major) kept alive by the caller until the
        job finished, or None for device-drawn inputs from `seed`."""
        d = JobDesc(_ptr(trace), _ptr(rand), seed, cols, n, k, toff, goff, m)
        jid = C.c_uint64()
        self._call("tvh_server_submit", srv, C.byref(d), C.byref(jid))
        return jid.value

    def server_poll(self, srv, jid):
        st = C.c_int()
        self._call("tvh_server_poll", srv, jid, C.byref(st))
        return st.value

    def server_wait(self, srv, jid):
        """-> (root as 5 ints, (submit_s, admit_s, done_s))"""
        root = np.zeros(5, np.uint64)
        t = JobTimes()
        self._call("tvh_server_wait", srv, jid, root.ctypes.data, C.byref(t))
        return [int(v) for v in root], (t.submit_s, t.admit_s, t.done_s)

    def server_open(self, srv, jid, leaves, cols, want_rows=True, want_auth=True):
        """-> (rows count x cols or None, auth structure digests k x 5 or None)"""
        li = np.ascontiguousarray(leaves, dtype=np.uint64)
        cnt = C.c_size_t()
        lp = li.ctypes.data if li.size else None
        auth = None
        if want_auth:
            self._call("tvh_server_open", srv, jid, lp, li.size, None, None, 0, C.byref(cnt))
            auth = np.zeros((max(1, cnt.value), 5), np.uint64)
        rows = np.zeros((max(1, li.size), max(1, cols)), np.uint64) if want_rows else None
        self._call("tvh_server_open", srv, jid, lp, li.size, _ptr(rows), _ptr(auth),
                   cnt.value, C.byref(cnt))
        return (rows[: li.size, :cols] if want_rows else None), (auth[: cnt.value] if want_auth
                                                                 else None)

    def server_paths(self, srv, jid, leaves, m):
        li = np.ascontiguousarray(leaves, dtype=np.uint64)
        h = int(m).bit_length() - 1
        out = np.zeros((max(1, li.size), max(1, h), 5), np.uint64)
        self._call("tvh_server_paths", srv, jid, li.ctypes.data if li.size else None, li.size,
                   out.ctypes.data)
        return out[: li.size, :h]

    def server_buffers(self, srv, jid):
        cw, tree = C.c_void_p(), C.c_void_p()
        self._call("tvh_server_buffers", srv, jid, C.byref(cw), C.byref(tree))
        return cw.value, tree.value

    def server_release(self, srv, jid):
        self._call("tvh_server_release", srv, jid)

    def server_stats(self, srv):
        v = [C.c_size_t() for _ in range(4)]
        self._call("tvh_server_stats", srv, *[C.byref(x) for x in v])
        return dict(zip(("arena_used", "arena_size", "queued", "running"), (x.value for x in v)))

    def server_destroy(self, srv):
        self._call("tvh_server_destroy", srv)

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
