"""ctypes front end of libcl_ops.so for Python callers, tests and benchmarks.

The library itself is the C API in include/*.h; this module only declares its
signatures, turns GError results into exceptions and moves numpy arrays in and
out of device buffers.  It never computes anything on the host: if the HIP
library is missing it raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# CLOPS_LIB selects another build of the same library (the sanitizer build,
# lib/libcl_ops_san.so); the default is the in-tree build.
LIB_PATH = os.environ.get("CLOPS_LIB") or os.path.join(HERE, "lib", "libcl_ops.so")

CL_QUEUE_PROFILING_ENABLE = 1 << 1

# CloType (include/clo_common.h)
CLO_CHAR, CLO_UCHAR, CLO_SHORT, CLO_USHORT, CLO_INT, CLO_UINT, CLO_LONG, CLO_ULONG = range(8)
CLO_HALF, CLO_FLOAT, CLO_DOUBLE = 8, 9, 10

# enum clo_error_codes
CLO_SUCCESS, CLO_ERROR_OPENFILE, CLO_ERROR_ARGS, CLO_ERROR_STREAM_WRITE = 0, 1, 2, 3
CLO_ERROR_IMPL_NOT_FOUND, CLO_ERROR_UNKNOWN_TYPE, CLO_ERROR_LIBRARY = 5, 6, 7

# CloRngSeedType
SEED_DEV_GID, SEED_HOST_MT, SEED_EXT_DEV, SEED_EXT_HOST = 0, 1, 2, 3

# CloGenKind (include/clo_util.h)
GEN_U64, GEN_U32_HI, GEN_U64_MASK, GEN_U32_MASK, GEN_IOTA_U32, GEN_ZIPF_U64 = range(6)
GEN_U32_MOD, GEN_U64_MOD = 6, 7
GEN_NEARLY_SORTED_U32, GEN_NEARLY_SORTED_U64, GEN_REVERSED_U32, GEN_REVERSED_U64 = 8, 9, 10, 11

_DTYPE_TO_CLO = {
    np.dtype(np.int8): CLO_CHAR, np.dtype(np.uint8): CLO_UCHAR,
    np.dtype(np.int16): CLO_SHORT, np.dtype(np.uint16): CLO_USHORT,
    np.dtype(np.int32): CLO_INT, np.dtype(np.uint32): CLO_UINT,
    np.dtype(np.int64): CLO_LONG, np.dtype(np.uint64): CLO_ULONG,
    np.dtype(np.float16): CLO_HALF, np.dtype(np.float32): CLO_FLOAT,
    np.dtype(np.float64): CLO_DOUBLE,
}
_CLO_TO_DTYPE = {v: k for k, v in _DTYPE_TO_CLO.items()}


class GError(ctypes.Structure):
    _fields_ = [("domain", ctypes.c_uint32), ("code", ctypes.c_int), ("message", ctypes.c_char_p)]


GErrorP = ctypes.POINTER(GError)


class ClopsError(RuntimeError):
    def __init__(self, domain, code, message):
        super().__init__("[%#x/%d] %s" % (domain, code, message))
        self.domain = domain
        self.code = code
        self.message = message


_lib = None


def lib():
    """Loads libcl_ops.so (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libcl_ops.so not built: run __graft_entry__.build() or "
                           "make -C cl-ops_amd")
    L = ctypes.CDLL(LIB_PATH)
    vp, sz, u32, u64, i32, cp = (ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint32,
                                 ctypes.c_uint64, ctypes.c_int, ctypes.c_char_p)
    ep = ctypes.POINTER(GErrorP)
    d = ctypes.c_double
    sig = {
        # errors and common
        "clo_error_free": (None, [GErrorP]),
        "clo_error_quark": (u32, []),
        "ccl_error_quark": (u32, []),
        "clo_type_get_name": (cp, [i32]),
        "clo_type_sizeof": (sz, [i32]),
        "clo_type_by_name": (i32, [cp, ep]),
        "clo_nlpo2": (ctypes.c_uint, [ctypes.c_uint]),
        "clo_ones32": (ctypes.c_uint, [ctypes.c_uint]),
        "clo_tzc": (ctypes.c_uint, [ctypes.c_int]),
        "clo_sum": (ctypes.c_uint, [ctypes.c_uint]),
        "ccl_debug_fail_next_allocs": (None, [i32]),
        "clo_scan_debug_lookback_polls": (None, [i32]),
        # thin HIP layer
        "ccl_context_new_from_device_index": (vp, [i32, ep]),
        "ccl_context_destroy": (None, [vp]),
        "ccl_queue_new": (vp, [vp, vp, u64, ep]),
        "ccl_queue_finish": (u32, [vp, ep]),
        "ccl_queue_gc": (None, [vp]),
        "ccl_queue_destroy": (None, [vp]),
        "ccl_queue_enqueue_marker": (vp, [vp, ep]),
        "ccl_enqueue_barrier": (vp, [vp, vp, ep]),
        "ccl_event_wait_list_add": (None, [vp, vp, vp]),
        "ccl_event_wait_list_clear": (None, [vp]),
        "ccl_queue_get_stream": (vp, [vp]),
        "ccl_buffer_new": (vp, [vp, u64, sz, vp, ep]),
        "ccl_buffer_destroy": (None, [vp]),
        "ccl_buffer_get_device_ptr": (vp, [vp]),
        "ccl_buffer_get_size": (sz, [vp]),
        "ccl_buffer_new_wrap": (vp, [vp, vp, sz]),
        "ccl_buffer_enqueue_write": (vp, [vp, vp, u32, sz, sz, vp, vp, ep]),
        "ccl_buffer_enqueue_read": (vp, [vp, vp, u32, sz, sz, vp, vp, ep]),
        "ccl_buffer_enqueue_copy": (vp, [vp, vp, vp, sz, sz, sz, vp, ep]),
        "ccl_event_elapsed_ms": (d, [vp, vp, ep]),
        "ccl_event_get_name": (cp, [vp]),
        "ccl_prof_reset": (None, [vp]),
        "ccl_prof_get": (sz, [vp, cp, ctypes.POINTER(d), ep]),
        "ccl_prof_get_name": (cp, [vp, sz]),
        "ccl_prof_set_filter": (None, [vp, cp]),
        # sort
        "clo_sort_new": (vp, [cp, cp, vp, ctypes.POINTER(i32), ctypes.POINTER(i32), cp, cp, cp,
                              ep]),
        "clo_sort_destroy": (None, [vp]),
        "clo_sort_with_device_data": (vp, [vp, vp, vp, vp, vp, sz, sz, ep]),
        "clo_sort_with_host_data": (u32, [vp, vp, vp, vp, vp, sz, sz, ep]),
        "clo_sort_kv_with_device_data": (vp, [vp, vp, vp, vp, vp, vp, vp, sz, sz, ep]),
        "clo_sort_kv_with_host_data": (u32, [vp, vp, vp, vp, vp, sz, sz, ep]),
        "clo_sort_get_last_passes": (u32, [vp, ep]),
        "clo_sort_segmented_with_device_data": (vp, [vp, vp, vp, vp, vp, vp, vp, vp, sz, sz, ep]),
        "clo_sort_get_element_type": (i32, [vp]),
        "clo_sort_get_element_size": (sz, [vp]),
        "clo_sort_get_key_type": (i32, [vp]),
        "clo_sort_get_source": (cp, [vp]),
        "clo_sort_get_key_size": (sz, [vp]),
        "clo_sort_get_num_kernels": (u32, [vp, ep]),
        "clo_sort_get_kernel_name": (cp, [vp, u32, ep]),
        "clo_sort_get_localmem_usage": (sz, [vp, u32, sz, sz, ep]),
        "clo_sort_get_context": (vp, [vp]),
        "clo_sort_abitonic_get_last_sweeps": (u32, [vp]),
        # scan
        "clo_scan_new": (vp, [cp, cp, vp, i32, i32, cp, ep]),
        "clo_scan_destroy": (None, [vp]),
        "clo_scan_with_device_data": (vp, [vp, vp, vp, vp, vp, sz, sz, ep]),
        "clo_scan_with_host_data": (u32, [vp, vp, vp, vp, vp, sz, sz, ep]),
        "clo_scan_segmented_with_device_data": (vp, [vp, vp, vp, vp, vp, vp, sz, sz, ep]),
        "clo_scan_get_elem_type": (i32, [vp]),
        "clo_scan_get_sum_type": (i32, [vp]),
        "clo_scan_get_element_size": (sz, [vp]),
        "clo_scan_get_sum_size": (sz, [vp]),
        "clo_scan_get_num_kernels": (u32, [vp, ep]),
        "clo_scan_get_kernel_name": (cp, [vp, u32, ep]),
        "clo_scan_get_localmem_usage": (sz, [vp, u32, sz, sz, ep]),
        # rng
        "clo_rng_new": (vp, [cp, i32, vp, sz, u64, cp, vp, vp, ep]),
        "clo_rng_destroy": (None, [vp]),
        "clo_rng_get_source": (cp, [vp]),
        "clo_rng_get_device_seeds": (vp, [vp]),
        "clo_rng_get_size": (sz, [vp]),
        "clo_rng_fill": (vp, [vp, vp, vp, u32, u32, u32, ep]),
        # util
        "clo_util_generate": (vp, [vp, vp, sz, i32, u64, u64, u64, ep]),
        "clo_util_generate_zipf": (vp, [vp, vp, sz, vp, sz, u64, ep]),
        "clo_util_generate_zipf_u32": (vp, [vp, vp, sz, vp, sz, u64, ep]),
        "clo_util_zipf_cdf": (None, [vp, sz, d]),
        "clo_util_copy": (vp, [vp, vp, vp, sz, ep]),
        "clo_util_stream_copy": (vp, [vp, vp, vp, sz, ep]),
        "clo_util_scatter_runs": (vp, [vp, vp, vp, vp, vp, sz, u32, ep]),
        "clo_util_multiset_hash": (u64, [vp, vp, sz, sz, ep]),
        "clo_util_check_sorted": (sz, [vp, vp, vp, sz, sz, ep]),
        "clo_util_build_info": (cp, []),
        "clo_util_count_diff": (sz, [vp, vp, vp, sz, ep]),
        "clo_util_check_pairs": (sz, [vp, vp, vp, vp, sz, sz, ep]),
        "clo_scan_with_device_data_noalloc": (u32, [vp, vp, vp, vp, sz, ep]),
        "clo_scan_noalloc_max_numel": (sz, [vp]),
        # dispatcher
        "clo_dispatch_new": (vp, [vp, ep]),
        "clo_dispatch_destroy": (None, [vp]),
        "clo_dispatch_config_name": (cp, [u32]),
        "clo_dispatch_config_feasible": (u32, [u32, i32, sz, sz]),
        "clo_dispatch_fixed_config": (u32, [i32, sz, sz]),
        "clo_dispatch_features": (u32, [vp, vp, i32, vp, sz, sz, vp, ep]),
        "clo_dispatch_model_row": (None, [vp, u32, vp]),
        "clo_dispatch_run": (vp, [vp, vp, u32, i32, vp, vp, sz, sz, ep]),
        "clo_dispatch_tune": (u32, [vp, vp, i32, vp, vp, sz, sz, u32, vp, ep]),
        "clo_dispatch_table_add": (None, [vp, i32, sz, sz, vp, vp]),
        "clo_dispatch_table_choose": (i32, [vp, i32, sz, sz, vp]),
        "clo_dispatch_table_save": (u32, [vp, cp, ep]),
        "clo_dispatch_table_load": (u32, [vp, cp, ep]),
        "clo_dispatch_set_model": (u32, [vp, vp, ep]),
        "clo_dispatch_set_margin": (None, [vp, ctypes.c_float]),
        "clo_dispatch_choose": (u32, [vp, vp, i32, sz, vp, vp, vp, vp, vp, ep]),
        "clo_dispatch_features_segmented": (u32, [vp, vp, vp, vp, sz, sz, sz, vp, ep]),
        "clo_dispatch_features_batch": (u32, [vp, vp, sz, vp, vp, vp, vp, vp, ep]),
        "clo_dispatch_run_segmented": (vp, [vp, vp, u32, vp, vp, vp, sz, sz, sz, vp, ep]),
        "clo_dispatch_tune_segmented": (u32, [vp, vp, vp, vp, vp, sz, sz, sz, u32, vp, ep]),
        "clo_dispatch_seg_bin": (u32, [sz]),
        "clo_dispatch_seg_bin_length": (sz, [u32]),
        "clo_dispatch_tune_segment_class": (u32, [vp, vp, u32, sz, sz, u32, vp, ep]),
        "clo_dispatch_segment_row": (None, [sz, sz, u32, u32, vp]),
        "clo_dispatch_set_segment_model": (u32, [vp, vp, ep]),
        "clo_dispatch_segment_classes": (u32, [vp, vp, u32, sz, ctypes.c_float, vp, ep]),
        # cost-model MLP
        "clo_mlp_new": (vp, [vp, u32, u64, ep]),
        "clo_mlp_destroy": (None, [vp]),
        "clo_mlp_num_params": (u32, [vp]),
        "clo_mlp_num_inputs": (u32, [vp]),
        "clo_mlp_num_hidden": (u32, [vp]),
        "clo_mlp_get_params": (u32, [vp, vp, ep]),
        "clo_mlp_set_params": (u32, [vp, vp, ep]),
        "clo_mlp_gradient": (u32, [vp, vp, vp, vp, sz, vp, vp, ep]),
        "clo_mlp_train": (u32, [vp, vp, vp, vp, sz, u32, ctypes.c_float, vp, ep]),
        "clo_mlp_predict": (u32, [vp, vp, vp, sz, vp, ep]),
        "clo_mlp_predict_device": (vp, [vp, vp, vp, sz, vp, ep]),
        "clo_mlp_save": (u32, [vp, cp, ep]),
        "clo_mlp_load": (vp, [vp, cp, ep]),
        # graphs
        "clo_graph_begin_capture": (u32, [vp, ep]),
        "clo_graph_end_capture": (vp, [vp, ep]),
        "clo_graph_launch": (vp, [vp, vp, ep]),
        "clo_graph_get_num_nodes": (sz, [vp]),
        "clo_graph_destroy": (None, [vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _call(fn, *args):
    """Calls an API function whose last argument is GError**; raises on error."""
    err = GErrorP()
    r = fn(*args, ctypes.byref(err))
    if err:
        e = err.contents
        exc = ClopsError(e.domain, e.code, (e.message or b"").decode())
        lib().clo_error_free(err)
        raise exc
    return r


def clo_type(dtype):
    return _DTYPE_TO_CLO[np.dtype(dtype)]


def dtype_of(clo_t):
    return _CLO_TO_DTYPE[clo_t]


class Context:
    def __init__(self, device=0):
        self.L = lib()
        self.ptr = _call(self.L.ccl_context_new_from_device_index, device)

    def close(self):
        if self.ptr:
            self.L.ccl_context_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Queue:
    def __init__(self, ctx, profiling=False):
        self.L = lib()
        self.ctx = ctx
        self.ptr = _call(self.L.ccl_queue_new, ctx.ptr, None,
                         CL_QUEUE_PROFILING_ENABLE if profiling else 0)

    def finish(self):
        _call(self.L.ccl_queue_finish, self.ptr)

    def gc(self):
        """Waits, then releases the events recorded on the queue so far."""
        self.L.ccl_queue_gc(self.ptr)

    def marker(self):
        return _call(self.L.ccl_queue_enqueue_marker, self.ptr)

    def wait_for(self, *events):
        """Later work on this queue waits for the events (of any queue)."""
        ewl = ctypes.c_void_p(None)
        for e in events:
            self.L.ccl_event_wait_list_add(ctypes.byref(ewl), e, None)
        return _call(self.L.ccl_enqueue_barrier, self.ptr, ctypes.byref(ewl))

    def elapsed_ms(self, e0, e1):
        return _call(self.L.ccl_event_elapsed_ms, e0, e1)

    def prof_reset(self):
        self.L.ccl_prof_reset(self.ptr)

    def prof(self, prefix=""):
        """(count, total_ms) of the profiled kernels whose name starts with prefix."""
        tot = ctypes.c_double()
        n = _call(self.L.ccl_prof_get, self.ptr, prefix.encode(), ctypes.byref(tot))
        return n, tot.value

    def prof_filter(self, prefix=""):
        """Record events only around kernels whose name starts with prefix
        ("" = every kernel)."""
        self.L.ccl_prof_set_filter(self.ptr, prefix.encode())

    def prof_names(self):
        out, i = [], 0
        while True:
            s = self.L.ccl_prof_get_name(self.ptr, i)
            if s is None:
                return out
            out.append(s.decode())
            i += 1

    def close(self):
        if self.ptr:
            self.L.ccl_queue_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _f32(a, cols=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if cols is not None:
        a = a.reshape(-1, cols)
    return a


class Mlp:
    """FP32 cost model n_in -> 32 -> 32 -> 1 trained and run on the GPU."""

    def __init__(self, ctx, n_in=None, seed=0, path=None):
        self.L = lib()
        self.ctx = ctx
        if path is not None:
            self.ptr = _call(self.L.clo_mlp_load, ctx.ptr, _b(path))
        else:
            self.ptr = _call(self.L.clo_mlp_new, ctx.ptr, n_in, seed)
        self.n_in = self.L.clo_mlp_num_inputs(self.ptr)
        self.count = self.L.clo_mlp_num_params(self.ptr)

    def params(self):
        p = np.empty(self.count, dtype=np.float32)
        _call(self.L.clo_mlp_get_params, self.ptr, p.ctypes.data_as(ctypes.c_void_p))
        return p

    def set_params(self, p):
        p = _f32(p)
        assert p.size == self.count
        _call(self.L.clo_mlp_set_params, self.ptr, p.ctypes.data_as(ctypes.c_void_p))

    def gradient(self, q, X, t):
        X, t = _f32(X, self.n_in), _f32(t)
        g = np.empty(self.count, dtype=np.float32)
        loss = ctypes.c_float()
        _call(self.L.clo_mlp_gradient, self.ptr, q.ptr, X.ctypes.data_as(ctypes.c_void_p),
              t.ctypes.data_as(ctypes.c_void_p), t.size, g.ctypes.data_as(ctypes.c_void_p),
              ctypes.byref(loss))
        return float(loss.value), g

    def train(self, q, X, t, steps, lr):
        X, t = _f32(X, self.n_in), _f32(t)
        loss = ctypes.c_float()
        _call(self.L.clo_mlp_train, self.ptr, q.ptr, X.ctypes.data_as(ctypes.c_void_p),
              t.ctypes.data_as(ctypes.c_void_p), t.size, steps, lr, ctypes.byref(loss))
        return float(loss.value)

    def predict(self, q, X):
        X = _f32(X, self.n_in)
        out = np.empty(X.shape[0], dtype=np.float32)
        _call(self.L.clo_mlp_predict, self.ptr, q.ptr, X.ctypes.data_as(ctypes.c_void_p),
              X.shape[0], out.ctypes.data_as(ctypes.c_void_p))
        return out

    def save(self, path):
        _call(self.L.clo_mlp_save, self.ptr, _b(path))

    def close(self):
        if self.ptr:
            self.L.clo_mlp_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


OP_SORT, OP_SCAN, OP_SEGSORT = 0, 1, 2
POLICY_FIXED, POLICY_TABLE, POLICY_MODEL = 0, 1, 2
N_CONFIGS, N_FEATURES = 24, 10
MODEL_INPUTS = N_FEATURES + N_CONFIGS
SEG_BINS, SEG_INPUTS = 64, 7


def _vp(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Dispatch:
    """Configuration choice for sort, scan and segmented sort requests
    (clo_dispatch_*), and the per-segment work class."""

    def __init__(self, ctx):
        self.L = lib()
        self.ctx = ctx
        self.ptr = _call(self.L.clo_dispatch_new, ctx.ptr)
        self.model = None
        self.seg_model = None

    @staticmethod
    def config_name(c):
        return lib().clo_dispatch_config_name(c).decode()

    @staticmethod
    def feasible(c, op, n, width):
        return bool(lib().clo_dispatch_config_feasible(c, op, n, width))

    @staticmethod
    def fixed(op, n, width):
        return lib().clo_dispatch_fixed_config(op, n, width)

    @staticmethod
    def model_row(features, cfg):
        f = _f32(features)
        row = np.empty(MODEL_INPUTS, dtype=np.float32)
        lib().clo_dispatch_model_row(_vp(f), cfg, _vp(row))
        return row

    def features(self, q, op, buf, n, width):
        f = np.empty(N_FEATURES, dtype=np.float32)
        _call(self.L.clo_dispatch_features, self.ptr, q.ptr, op, buf.ptr, n, width, _vp(f))
        return f

    def run(self, q, cfg, op, bin_, bout, n, width):
        return _call(self.L.clo_dispatch_run, self.ptr, q.ptr, cfg, op, bin_.ptr, bout.ptr, n,
                     width)

    def tune(self, q, op, bin_, bout, n, width, reps=3):
        t = np.empty(N_CONFIGS, dtype=np.float32)
        _call(self.L.clo_dispatch_tune, self.ptr, q.ptr, op, bin_.ptr, bout.ptr, n, width, reps,
              _vp(t))
        return t

    def features_batch(self, q, ops, bufs, ns, widths):
        count = len(bufs)
        f = np.empty((count, N_FEATURES), dtype=np.float32)
        if count == 0:
            return f
        o = np.ascontiguousarray(ops, dtype=np.int32)
        p = (ctypes.c_void_p * count)(*[b.ptr for b in bufs])
        n = np.ascontiguousarray(ns, dtype=np.uint64)
        w = np.ascontiguousarray(widths, dtype=np.uint64)
        _call(self.L.clo_dispatch_features_batch, self.ptr, q.ptr, count, _vp(o), p, _vp(n),
              _vp(w), _vp(f))
        return f

    def features_segmented(self, q, buf, off, nseg, n, width):
        f = np.empty(N_FEATURES, dtype=np.float32)
        _call(self.L.clo_dispatch_features_segmented, self.ptr, q.ptr, buf.ptr, off.ptr, nseg, n,
              width, _vp(f))
        return f

    def run_segmented(self, q, cfg, bin_, bout, off, nseg, n, width, classes=None):
        c = None
        if classes is not None:
            c = np.ascontiguousarray(classes, dtype=np.uint8)
            assert c.size == SEG_BINS
        return _call(self.L.clo_dispatch_run_segmented, self.ptr, q.ptr, cfg, bin_.ptr, bout.ptr,
                     off.ptr, nseg, n, width, _vp(c) if c is not None else None)

    def tune_segmented(self, q, bin_, bout, off, nseg, n, width, reps=3):
        t = np.empty(N_CONFIGS, dtype=np.float32)
        _call(self.L.clo_dispatch_tune_segmented, self.ptr, q.ptr, bin_.ptr, bout.ptr, off.ptr,
              nseg, n, width, reps, _vp(t))
        return t

    @staticmethod
    def seg_bin(length):
        return lib().clo_dispatch_seg_bin(length)

    @staticmethod
    def seg_bin_length(b):
        return lib().clo_dispatch_seg_bin_length(b)

    def tune_segment_class(self, q, cfg, width, length, reps=3):
        t = np.empty(3, dtype=np.float32)
        _call(self.L.clo_dispatch_tune_segment_class, self.ptr, q.ptr, cfg, width, length, reps,
              _vp(t))
        return t

    @staticmethod
    def segment_row(length, width, cfg, cls):
        row = np.empty(SEG_INPUTS, dtype=np.float32)
        lib().clo_dispatch_segment_row(length, width, cfg, cls, _vp(row))
        return row

    def set_segment_model(self, mlp):
        self.seg_model = mlp
        _call(self.L.clo_dispatch_set_segment_model, self.ptr, mlp.ptr if mlp else None)

    def segment_classes(self, q, cfg, width, log2_margin=0.0):
        c = np.empty(SEG_BINS, dtype=np.uint8)
        _call(self.L.clo_dispatch_segment_classes, self.ptr, q.ptr if q else None, cfg, width,
              log2_margin, _vp(c))
        return c

    def table_add(self, op, n, width, features, times):
        f, t = _f32(features), _f32(times)
        self.L.clo_dispatch_table_add(self.ptr, op, n, width, _vp(f), _vp(t))

    def table_choose(self, op, n, width, features):
        f = _f32(features)
        return self.L.clo_dispatch_table_choose(self.ptr, op, n, width, _vp(f))

    def table_save(self, path):
        _call(self.L.clo_dispatch_table_save, self.ptr, _b(path))

    def table_load(self, path):
        _call(self.L.clo_dispatch_table_load, self.ptr, _b(path))

    def set_model(self, mlp):
        self.model = mlp
        _call(self.L.clo_dispatch_set_model, self.ptr, mlp.ptr if mlp else None)

    def set_margin(self, log2_margin):
        self.L.clo_dispatch_set_margin(self.ptr, log2_margin)

    def choose(self, q, policy, ops, ns, widths, features):
        ops = np.ascontiguousarray(ops, dtype=np.int32)
        ns = np.ascontiguousarray(ns, dtype=np.uint64)
        ws = np.ascontiguousarray(widths, dtype=np.uint64)
        f = _f32(features, N_FEATURES)
        out = np.empty(ops.size, dtype=np.int32)
        _call(self.L.clo_dispatch_choose, self.ptr, q.ptr if q else None, policy, ops.size,
              _vp(ops), _vp(ns), _vp(ws), _vp(f), _vp(out))
        return out

    def close(self):
        if self.ptr:
            self.L.clo_dispatch_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def count_diff(q, a, b, words):
    return _call(lib().clo_util_count_diff, q.ptr, a.ptr, b.ptr, words)


class Graph:
    """Work enqueued on `q` between begin() and end() (clo_graph_*)."""

    def __init__(self, q):
        self.L = lib()
        self.q = q
        self.ptr = None

    def begin(self):
        # A finaliser that frees device memory (hipFree) on this thread while
        # the capture is open would invalidate it: collect garbage now and
        # keep the cyclic collector off until end().
        import gc
        gc.collect()
        self._gc_was_enabled = gc.isenabled()
        gc.disable()
        try:
            _call(self.L.clo_graph_begin_capture, self.q.ptr)
        except Exception:
            if self._gc_was_enabled:
                gc.enable()
            raise

    def end(self):
        import gc
        try:
            self.ptr = _call(self.L.clo_graph_end_capture, self.q.ptr)
        finally:
            if getattr(self, "_gc_was_enabled", False):
                gc.enable()
        return self

    def launch(self, q=None):
        return _call(self.L.clo_graph_launch, self.ptr, (q or self.q).ptr)

    def num_nodes(self):
        return self.L.clo_graph_get_num_nodes(self.ptr)

    def close(self):
        if self.ptr:
            self.L.clo_graph_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Buffer:
    def __init__(self, ctx, nbytes, host=None):
        self.L = lib()
        self.ctx = ctx
        self.nbytes = int(nbytes)
        hp = None
        if host is not None:
            host = np.ascontiguousarray(host)
            assert host.nbytes <= self.nbytes
            self._keep = host
            hp = host.ctypes.data_as(ctypes.c_void_p)
        self.ptr = _call(self.L.ccl_buffer_new, ctx.ptr, 1, self.nbytes, hp)

    @classmethod
    def from_array(cls, ctx, a):
        a = np.ascontiguousarray(a)
        return cls(ctx, max(a.nbytes, 1), a)

    def view(self, offset, nbytes):
        """A buffer over bytes [offset, offset+nbytes) of this one (no copy;
        the view must not outlive this buffer)."""
        v = Buffer.__new__(Buffer)
        v.L, v.ctx, v.nbytes, v._parent = self.L, self.ctx, int(nbytes), self
        base = self.L.ccl_buffer_get_device_ptr(self.ptr)
        v.ptr = self.L.ccl_buffer_new_wrap(self.ctx.ptr, ctypes.c_void_p(base + int(offset)),
                                           max(int(nbytes), 1))
        if not v.ptr:
            raise ClopsError(0, 0, "ccl_buffer_new_wrap failed")
        return v

    def write(self, q, a, offset=0):
        a = np.ascontiguousarray(a)
        _call(self.L.ccl_buffer_enqueue_write, self.ptr, q.ptr, 1, offset, a.nbytes,
              a.ctypes.data_as(ctypes.c_void_p), None)

    def read(self, q, dtype, count, offset=0):
        out = np.empty(count, dtype=dtype)
        if out.nbytes:
            _call(self.L.ccl_buffer_enqueue_read, self.ptr, q.ptr, 1, offset, out.nbytes,
                  out.ctypes.data_as(ctypes.c_void_p), None)
        return out

    def close(self):
        if self.ptr:
            self.L.ccl_buffer_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _b(s):
    return None if s is None else s.encode()


class Sorter:
    def __init__(self, ctx, algo, dtype, options=None, compare=None, get_key=None,
                 key_dtype=None, compiler_opts=None):
        self.L = lib()
        self.ctx = ctx
        self.dtype = np.dtype(dtype)
        et = ctypes.c_int(clo_type(dtype))
        kt = None if key_dtype is None else ctypes.byref(ctypes.c_int(clo_type(key_dtype)))
        self.ptr = _call(self.L.clo_sort_new, _b(algo), _b(options), ctx.ptr, ctypes.byref(et),
                         kt, _b(compare), _b(get_key), _b(compiler_opts))

    def source(self):
        """HIP source of the run-time compiled kernels (None: precompiled)."""
        s = self.L.clo_sort_get_source(self.ptr)
        return None if s is None else s.decode()

    def sort_device(self, q, bin_, bout, n, q_comm=None):
        return _call(self.L.clo_sort_with_device_data, self.ptr, q.ptr,
                     q_comm.ptr if q_comm else None, bin_.ptr, bout.ptr if bout else None, n, 0)

    def sort_kv_device(self, q, kin, vin, kout=None, vout=None, n=0):
        return _call(self.L.clo_sort_kv_with_device_data, self.ptr, q.ptr, None, kin.ptr, vin.ptr,
                     kout.ptr if kout else None, vout.ptr if vout else None, n, 0)

    def sort_host(self, a, q=None, q_comm=None):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        out = np.empty_like(a)
        _call(self.L.clo_sort_with_host_data, self.ptr, q.ptr if q else None,
              q_comm.ptr if q_comm else None,
              a.ctypes.data_as(ctypes.c_void_p), out.ctypes.data_as(ctypes.c_void_p), a.size, 0)
        return out

    def sort_kv_host(self, keys, vals, q=None, q_comm=None):
        k = np.array(keys, dtype=self.dtype, copy=True)
        v = np.array(vals, dtype=np.uint32, copy=True)
        _call(self.L.clo_sort_kv_with_host_data, self.ptr, q.ptr if q else None,
              q_comm.ptr if q_comm else None,
              k.ctypes.data_as(ctypes.c_void_p), v.ctypes.data_as(ctypes.c_void_p), k.size, 0)
        return k, v

    def sort_segmented_device(self, q, kin, offsets, nseg, n, vin=None, kout=None, vout=None):
        return _call(self.L.clo_sort_segmented_with_device_data, self.ptr, q.ptr, None, kin.ptr,
                     vin.ptr if vin else None, kout.ptr if kout else None,
                     vout.ptr if vout else None, offsets.ptr, nseg, n)

    def last_passes(self):
        return _call(self.L.clo_sort_get_last_passes, self.ptr)

    def last_sweeps(self):
        return self.L.clo_sort_abitonic_get_last_sweeps(self.ptr)

    def kernel_names(self):
        n = _call(self.L.clo_sort_get_num_kernels, self.ptr)
        return [_call(self.L.clo_sort_get_kernel_name, self.ptr, i).decode() for i in range(n)]

    def close(self):
        if self.ptr:
            self.L.clo_sort_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Scanner:
    def __init__(self, ctx, elem_dtype, sum_dtype, algo="blelloch", options=None):
        self.L = lib()
        self.ctx = ctx
        self.elem = np.dtype(elem_dtype)
        self.sum = np.dtype(sum_dtype)
        self.ptr = _call(self.L.clo_scan_new, _b(algo), _b(options), ctx.ptr,
                         clo_type(elem_dtype), clo_type(sum_dtype), None)

    def scan_device(self, q, bin_, bout, n):
        return _call(self.L.clo_scan_with_device_data, self.ptr, q.ptr, None, bin_.ptr, bout.ptr,
                     n, 0)

    def scan_segmented_device(self, q, bin_, bout, offsets, nseg, n):
        return _call(self.L.clo_scan_segmented_with_device_data, self.ptr, q.ptr, None, bin_.ptr,
                     bout.ptr, offsets.ptr, nseg, n)

    def scan_host(self, a, q=None, q_comm=None):
        a = np.ascontiguousarray(a, dtype=self.elem)
        out = np.empty(a.size, dtype=self.sum)
        _call(self.L.clo_scan_with_host_data, self.ptr, q.ptr if q else None,
              q_comm.ptr if q_comm else None,
              a.ctypes.data_as(ctypes.c_void_p), out.ctypes.data_as(ctypes.c_void_p), a.size, 0)
        return out

    def close(self):
        if self.ptr:
            self.L.clo_scan_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Rng:
    def __init__(self, ctx, name, seed_type, count, main_seed=0, hash=None, seeds=None, q=None):
        self.L = lib()
        self.ctx = ctx
        self.name = name
        self.count = count
        sp = None
        if seed_type == SEED_EXT_HOST:
            self._keep = np.ascontiguousarray(seeds)
            sp = self._keep.ctypes.data_as(ctypes.c_void_p)
        elif seed_type == SEED_EXT_DEV:
            sp = seeds.ptr
        self.ptr = _call(self.L.clo_rng_new, _b(name), seed_type, sp, count, main_seed, _b(hash),
                         ctx.ptr, q.ptr if q else None)

    def source(self):
        return self.L.clo_rng_get_source(self.ptr).decode()

    def size(self):
        return self.L.clo_rng_get_size(self.ptr)

    def seeds_ptr(self):
        return self.L.clo_rng_get_device_seeds(self.ptr)

    def read_states(self, q):
        buf = self.L.clo_rng_get_device_seeds(self.ptr)
        n = self.size()
        out = np.empty(n, dtype=np.uint8)
        _call(self.L.ccl_buffer_enqueue_read, buf, q.ptr, 1, 0, n,
              out.ctypes.data_as(ctypes.c_void_p), None)
        return out

    def fill(self, q, out_buf, draws, bits=32, maxint=0):
        return _call(self.L.clo_rng_fill, self.ptr, q.ptr, out_buf.ptr, draws, bits, maxint)

    def close(self):
        if self.ptr:
            self.L.clo_rng_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def generate(q, buf, n, kind, seed, param=0, start=0):
    return _call(lib().clo_util_generate, q.ptr, buf.ptr, n, kind, seed, param, start)


def multiset_hash(q, buf, n, width):
    return _call(lib().clo_util_multiset_hash, q.ptr, buf.ptr, n, width)


def check_pairs(q, keys_in, keys_out, vals_out, n, width):
    """Violations of: vals_out is a permutation of 0..n-1 and
    keys_out[i] == keys_in[vals_out[i]] (computed on the device)."""
    return _call(lib().clo_util_check_pairs, q.ptr, keys_in.ptr, keys_out.ptr, vals_out.ptr, n,
                 width)


def check_sorted(q, keys, n, width, vals=None):
    return _call(lib().clo_util_check_sorted, q.ptr, keys.ptr, vals.ptr if vals else None, n,
                 width)
