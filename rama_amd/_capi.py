"""ctypes binding of librama_hip.so (include/rama_hip.h).

The product path: every call runs on the GPU through the C ABI.  If the
library is missing or cannot load, import fails loudly — there is no CPU
fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librama_hip.so")

I32P = ctypes.POINTER(ctypes.c_int32)
F32P = ctypes.POINTER(ctypes.c_float)
I64P = ctypes.POINTER(ctypes.c_int64)
F64P = ctypes.POINTER(ctypes.c_double)
VP = ctypes.c_void_p
I32 = ctypes.c_int32
I64 = ctypes.c_int64
F32 = ctypes.c_float
F64 = ctypes.c_double


class Options(ctypes.Structure):
    """rama_options_c: field-for-field multicut_solver_options (include/multicut_solver_options.h:4-19)."""

    _fields_ = [
        ("max_cycle_length_lb", ctypes.c_int32),
        ("num_dual_itr_lb", ctypes.c_int32),
        ("max_cycle_length_primal", ctypes.c_int32),
        ("num_dual_itr_primal", ctypes.c_int32),
        ("num_outer_itr_dual", ctypes.c_int32),
        ("mean_multiplier_mm", ctypes.c_float),
        ("matching_thresh_crossover_ratio", ctypes.c_float),
        ("tri_memory_factor", ctypes.c_float),
        ("only_compute_lb", ctypes.c_int32),
        ("max_time_sec", ctypes.c_int32),
        ("dump_timeline", ctypes.c_int32),
        ("verbose", ctypes.c_int32),
        ("sanitize_graph", ctypes.c_int32),
    ]


OPTP = ctypes.POINTER(Options)
TLP = ctypes.c_void_p

# Every symbol declared in include/rama_hip.h with its argument types.
SIGNATURES = {
    "rama_last_error": ([], ctypes.c_char_p),
    "rama_version": ([], ctypes.c_char_p),
    "rama_options_default": ([OPTP], None),
    "rama_options_preset": ([ctypes.c_char_p, OPTP], I32),
    "rama_solve_host": ([I32P, I32P, F32P, I64, OPTP, I32P, I64, I64P, F64P, I32P, ctypes.POINTER(TLP)], I32),
    "rama_solve_device": ([VP, VP, VP, I64, OPTP, I32, VP, VP, I64, I64P, F64P, I32P, ctypes.POINTER(TLP)], I32),
    "rama_solve_device_noalloc": ([VP, VP, VP, I64, OPTP, I32, VP, VP, I64, I64P, F64P, I32P, I64P], I32),
    "rama_timeline_count": ([TLP], I64),
    "rama_timeline_length": ([TLP, I64], I64),
    "rama_timeline_labels": ([TLP, I64], I32P),
    "rama_timeline_free": ([TLP], None),
    "rama_objective": ([I32P, I64, I32P, I32P, F32P, I64], F64),
    "rama_stage_dcoo_init": ([I32P, I32P, F32P, I64, I32P, I32P, F32P, I32P], I32),
    "rama_stage_contract": ([I32P, I32P, F32P, I64, I32P, I64, I32P, I32P, F32P, I64P, I32P, I32P, F64P], I32),
    "rama_stage_contract_symmetric": ([I32P, I32P, F32P, I64, I32P, I64, I32P, I32P, I64, I32P, F32P, I64, I64P], I32),
    "rama_stage_compute_offsets": ([I32P, I64, I32, I32P], I32),
    "rama_stage_symmetric_csr": ([I32P, I32P, F32P, I64, I32, I32, F32, I32P, I32P, F32P, I64, I64P], I32),
    "rama_stage_compress_labels": ([I32P, I64, I32, I32P], I32),
    "rama_stage_matching": ([I32P, I32P, F32P, I64, I32, I32, F32, I32P, I64P, I32P], I32),
    "rama_stage_conflicted_cycles": ([I32P, I32P, F32P, I64, I32, I32, I32, F32, F32, I32P, I32P, I32P, I64, I64P,
                                      I64P], I32),
    "rama_stage_message_passing": ([I32P, I32P, F32P, I64, I32P, I32P, I32P, I64, I32, I32, I32P, I32P, F32P, I64,
                                    I64P, F64P], I32),
    "rama_stage_dual_solver": ([I32P, I32P, F32P, I64, I32, I32, I32, I32, F32, I32, F32, I32P, I32P, F32P, I64, I64P,
                                I32P, F64P], I32),
    "rama_stage_connected_components": ([I32P, I32P, I32, I64, I32P], I32),
    "rama_stage_max_spanning_forest": ([I32P, I32P, F32P, I64, I32P, I32P, F32P, I64P], I32),
    "rama_stage_woc_mapping": ([I32P, I32P, F32P, I64, I32, I32, I32P, I64P, I32P], I32),
    "rama_solve_batch_host": ([I32, I64P, I32P, I32P, F32P, OPTP, I32, I32P, I64P, I64P, F64P, I32P], I32),
    "rama_server_create": ([I32, I32, I64, OPTP], VP),
    "rama_server_submit": ([VP, I32P, I32P, F32P, I64], I64),
    "rama_server_poll": ([VP, I64], I32),
    "rama_server_wait": ([VP, I64, I32P, I64, I64P, F64P, F64P, F64P], I32),
    "rama_tuning_load": ([ctypes.c_char_p], I32),
    "rama_tuning_set": ([ctypes.c_char_p, I64, I64], I32),
    "rama_tuning_get": ([ctypes.c_char_p, I64], I64),
    "rama_tuning_source": ([], ctypes.c_char_p),
    "rama_tune_time_sort": ([I64, I32, I32, I32, I32, I32, F64P], I32),
    "rama_stream_copy_gbps": ([I64, I32, F64P], I32),
    "rama_server_destroy": ([VP], None),
    "rama_profile_enable": ([I32], None),
    "rama_profile_only": ([ctypes.c_char_p], None),
    "rama_release_cached": ([I32], None),
    "rama_profile_reset": ([], None),
    "rama_profile_count": ([], I64),
    "rama_profile_entry": ([I64, ctypes.POINTER(ctypes.c_char_p), I64P, F64P, F64P], I32),
    "rama_profile_work": ([I64, F64P, F64P], I32),
}

_lib = None


class RamaError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(msg)
        self.status = status


def lib():
    """Load librama_hip.so (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("librama_hip.so is not built; run `python -m rama_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise RamaError(rc, lib().rama_last_error().decode())


def _p(a, t):
    return a.ctypes.data_as(t)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


_PRESET_NAMES = ("P", "PD", "PD+", "D")


def options(preset="PD", **kw):
    o = Options()
    if preset is None:
        lib().rama_options_default(ctypes.byref(o))
    else:
        lib().rama_options_preset(preset.encode(), ctypes.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def _timeline(tl):
    out = []
    if not tl:
        return out
    L = lib()
    try:
        for k in range(L.rama_timeline_count(tl)):
            n = L.rama_timeline_length(tl, k)
            p = L.rama_timeline_labels(tl, k)
            out.append(np.ctypeslib.as_array(p, shape=(n,)).copy() if n else np.zeros(0, np.int32))
    finally:
        L.rama_timeline_free(tl)
    return out


def solve(i, j, c, opts=None, want_timeline=False):
    """rama_cuda(i, j, costs, opts) on host arrays -> (labels, lb, time_ms, timeline)."""
    opts = opts or options()
    i, j, c = _i32(i), _i32(j), _f32(c)
    cap = int(max(i.max(initial=0), j.max(initial=0))) + 2
    lab = np.zeros(cap, np.int32)
    n = ctypes.c_int64()
    lb = ctypes.c_double()
    ms = ctypes.c_int32()
    tl = TLP()
    _check(lib().rama_solve_host(_p(i, I32P), _p(j, I32P), _p(c, F32P), len(c), ctypes.byref(opts), _p(lab, I32P), cap,
                                 ctypes.byref(n), ctypes.byref(lb), ctypes.byref(ms),
                                 ctypes.byref(tl) if want_timeline else None))
    return lab[: n.value].copy(), lb.value, ms.value, _timeline(tl)


def solve_device(i_ptr, j_ptr, c_ptr, num_edges, labels_ptr, labels_cap, opts=None, device=0, stream=None,
                 want_timeline=False):
    """Device-pointer solve (rama_solve_device). Returns (num_labels, lb, time_ms, timeline)."""
    opts = opts or options()
    n = ctypes.c_int64()
    lb = ctypes.c_double()
    ms = ctypes.c_int32()
    tl = TLP()
    _check(lib().rama_solve_device(VP(i_ptr), VP(j_ptr), VP(c_ptr), num_edges, ctypes.byref(opts), device,
                                   VP(stream or 0), VP(labels_ptr), labels_cap, ctypes.byref(n), ctypes.byref(lb),
                                   ctypes.byref(ms), ctypes.byref(tl) if want_timeline else None))
    return n.value, lb.value, ms.value, _timeline(tl)


def solve_device_noalloc(i_ptr, j_ptr, c_ptr, num_edges, labels_ptr, labels_cap, opts=None, device=0, stream=None):
    """Allocation-free device-pointer solve (rama_solve_device_noalloc).
    Returns (num_labels, lb, time_ms, scratch_peak_bytes)."""
    opts = opts or options()
    n = ctypes.c_int64()
    lb = ctypes.c_double()
    ms = ctypes.c_int32()
    peak = ctypes.c_int64()
    _check(lib().rama_solve_device_noalloc(VP(i_ptr), VP(j_ptr), VP(c_ptr), num_edges, ctypes.byref(opts), device,
                                           VP(stream or 0), VP(labels_ptr), labels_cap, ctypes.byref(n),
                                           ctypes.byref(lb), ctypes.byref(ms), ctypes.byref(peak)))
    return n.value, lb.value, ms.value, peak.value


NOALLOC_MAX_EDGES = 1000000  # RAMA_NOALLOC_MAX_EDGES (include/rama_hip.h)


def objective(labels, i, j, c):
    labels, i, j, c = _i32(labels), _i32(i), _i32(j), _f32(c)
    return lib().rama_objective(_p(labels, I32P), len(labels), _p(i, I32P), _p(j, I32P), _p(c, F32P), len(c))


# ---------------------------------------------------------------- stages

def dcoo_init(i, j, c):
    i, j, c = _i32(i), _i32(j), _f32(c)
    m = len(c)
    r, co, v = np.zeros(m, np.int32), np.zeros(m, np.int32), np.zeros(m, np.float32)
    dims = np.zeros(2, np.int32)
    _check(lib().rama_stage_dcoo_init(_p(i, I32P), _p(j, I32P), _p(c, F32P), m, _p(r, I32P), _p(co, I32P),
                                      _p(v, F32P), _p(dims, I32P)))
    return r, co, v, (int(dims[0]), int(dims[1]))


def contract(row, col, val, node_map):
    row, col, val, node_map = _i32(row), _i32(col), _f32(val), _i32(node_map)
    m = len(val)
    r, co, v = np.zeros(m, np.int32), np.zeros(m, np.int32), np.zeros(m, np.float32)
    mo = ctypes.c_int64()
    dims = np.zeros(2, np.int32)
    bad = ctypes.c_int32()
    ds = ctypes.c_double()
    _check(lib().rama_stage_contract(_p(row, I32P), _p(col, I32P), _p(val, F32P), m, _p(node_map, I32P), len(node_map),
                                     _p(r, I32P), _p(co, I32P), _p(v, F32P), ctypes.byref(mo), _p(dims, I32P),
                                     ctypes.byref(bad), ctypes.byref(ds)))
    k = mo.value
    return r[:k].copy(), co[:k].copy(), v[:k].copy(), (int(dims[0]), int(dims[1])), bool(bad.value), ds.value


def contract_symmetric(row, col, val, node_map):
    """Symmetric CSR (offsets, neighbours, costs) of the contracted graph,
    max(rows, cols) rows, as the next cycle search receives it."""
    row, col, val, node_map = _i32(row), _i32(col), _f32(val), _i32(node_map)
    m = len(val)
    off_cap = (int(node_map.max()) if len(node_map) else 0) + 2
    off = np.zeros(off_cap, np.int32)
    cap = 2 * m + 1
    adj = np.zeros(cap, np.int32)
    w = np.zeros(cap, np.float32)
    n = ctypes.c_int32()
    nnz = ctypes.c_int64()
    _check(lib().rama_stage_contract_symmetric(_p(row, I32P), _p(col, I32P), _p(val, F32P), m, _p(node_map, I32P),
                                               len(node_map), ctypes.byref(n), _p(off, I32P), off_cap, _p(adj, I32P),
                                               _p(w, F32P), cap, ctypes.byref(nnz)))
    k = nnz.value
    return off[: n.value + 1].copy(), adj[:k].copy(), w[:k].copy()


def compute_offsets(ids, max_value):
    ids = _i32(ids)
    out = np.zeros(max_value + 2, np.int32)
    _check(lib().rama_stage_compute_offsets(_p(ids, I32P), len(ids), max_value, _p(out, I32P)))
    return out


def symmetric_csr(row, col, val, n_nodes, mode=0, keep_above=0.0):
    """Symmetric CSR (offsets, neighbours, costs) of a (row, col)-sorted
    edge list with row < col; mode 1 keeps the edges with cost > keep_above,
    mode 2 finds the row offsets one thread per row."""
    row, col, val = _i32(row), _i32(col), _f32(val)
    off = np.zeros(n_nodes + 1, np.int32)
    cap = 2 * len(val) + 1
    adj = np.zeros(cap, np.int32)
    w = np.zeros(cap, np.float32)
    nnz = ctypes.c_int64()
    _check(lib().rama_stage_symmetric_csr(_p(row, I32P), _p(col, I32P), _p(val, F32P), len(val), n_nodes, mode,
                                          keep_above, _p(off, I32P), _p(adj, I32P), _p(w, F32P), cap,
                                          ctypes.byref(nnz)))
    k = nnz.value
    return off, adj[:k].copy(), w[:k].copy()


def compress_labels(data, max_label):
    data = _i32(data)
    out = np.zeros(len(data), np.int32)
    _check(lib().rama_stage_compress_labels(_p(data, I32P), len(data), max_label, _p(out, I32P)))
    return out


def matching(row, col, val, rows, cols, mm):
    row, col, val = _i32(row), _i32(col), _f32(val)
    out = np.zeros(rows + cols + 1, np.int32)
    n = ctypes.c_int64()
    matched = ctypes.c_int32()
    _check(lib().rama_stage_matching(_p(row, I32P), _p(col, I32P), _p(val, F32P), len(val), rows, cols, mm,
                                     _p(out, I32P), ctypes.byref(n), ctypes.byref(matched)))
    return out[: n.value].copy(), matched.value


def conflicted_cycles(row, col, val, rows, cols, max_len, factor=2.0, tol=1e-4):
    row, col, val = _i32(row), _i32(col), _f32(val)
    cap = int(factor * len(val)) + 16
    t = [np.zeros(cap, np.int32) for _ in range(3)]
    n = ctypes.c_int64()
    raw = ctypes.c_int64()
    _check(lib().rama_stage_conflicted_cycles(_p(row, I32P), _p(col, I32P), _p(val, F32P), len(val), rows, cols,
                                              max_len, factor, tol, _p(t[0], I32P), _p(t[1], I32P), _p(t[2], I32P),
                                              cap, ctypes.byref(n), ctypes.byref(raw)))
    k = n.value
    return t[0][:k].copy(), t[1][:k].copy(), t[2][:k].copy(), raw.value


def message_passing(row, col, val, t1, t2, t3, pre=0, iters=0):
    row, col, val = _i32(row), _i32(col), _f32(val)
    t1, t2, t3 = _i32(t1), _i32(t2), _i32(t3)
    cap = len(val) + 3 * len(t1) + 1
    oi, oj, oc = np.zeros(cap, np.int32), np.zeros(cap, np.int32), np.zeros(cap, np.float32)
    om = ctypes.c_int64()
    lbs = np.zeros(iters + 1, np.float64)
    _check(lib().rama_stage_message_passing(_p(row, I32P), _p(col, I32P), _p(val, F32P), len(val), _p(t1, I32P),
                                            _p(t2, I32P), _p(t3, I32P), len(t1), pre, iters, _p(oi, I32P),
                                            _p(oj, I32P), _p(oc, F32P), cap, ctypes.byref(om), _p(lbs, F64P)))
    k = om.value
    return oi[:k].copy(), oj[:k].copy(), oc[:k].copy(), lbs


def dual_solver(row, col, val, rows, cols, max_len, num_iter, factor=2.0, num_outer=1, tol=1e-4):
    row, col, val = _i32(row), _i32(col), _f32(val)
    cap = len(val) * (1 + int(3 * factor) * 3 * max(1, num_outer)) + 16
    r, c, v = np.zeros(cap, np.int32), np.zeros(cap, np.int32), np.zeros(cap, np.float32)
    om = ctypes.c_int64()
    dims = np.zeros(2, np.int32)
    lb = ctypes.c_double()
    _check(lib().rama_stage_dual_solver(_p(row, I32P), _p(col, I32P), _p(val, F32P), len(val), rows, cols, max_len,
                                        num_iter, factor, num_outer, tol, _p(r, I32P), _p(c, I32P), _p(v, F32P), cap,
                                        ctypes.byref(om), _p(dims, I32P), ctypes.byref(lb)))
    k = om.value
    return r[:k].copy(), c[:k].copy(), v[:k].copy(), (int(dims[0]), int(dims[1])), lb.value


def connected_components(offsets, cols):
    offsets, cols = _i32(offsets), _i32(cols)
    n = len(offsets) - 1
    out = np.zeros(max(n, 1), np.int32)
    _check(lib().rama_stage_connected_components(_p(offsets, I32P), _p(cols, I32P), n, len(cols), _p(out, I32P)))
    return out[:n]


def max_spanning_forest(row, col, val):
    row, col, val = _i32(row), _i32(col), _f32(val)
    m = len(val)
    r, c, v = np.zeros(m, np.int32), np.zeros(m, np.int32), np.zeros(m, np.float32)
    om = ctypes.c_int64()
    _check(lib().rama_stage_max_spanning_forest(_p(row, I32P), _p(col, I32P), _p(val, F32P), m, _p(r, I32P),
                                                _p(c, I32P), _p(v, F32P), ctypes.byref(om)))
    k = om.value
    return r[:k].copy(), c[:k].copy(), v[:k].copy()


def woc_mapping(row, col, val, rows, cols):
    row, col, val = _i32(row), _i32(col), _f32(val)
    out = np.zeros(max(rows, cols) + 1, np.int32)
    n = ctypes.c_int64()
    cnt = ctypes.c_int32()
    _check(lib().rama_stage_woc_mapping(_p(row, I32P), _p(col, I32P), _p(val, F32P), len(val), rows, cols,
                                        _p(out, I32P), ctypes.byref(n), ctypes.byref(cnt)))
    return out[: n.value].copy(), cnt.value


# ------------------------------------------------------- batch and serving

def solve_batch(instances, opts=None, workers=8):
    """Solve a list of (i, j, c) instances concurrently on one GPU.
    Returns a list of (labels, lb)."""
    opts = opts or options()
    k = len(instances)
    edge_off = np.zeros(k + 1, np.int64)
    label_off = np.zeros(k + 1, np.int64)
    for q, (i, j, c) in enumerate(instances):
        edge_off[q + 1] = edge_off[q] + len(c)
        label_off[q + 1] = label_off[q] + int(max(np.max(i), np.max(j))) + 2
    I = _i32(np.concatenate([np.asarray(x[0]) for x in instances])) if k else np.zeros(0, np.int32)
    J = _i32(np.concatenate([np.asarray(x[1]) for x in instances])) if k else np.zeros(0, np.int32)
    C = _f32(np.concatenate([np.asarray(x[2]) for x in instances])) if k else np.zeros(0, np.float32)
    labels = np.zeros(max(1, int(label_off[-1])), np.int32)
    n = np.zeros(max(1, k), np.int64)
    lbs = np.zeros(max(1, k), np.float64)
    st = np.zeros(max(1, k), np.int32)
    _check(lib().rama_solve_batch_host(k, _p(edge_off, I64P), _p(I, I32P), _p(J, I32P), _p(C, F32P),
                                       ctypes.byref(opts), workers, _p(labels, I32P), _p(label_off, I64P),
                                       _p(n, I64P), _p(lbs, F64P), _p(st, I32P)))
    return [(labels[label_off[q]: label_off[q] + n[q]].copy(), float(lbs[q])) for q in range(k)]


class Server:
    """Serving loop (rama_server_*): submit instances as they arrive, wait for
    results; admission keeps the in-flight device-memory estimate within the
    HBM budget.  mode "lockstep" (default where the options allow) packs the
    admitted requests and admits new ones between rounds; "workers" solves
    one request per worker thread (RAMA_SERVE_MODE)."""

    def __init__(self, workers=8, hbm_budget_bytes=200 << 30, opts=None, device=0, mode=None):
        self._opts = opts or options(verbose=0)
        if mode not in (None, "lockstep", "workers"):
            raise ValueError("mode must be 'lockstep' or 'workers'")
        old = os.environ.get("RAMA_SERVE_MODE")
        try:
            if mode is not None:
                os.environ["RAMA_SERVE_MODE"] = mode
            self._h = lib().rama_server_create(device, workers, hbm_budget_bytes, ctypes.byref(self._opts))
        finally:
            if mode is not None:
                if old is None:
                    os.environ.pop("RAMA_SERVE_MODE", None)
                else:
                    os.environ["RAMA_SERVE_MODE"] = old
        if not self._h:
            raise RamaError(RAMA_ERR_INTERNAL_STATUS, lib().rama_last_error().decode())
        self._caps = {}

    def submit(self, i, j, c):
        i, j, c = _i32(i), _i32(j), _f32(c)
        t = lib().rama_server_submit(self._h, _p(i, I32P), _p(j, I32P), _p(c, F32P), len(c))
        if t < 0:
            raise RamaError(1, lib().rama_last_error().decode())
        self._caps[t] = int(max(i.max(initial=0), j.max(initial=0))) + 2
        return t

    def done(self, ticket):
        return lib().rama_server_poll(self._h, ticket) == 1

    def wait(self, ticket):
        """-> (labels, lb, latency_ms, solve_ms)"""
        cap = self._caps.pop(ticket)
        lab = np.zeros(cap, np.int32)
        n = ctypes.c_int64()
        lb = ctypes.c_double()
        lat = ctypes.c_double()
        sms = ctypes.c_double()
        _check(lib().rama_server_wait(self._h, ticket, _p(lab, I32P), cap, ctypes.byref(n), ctypes.byref(lb),
                                      ctypes.byref(lat), ctypes.byref(sms)))
        return lab[: n.value].copy(), lb.value, lat.value, sms.value

    def close(self):
        if self._h:
            lib().rama_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


RAMA_ERR_INTERNAL_STATUS = 6
