"""ctypes binding of liboka.so (declared in include/oka.h).

The library is the product's only compute path: every class and function in
this package that evaluates the model, fits, traces or rasterizes calls it.  If
the library is missing, or no GPU is visible, calls raise ``OkaLibraryError``;
nothing falls back to a CPU implementation.
"""
import ctypes as C
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboka.so")

OK = 0
ERR_ARG, ERR_RANGE, ERR_AQUIFER, ERR_LINALG, ERR_DEVICE, ERR_NOMEM, ERR_CAPACITY = (
    -1, -2, -3, -4, -5, -6, -7)

TRACE_OK, TRACE_TRUNCATED, TRACE_STALLED, TRACE_STEP_CAP = 0, 1, 2, 3
FP64_EXACT, FP32 = 0, 1
(EVAL_POTENTIAL, EVAL_POTENTIAL_WELLS, EVAL_DISCHARGE, EVAL_HEAD, EVAL_VELOCITY,
 EVAL_VELOCITY_CONFINED) = range(6)
PTR_DEVICE = 0x1
ASYNC = 0x2
RASTER_FP32 = 0x4      # oka_rasterize: the FP32 tracers' rasterizer
ANTITHETIC = 0x8       # draws / surrogate sums in antithetic pairs (2k, 2k + 1)
QMC = 0x10             # draws / surrogate sums from scrambled Sobol' points

_d = C.c_double
_i32 = C.c_int32
_i64 = C.c_int64
_u32 = C.c_uint32
_vp = C.c_void_p

# Every symbol include/oka.h declares (the CPU test checks the export table).
EXPORTED = (
    "oka_abi_version", "oka_ctx_device_info", "oka_ctx_create", "oka_ctx_destroy", "oka_last_error",
    "oka_ctx_set_stream", "oka_ctx_stream", "oka_ctx_synchronize", "oka_model_eval",
    "oka_fit_draw", "oka_trace_chronicle", "oka_grid_history", "oka_trace_vertices",
    "oka_rasterize", "oka_pow_fifth", "oka_log", "oka_trace_marks", "oka_cv_node_sums",
    "oka_cv_node_trace", "oka_cv_node_reduce", "oka_node_set_destroy", "oka_mlp_create",
    "oka_mlp_destroy", "oka_mlp_param_count", "oka_mlp_set_params", "oka_mlp_get_params",
    "oka_mlp_grad", "oka_mlp_train", "oka_mlp_predict", "oka_mlp_accumulate",
    "oka_draw_realizations", "oka_cv_surrogate_sums", "oka_cv_surrogate_sums_qmc",
    "oka_mark_moments", "oka_batch_create",
    "oka_batch_tune", "oka_batch_run", "oka_batch_results", "oka_batch_destroy",
    "oka_count_histogram", "oka_trace_chronicle_noalloc", "oka_field_smooth",
    "oka_field_sorted_probabilities",
)

DIST_CONST, DIST_UNIFORM, DIST_TRIANGULAR, DIST_LOGNORMAL = 0, 1, 2, 3


class Dist_(C.Structure):
    _fields_ = [("kind", _i32), ("reserved", _i32), ("a", _d), ("b", _d), ("c", _d),
                ("p_negative", _d)]


class OkaLibraryError(RuntimeError):
    """liboka.so is missing, failed to load, or no GPU is available."""


class OkaError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"liboka error {code}: {msg}")
        self.code = code


class Model_(C.Structure):
    _fields_ = [("base", _d), ("conductivity", _d), ("porosity", _d), ("thickness", _d),
                ("xo", _d), ("yo", _d), ("coef", _d * 6), ("nwells", _i32), ("reserved", _i32),
                ("wells", _vp)]


class Site_(C.Structure):
    _fields_ = [("nwells", _i32), ("nobs", _i32), ("wells_xy", _vp), ("obs", _vp), ("xo", _d),
                ("yo", _d), ("base", _d)]


class TraceOpts_(C.Structure):
    _fields_ = [("npaths", _i32), ("precision", _i32), ("confined", _i32), ("schedule", _i32),
                ("duration", _d), ("tol", _d), ("maxstep", _d), ("umbra", _d),
                ("max_steps", _i64), ("start_xy", _vp), ("bitmap_mode", _i32),
                ("tile_slots", _i32), ("pool_slots", _i32), ("pool_waves_per_cu", _i32),
                ("pool_chunk", _i32), ("far_order", _i32)]


class DeviceInfo_(C.Structure):
    _fields_ = [("num_cu", _i32), ("warp_size", _i32), ("lds_per_workgroup", _i64),
                ("lds_per_cu", _i64)]


class Grid_(C.Structure):
    _fields_ = [("xmin", _d), ("ymin", _d), ("xmax", _d), ("ymax", _d), ("dx", _d), ("dy", _d),
                ("nrows", _i32), ("ncols", _i32)]


class BatchSite_(C.Structure):
    _fields_ = [("site", C.POINTER(Site_)), ("nreal", _i32), ("K", _vp), ("n", _vp), ("b", _vp),
                ("q", _vp), ("z", _vp), ("opts", TraceOpts_), ("grid", Grid_)]


class Epoch_(C.Structure):
    _fields_ = [("first_trace", _i64), ("xmin", _d), ("ymin", _d), ("nrows", _i32),
                ("ncols", _i32), ("row_off", _i32), ("col_off", _i32)]


EPOCH_DTYPE = np.dtype([("first_trace", np.int64), ("xmin", np.float64), ("ymin", np.float64),
                        ("nrows", np.int32), ("ncols", np.int32), ("row_off", np.int32),
                        ("col_off", np.int32)])
assert EPOCH_DTYPE.itemsize == C.sizeof(Epoch_)


class Window_(C.Structure):
    _fields_ = [("row0", _i32), ("col0", _i32), ("cell", _i32), ("nrows", _i32), ("ncols", _i32)]


class TraceStats_(C.Structure):
    _fields_ = [("attempted", _i64), ("accepted", _i64), ("status_count", _i64 * 4),
                ("escaped", _i64), ("visited", _i64), ("candidates", _i64),
                ("kernel_ms", _d)]


_lib = None
_lock = threading.Lock()
_contexts = {}


def load():
    """Load liboka.so and declare its prototypes (no GPU needed)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OkaLibraryError(
            f"{LIB_PATH} is not built; run __graft_entry__.build() (or make -C onekapy_amd/csrc)")
    try:
        L = C.CDLL(LIB_PATH)
    except OSError as e:
        raise OkaLibraryError(f"cannot load {LIB_PATH}: {e}") from e
    L.oka_abi_version.restype = C.c_int
    L.oka_abi_version.argtypes = []
    L.oka_ctx_device_info.restype = C.c_int
    L.oka_ctx_device_info.argtypes = [_vp, C.POINTER(DeviceInfo_)]
    L.oka_ctx_create.restype = C.c_int
    L.oka_ctx_create.argtypes = [C.POINTER(_vp), C.c_int]
    L.oka_ctx_destroy.restype = None
    L.oka_ctx_destroy.argtypes = [_vp]
    L.oka_last_error.restype = C.c_char_p
    L.oka_last_error.argtypes = [_vp]
    L.oka_ctx_set_stream.restype = C.c_int
    L.oka_ctx_set_stream.argtypes = [_vp, _vp]
    L.oka_ctx_stream.restype = _vp
    L.oka_ctx_stream.argtypes = [_vp]
    L.oka_ctx_synchronize.restype = C.c_int
    L.oka_ctx_synchronize.argtypes = [_vp]
    L.oka_model_eval.restype = C.c_int
    L.oka_model_eval.argtypes = [_vp, C.POINTER(Model_), _i32, _i64, _vp, _vp, _vp, _u32]
    L.oka_fit_draw.restype = C.c_int
    L.oka_fit_draw.argtypes = [_vp, C.POINTER(Site_), _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                               _vp, _u32]
    L.oka_trace_chronicle.restype = C.c_int
    L.oka_trace_chronicle.argtypes = [_vp, C.POINTER(Site_), _i32, _vp, _vp, _vp, _vp, _vp,
                                      C.POINTER(TraceOpts_), C.POINTER(Grid_), _i32, _vp, _i32,
                                      _vp, _vp, _vp, _vp, C.POINTER(TraceStats_), _u32]
    L.oka_trace_chronicle_noalloc.restype = C.c_int
    L.oka_trace_chronicle_noalloc.argtypes = L.oka_trace_chronicle.argtypes
    L.oka_grid_history.restype = C.c_int
    L.oka_grid_history.argtypes = [C.POINTER(Grid_), _i64, _vp, C.POINTER(Grid_), _vp, _i64,
                                   C.POINTER(_i64)]
    L.oka_trace_vertices.restype = C.c_int
    L.oka_trace_vertices.argtypes = [_vp, C.POINTER(Site_), _i32, _vp, _vp, _vp, _vp, _vp,
                                     C.POINTER(TraceOpts_), _i64, _vp, _vp, _vp, _vp, _u32]
    L.oka_rasterize.restype = C.c_int
    L.oka_rasterize.argtypes = [_vp, C.POINTER(Grid_), _i64, _vp, _d, _vp, _u32]
    L.oka_trace_marks.restype = C.c_int
    L.oka_trace_marks.argtypes = [_vp, C.POINTER(Site_), _i32, _vp, _vp, _vp, _vp, _vp,
                                  C.POINTER(TraceOpts_), C.POINTER(Grid_), C.POINTER(Window_),
                                  _vp, _vp, _vp, C.POINTER(TraceStats_), _u32]
    L.oka_cv_node_trace.restype = C.c_int
    L.oka_cv_node_trace.argtypes = [_vp, C.POINTER(Site_), _i32, _vp, _vp, _vp, _vp, _vp,
                                    C.POINTER(TraceOpts_), C.POINTER(Grid_), C.POINTER(Window_),
                                    _vp, C.POINTER(_vp), C.POINTER(TraceStats_), _u32]
    L.oka_cv_node_reduce.restype = C.c_int
    L.oka_cv_node_reduce.argtypes = [_vp, _vp, _vp, _i32, _vp, _vp, _vp, _vp, _u32]
    L.oka_node_set_destroy.restype = None
    L.oka_node_set_destroy.argtypes = [_vp]
    L.oka_cv_node_sums.restype = C.c_int
    L.oka_cv_node_sums.argtypes = [_vp, _vp, C.POINTER(Site_), _i32, _vp, _vp, _vp, _vp, _vp,
                                   _vp, C.POINTER(TraceOpts_), C.POINTER(Grid_),
                                   C.POINTER(Window_), _i32, _vp, _vp, _vp, _vp, _vp,
                                   C.POINTER(TraceStats_), _u32]
    L.oka_mlp_create.restype = C.c_int
    L.oka_mlp_create.argtypes = [_vp, _i32, _i32, _i32, _i32, C.POINTER(_vp)]
    L.oka_mlp_destroy.restype = None
    L.oka_mlp_destroy.argtypes = [_vp]
    L.oka_mlp_param_count.restype = _i64
    L.oka_mlp_param_count.argtypes = [_vp]
    L.oka_mlp_set_params.restype = C.c_int
    L.oka_mlp_set_params.argtypes = [_vp, _vp, _u32]
    L.oka_mlp_get_params.restype = C.c_int
    L.oka_mlp_get_params.argtypes = [_vp, _vp, _u32]
    L.oka_mlp_grad.restype = C.c_int
    L.oka_mlp_grad.argtypes = [_vp, _i32, _vp, _vp, C.c_float, _vp, C.POINTER(C.c_float), _u32]
    L.oka_mlp_train.restype = C.c_int
    L.oka_mlp_train.argtypes = [_vp, _i64, _vp, _vp, C.c_float, _vp, _i32, _i32, C.c_float,
                                C.c_float, C.c_float, C.c_float, _vp, _u32]
    L.oka_mlp_predict.restype = C.c_int
    L.oka_mlp_predict.argtypes = [_vp, _i64, _vp, _vp, _u32]
    L.oka_mlp_accumulate.restype = C.c_int
    L.oka_mlp_accumulate.argtypes = [_vp, _i64, _vp, _vp, C.c_float, _vp, _vp, _vp, _vp, _u32]
    L.oka_draw_realizations.restype = C.c_int
    L.oka_draw_realizations.argtypes = [_vp, _i32, _vp, _i64, C.c_uint64, _i64, _vp, _vp, _vp,
                                        _vp, _vp, _u32]
    L.oka_cv_surrogate_sums.restype = C.c_int
    L.oka_cv_surrogate_sums.argtypes = [_vp, _vp, C.POINTER(Site_), _vp, _i64, C.c_uint64, _i64,
                                        _vp, _vp, _vp, _vp, _u32]
    L.oka_cv_surrogate_sums_qmc.restype = C.c_int
    L.oka_cv_surrogate_sums_qmc.argtypes = [_vp, _vp, C.POINTER(Site_), _vp, _i64, _i32, _vp,
                                            _vp, _vp, _vp, _u32]
    L.oka_batch_create.restype = C.c_int
    L.oka_batch_create.argtypes = [_vp, _i32, _vp, C.POINTER(TraceOpts_), C.POINTER(_vp)]
    L.oka_batch_tune.restype = C.c_int
    L.oka_batch_tune.argtypes = [_vp, C.POINTER(TraceOpts_)]
    L.oka_batch_run.restype = C.c_int
    L.oka_batch_run.argtypes = [_vp, C.POINTER(C.c_float)]
    L.oka_batch_results.restype = C.c_int
    L.oka_batch_results.argtypes = [_vp, _i32, _vp, _vp, C.POINTER(TraceStats_)]
    L.oka_batch_destroy.restype = None
    L.oka_batch_destroy.argtypes = [_vp]
    L.oka_count_histogram.restype = C.c_int
    L.oka_count_histogram.argtypes = [_vp, _i64, _vp, _i32, _vp, _u32]
    L.oka_field_smooth.restype = C.c_int
    L.oka_field_smooth.argtypes = [_vp, _i32, _i32, _vp, _d, _i32, _vp, _vp, _u32]
    L.oka_field_sorted_probabilities.restype = C.c_int
    L.oka_field_sorted_probabilities.argtypes = [_vp, _i64, _vp, _d, _vp, _u32]
    L.oka_mark_moments.restype = C.c_int
    L.oka_mark_moments.argtypes = [_vp, _i64, _i32, _vp, _vp, _vp, _u32]
    L.oka_pow_fifth.restype = _d
    L.oka_pow_fifth.argtypes = [_d]
    L.oka_log.restype = _d
    L.oka_log.argtypes = [_d]
    _lib = L
    return L


class Context:
    """One liboka context (device, stream, staging buffers) per device."""

    def __init__(self, device=0):
        L = load()
        h = _vp()
        rc = L.oka_ctx_create(C.byref(h), int(device))
        if rc != OK:
            raise OkaLibraryError(
                f"oka_ctx_create(device={device}) failed with {rc}: no usable GPU for liboka")
        self.h = h
        self.device = device
        self.lib = L

    def check(self, rc):
        if rc != OK:
            msg = self.lib.oka_last_error(self.h)
            raise OkaError(rc, msg.decode() if msg else "")
        return rc

    def set_stream(self, stream_ptr):
        self.check(self.lib.oka_ctx_set_stream(self.h, stream_ptr))

    def device_info(self):
        """dict(num_cu, warp_size, lds_per_workgroup, lds_per_cu) of the context's device."""
        d = DeviceInfo_()
        self.check(self.lib.oka_ctx_device_info(self.h, C.byref(d)))
        return {k: getattr(d, k) for k, _ in DeviceInfo_._fields_}

    def synchronize(self):
        self.check(self.lib.oka_ctx_synchronize(self.h))

    def __del__(self):
        try:
            if self.h:
                self.lib.oka_ctx_destroy(self.h)
        except Exception:
            pass


def context(device=0):
    """The process-wide context for `device`, created on first use."""
    with _lock:
        ctx = _contexts.get(device)
        if ctx is None:
            ctx = Context(device)
            _contexts[device] = ctx
        return ctx


def f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        a = a.reshape(shape)
    return a


def ptr(a):
    """Address of a numpy array's data (None for None)."""
    if a is None:
        return None
    return a.ctypes.data


def make_grid(xmin, ymin, xmax, ymax, dx, dy, nrows, ncols):
    g = Grid_()
    g.xmin, g.ymin, g.xmax, g.ymax = float(xmin), float(ymin), float(xmax), float(ymax)
    g.dx, g.dy = float(dx), float(dy)
    g.nrows, g.ncols = int(nrows), int(ncols)
    return g


def make_site(wells_xy, xo, yo, base=0.0, obs=None, wells_ptr=None, obs_ptr=None):
    """Site_ from host arrays (kept alive by the caller) or raw device pointers."""
    s = Site_()
    s.nwells = int(len(wells_xy)) if wells_xy is not None else 0
    s.nobs = int(len(obs)) if obs is not None else 0
    s.wells_xy = wells_ptr if wells_ptr is not None else ptr(wells_xy)
    s.obs = obs_ptr if obs_ptr is not None else ptr(obs)
    s.xo, s.yo, s.base = float(xo), float(yo), float(base)
    return s


BITMAP_AUTO, BITMAP_LDS_FLAT, BITMAP_TILED, BITMAP_HBM_FLAT = 0, 1, 2, 3
SCHED_AUTO, SCHED_WORKGROUP, SCHED_POOL = 0, 1, 2


def make_trace_opts(npaths, duration, tol, maxstep, umbra, start_xy_ptr, precision=FP64_EXACT,
                    confined=True, max_steps=0, bitmap_mode=BITMAP_AUTO, tile_slots=0,
                    schedule=SCHED_AUTO, pool_slots=0, pool_waves_per_cu=0, pool_chunk=0,
                    far_order=0):
    o = TraceOpts_()
    o.npaths = int(npaths)
    o.precision = int(precision)
    o.confined = 1 if confined else 0
    o.duration, o.tol, o.maxstep, o.umbra = float(duration), float(tol), float(maxstep), float(umbra)
    o.max_steps = int(max_steps)
    o.start_xy = start_xy_ptr
    o.bitmap_mode = int(bitmap_mode)
    o.tile_slots = int(tile_slots)
    o.schedule = int(schedule)
    o.pool_slots = int(pool_slots)
    o.pool_waves_per_cu = int(pool_waves_per_cu)
    o.pool_chunk = int(pool_chunk)
    o.far_order = int(far_order)
    return o


def grid_history(init_grid, bbox):
    """oka_grid_history: (final Grid_, epochs structured array)."""
    L = load()
    bbox = f64(bbox).reshape(-1, 4)
    final = Grid_()
    n = _i64(0)
    rc = L.oka_grid_history(C.byref(init_grid), bbox.shape[0], ptr(bbox), C.byref(final), None, 0,
                            C.byref(n))
    if rc != OK:
        raise OkaError(rc, "oka_grid_history failed")
    epochs = np.zeros(n.value, dtype=EPOCH_DTYPE)
    rc = L.oka_grid_history(C.byref(init_grid), bbox.shape[0], ptr(bbox), C.byref(final),
                            epochs.ctypes.data, n.value, C.byref(n))
    if rc != OK:
        raise OkaError(rc, "oka_grid_history failed")
    return final, epochs
