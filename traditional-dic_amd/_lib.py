"""ctypes binding of libdic_hip.so (the C-ABI declared in include/dic_hip.h).

The shared library is built in-tree by ``make -C traditional-dic_amd``.  There
is no fallback: if the library or a GPU is missing, every entry point raises.
"""
from __future__ import annotations

import ctypes
import os
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdic_hip.so")
SYNTH_PATH = os.path.join(_HERE, "libdic_synth.so")

# ---- status codes (include/dic_hip.h) ----
DIC_OK = 0
DIC_ERR_INVALID_ARGUMENT = 1
DIC_ERR_HIP = 2
DIC_ERR_UNSUPPORTED = 3
DIC_ERR_BUDGET = 4
DIC_ERR_NO_DEVICE = 5

# SolverStatus order (include/dic/core/types.hpp:35-40 of the reference)
SUCCESS, NOT_CONVERGED, INVALID_INPUT, NUMERICAL_FAILURE = 0, 1, 2, 3

PRECISION_FP64, PRECISION_MIXED = 0, 1
OBJECTIVE_ZNSSD, OBJECTIVE_SSD = 0, 1
OPTIMIZER_ICGN, OPTIMIZER_FGN = 0, 1
GRADIENT_CENTRAL, GRADIENT_BSPLINE_HALF = 0, 1

EXPORTED_SYMBOLS = (
    "dic_version", "dic_device_count", "dic_ctx_create", "dic_ctx_destroy", "dic_last_error",
    "dic_ctx_set_stream", "dic_ctx_stream", "dic_ctx_synchronize", "dic_ctx_bytes_in_use",
    "dic_image_upload", "dic_image_wrap_device", "dic_image_free", "dic_ref_gradient",
    "dic_grad_download", "dic_grad_free", "dic_bspline_prefilter", "dic_coeffs_info",
    "dic_coeffs_download", "dic_coeffs_free", "dic_bspline_sample", "dic_icgn_solve_batch",
    "dic_icgn_solve_batch_device", "dic_buffer_alloc", "dic_buffer_free", "dic_buffer_upload",
    "dic_buffer_download", "dic_buffer_copy", "dic_event_record", "dic_event_elapsed_ms", "dic_event_destroy",
    "dic_bspline_prefilter_into", "dic_ref_gradient_into",
    "dic_batch_create", "dic_batch_set_tuning", "dic_batch_get_tuning", "dic_batch_info", "dic_batch_run",
    "dic_batch_autotune", "dic_batch_destroy",
    "dic_host_alloc", "dic_host_free", "dic_server_create", "dic_server_submit", "dic_server_wait",
    "dic_server_frames_done", "dic_server_stats_get", "dic_server_now", "dic_server_error", "dic_server_destroy",
    "dic_propagate", "dic_integer_search", "dic_seed_candidates", "dic_select_seed", "dic_image_mirror_pad",
    "dic_image_download", "dic_icgn_solve_batch_device_noalloc",
)
NOALLOC_MAX_RADIUS = 102  # DIC_NOALLOC_MAX_RADIUS

# Displacement2D + iterations (dic_result); InitialDisplacement (dic_initial)
RESULT_DTYPE = np.dtype([
    ("x", "<f8"), ("y", "<f8"), ("p", "<f8", (12,)), ("correlation", "<f8"),
    ("status", "<i4"), ("valid", "<i4"), ("iterations", "<i4"), ("reserved", "<i4"),
])
INITIAL_DTYPE = np.dtype([
    ("u", "<f8"), ("v", "<f8"), ("du_dx", "<f8"), ("du_dy", "<f8"),
    ("dv_dx", "<f8"), ("dv_dy", "<f8"), ("confidence", "<f8"),
])
assert RESULT_DTYPE.itemsize == 136 and INITIAL_DTYPE.itemsize == 56


class IcgnParams(ctypes.Structure):
    _fields_ = [
        ("subset_radius", ctypes.c_int32),
        ("max_iterations", ctypes.c_int32),
        ("convergence_threshold", ctypes.c_double),
        ("shape_order", ctypes.c_int32),
        ("objective", ctypes.c_int32),
        ("optimizer", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("masked", ctypes.c_int32),
        ("min_valid_samples", ctypes.c_int32),
        ("min_valid_sample_ratio", ctypes.c_double),
    ]


BATCH_GRID, BATCH_PERSISTENT, BATCH_TILE = 0, 1, 2
BATCH_ORDER_NATURAL, BATCH_ORDER_COST, BATCH_ORDER_INTERLEAVE, BATCH_ORDER_XCD = 0, 1, 2, 3


class BatchItem(ctypes.Structure):
    """dic_batch_item: one (pair, parameter set) of a batched launch (device arrays)."""
    _fields_ = [
        ("reference", ctypes.c_void_p),
        ("gradient", ctypes.c_void_p),
        ("deformed", ctypes.c_void_p),
        ("poi_xy", ctypes.c_void_p),
        ("init", ctypes.c_void_p),
        ("out", ctypes.c_void_p),
        ("n", ctypes.c_int32),
        ("subset_radius", ctypes.c_int32),
        ("shape_order", ctypes.c_int32),
        ("max_iterations", ctypes.c_int32),
        ("convergence_threshold", ctypes.c_double),
        ("optimizer", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class BatchTuning(ctypes.Structure):
    _fields_ = [
        ("mapping", ctypes.c_int32),
        ("waves_per_cu", ctypes.c_int32),
        ("chunk", ctypes.c_int32),
        ("order", ctypes.c_int32),
    ]

    def as_dict(self):
        return {"mapping": ("grid", "persistent", "tile")[self.mapping], "waves_per_cu": self.waves_per_cu,
                "chunk": self.chunk, "order": ("natural", "cost", "interleave", "xcd")[self.order]}


class ServerConfig(ctypes.Structure):
    _fields_ = [
        ("hbm_budget", ctypes.c_uint64),
        ("max_active", ctypes.c_int32),
        ("max_width", ctypes.c_int32),
        ("max_height", ctypes.c_int32),
        ("subset_radius", ctypes.c_int32),
        ("shape_order", ctypes.c_int32),
        ("max_iterations", ctypes.c_int32),
        ("convergence_threshold", ctypes.c_double),
        ("degree", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("objective", ctypes.c_int32),
        ("optimizer", ctypes.c_int32),
        ("guess_lag", ctypes.c_int32),
        ("tick_pois", ctypes.c_int32),
        ("guess_extrapolate", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class Request(ctypes.Structure):
    _fields_ = [
        ("reference", ctypes.c_void_p),
        ("width", ctypes.c_int32),
        ("height", ctypes.c_int32),
        ("n_frames", ctypes.c_int32),
        ("n", ctypes.c_int32),
        ("frames", ctypes.c_void_p),
        ("poi_xy", ctypes.c_void_p),
        ("init", ctypes.c_void_p),
        ("results", ctypes.c_void_p),
        ("frame_done_s", ctypes.c_void_p),
    ]


class ServerStats(ctypes.Structure):
    _fields_ = [
        ("submitted", ctypes.c_int64), ("completed", ctypes.c_int64), ("frames_done", ctypes.c_int64),
        ("pois_done", ctypes.c_int64), ("ticks", ctypes.c_int64),
        ("active", ctypes.c_int32), ("pending", ctypes.c_int32), ("max_active_seen", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("bytes_held", ctypes.c_uint64), ("peak_bytes_held", ctypes.c_uint64),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


PROPAGATION_ORDERS = {"heap": 0, "frontier": 1}  # DIC_PROPAGATE_HEAP, DIC_PROPAGATE_FRONTIER


class PropagationParams(ctypes.Structure):
    _fields_ = [("spacing", ctypes.c_int32), ("order", ctypes.c_int32), ("max_znssd", ctypes.c_double)]


SEED_DTYPE = np.dtype([("x", "<f8"), ("y", "<f8"), ("displacement", INITIAL_DTYPE)])

SEARCH_GLOBAL, SEARCH_AROUND_PRIOR = 0, 1
QUALITY_ZNCC, QUALITY_ZNSSD, QUALITY_SSD = 0, 1, 2

# InitialDisplacement with every field (dic_estimate)
ESTIMATE_DTYPE = np.dtype([
    ("u", "<f8"), ("v", "<f8"), ("du_dx", "<f8"), ("du_dy", "<f8"), ("dv_dx", "<f8"), ("dv_dy", "<f8"),
    ("confidence", "<f8"), ("valid", "<i4"), ("reserved", "<i4"), ("zncc", "<f8"), ("znssd", "<f8"),
])
# SeedEvaluation (dic_seed_evaluation)
SEED_EVAL_DTYPE = np.dtype([
    ("x", "<f8"), ("y", "<f8"), ("displacement", RESULT_DTYPE), ("quality", "<f8"),
    ("displacement_norm", "<f8"), ("quality_passed", "<i4"), ("valid", "<i4"),
])


class IntegerSearchParams(ctypes.Structure):
    _fields_ = [("subset_radius", ctypes.c_int32), ("search_radius", ctypes.c_int32),
                ("pyramid_enabled", ctypes.c_int32), ("pyramid_scale", ctypes.c_int32),
                ("pyramid_refinement_radius", ctypes.c_int32), ("mode", ctypes.c_int32),
                ("masked", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class SeedParams(ctypes.Structure):
    _fields_ = [("search", IntegerSearchParams), ("subpixel_enabled", ctypes.c_int32),
                ("subpixel_shape_order", ctypes.c_int32), ("subpixel_optimizer", ctypes.c_int32),
                ("subpixel_objective", ctypes.c_int32), ("subpixel_radius", ctypes.c_int32),
                ("subpixel_max_iterations", ctypes.c_int32), ("subpixel_convergence_threshold", ctypes.c_double),
                ("seed_count", ctypes.c_int32), ("quality_metric", ctypes.c_int32), ("min_zncc", ctypes.c_double),
                ("max_znssd", ctypes.c_double), ("max_ssd", ctypes.c_double),
                ("min_displacement_norm", ctypes.c_double), ("min_texture_std", ctypes.c_double),
                ("kmeans_iterations", ctypes.c_int32), ("kmeans_sample_limit", ctypes.c_int32),
                ("truncate_roi_subsets", ctypes.c_int32), ("min_valid_samples", ctypes.c_int32),
                ("min_valid_sample_ratio", ctypes.c_double), ("degree", ctypes.c_int32), ("border", ctypes.c_int32)]


class DicError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"[dic_hip error {code}] {message}")
        self.code = code


_vp = ctypes.c_void_p
_i32 = ctypes.c_int
_dp = ctypes.POINTER(ctypes.c_double)

_SIGNATURES = {
    "dic_version": ([], ctypes.c_char_p),
    "dic_device_count": ([ctypes.POINTER(ctypes.c_int)], _i32),
    "dic_ctx_create": ([_i32, ctypes.c_uint64, ctypes.POINTER(_vp)], _i32),
    "dic_ctx_destroy": ([_vp], None),
    "dic_last_error": ([_vp], ctypes.c_char_p),
    "dic_ctx_set_stream": ([_vp, _vp], _i32),
    "dic_ctx_stream": ([_vp], _vp),
    "dic_ctx_synchronize": ([_vp], _i32),
    "dic_ctx_bytes_in_use": ([_vp], ctypes.c_uint64),
    "dic_image_upload": ([_vp, _vp, _i32, _i32, ctypes.POINTER(_vp)], _i32),
    "dic_image_wrap_device": ([_vp, _vp, _i32, _i32, ctypes.POINTER(_vp)], _i32),
    "dic_image_free": ([_vp], None),
    "dic_image_download": ([_vp, _vp, _vp], _i32),
    "dic_ref_gradient": ([_vp, _vp, _i32, _i32, _i32, _i32, ctypes.POINTER(_vp)], _i32),
    "dic_grad_download": ([_vp, _vp, _vp, _vp], _i32),
    "dic_grad_free": ([_vp], None),
    "dic_bspline_prefilter": ([_vp, _vp, _i32, _i32, _i32, _i32, ctypes.POINTER(_vp)], _i32),
    "dic_bspline_prefilter_into": ([_vp, _vp, _vp], _i32),
    "dic_ref_gradient_into": ([_vp, _vp, _vp], _i32),
    "dic_coeffs_info": ([_vp] + [ctypes.POINTER(ctypes.c_int)] * 5, _i32),
    "dic_coeffs_download": ([_vp, _vp, _vp], _i32),
    "dic_coeffs_free": ([_vp], None),
    "dic_bspline_sample": ([_vp, _vp, _vp, _i32, _vp, _vp], _i32),
    "dic_icgn_solve_batch": ([_vp, ctypes.POINTER(IcgnParams), _vp, _vp, _vp, _vp, _vp, _vp, _i32, _vp], _i32),
    "dic_icgn_solve_batch_device": (
        [_vp, ctypes.POINTER(IcgnParams), _vp, _vp, _vp, _vp, _vp, _vp, _i32, _vp, _vp], _i32),
    "dic_icgn_solve_batch_device_noalloc": (
        [_vp, ctypes.POINTER(IcgnParams), _vp, _vp, _vp, _vp, _vp, _vp, _i32, _vp, _vp], _i32),
    "dic_buffer_alloc": ([_vp, ctypes.c_size_t, ctypes.POINTER(_vp)], _i32),
    "dic_buffer_free": ([_vp, _vp], None),
    "dic_buffer_upload": ([_vp, _vp, _vp, ctypes.c_size_t], _i32),
    "dic_buffer_download": ([_vp, _vp, _vp, ctypes.c_size_t], _i32),
    "dic_buffer_copy": ([_vp, _vp, _vp, ctypes.c_size_t], _i32),
    "dic_event_record": ([_vp, ctypes.POINTER(_vp)], _i32),
    "dic_event_elapsed_ms": ([_vp, _vp, ctypes.POINTER(ctypes.c_float)], _i32),
    "dic_event_destroy": ([_vp], None),
    "dic_batch_create": ([_vp, ctypes.POINTER(BatchItem), _i32, _i32, ctypes.POINTER(_vp)], _i32),
    "dic_batch_set_tuning": ([_vp, ctypes.POINTER(BatchTuning)], _i32),
    "dic_batch_get_tuning": ([_vp, ctypes.POINTER(BatchTuning)], _i32),
    "dic_batch_info": ([_vp, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)], _i32),
    "dic_batch_run": ([_vp, _vp], _i32),
    "dic_batch_autotune": ([_vp, ctypes.POINTER(BatchTuning), _i32, _i32, ctypes.POINTER(BatchTuning),
                            ctypes.POINTER(ctypes.c_float)], _i32),
    "dic_batch_destroy": ([_vp], None),
    "dic_host_alloc": ([_vp, ctypes.c_size_t, ctypes.POINTER(_vp)], _i32),
    "dic_host_free": ([_vp], None),
    "dic_server_create": ([_vp, ctypes.POINTER(ServerConfig), ctypes.POINTER(_vp)], _i32),
    "dic_server_submit": ([_vp, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_int64),
                           ctypes.POINTER(ctypes.c_double)], _i32),
    "dic_server_wait": ([_vp, ctypes.c_int64, ctypes.c_double], _i32),
    "dic_server_frames_done": ([_vp, ctypes.c_int64, ctypes.POINTER(ctypes.c_int)], _i32),
    "dic_server_stats_get": ([_vp, ctypes.POINTER(ServerStats)], _i32),
    "dic_server_now": ([_vp], ctypes.c_double),
    "dic_server_error": ([_vp], ctypes.c_char_p),
    "dic_server_destroy": ([_vp], None),
    "dic_propagate": ([_vp, ctypes.POINTER(IcgnParams), ctypes.POINTER(PropagationParams), _vp, _vp, _vp, _vp, _vp,
                       _i32, _vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                       ctypes.POINTER(ctypes.c_int)], _i32),
    "dic_integer_search": ([_vp, ctypes.POINTER(IntegerSearchParams), _vp, _vp, _vp, _vp, _vp, _i32, _vp], _i32),
    "dic_seed_candidates": ([_vp, ctypes.POINTER(SeedParams), _vp, _vp, _vp, _i32, ctypes.POINTER(ctypes.c_int)],
                            _i32),
    "dic_select_seed": ([_vp, ctypes.POINTER(SeedParams), _vp, _vp, _vp, _vp, _i32, ctypes.POINTER(ctypes.c_int),
                         ctypes.POINTER(ctypes.c_int)], _i32),
    "dic_image_mirror_pad": ([_vp, _vp, _i32, ctypes.POINTER(_vp)], _i32),
}

_lib = None


def load_library(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load libdic_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(
            f"{path} is missing: build it with `make -C traditional-dic_amd` "
            "(or __graft_entry__.build()); there is no CPU fallback")
    lib = ctypes.CDLL(path)
    for name, (argtypes, restype) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = argtypes
        fn.restype = restype
    _lib = lib
    return lib


def device_count() -> int:
    lib = load_library()
    n = ctypes.c_int(0)
    rc = lib.dic_device_count(ctypes.byref(n))
    return n.value if rc == DIC_OK else 0


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def _checked_roi(roi: np.ndarray, img: "DeviceImage") -> np.ndarray:
    """The ROI as contiguous bytes of the image's shape (H, W): the C-ABI reads
    W * H bytes from the pointer, so any other shape is refused here."""
    r = np.asarray(roi)
    if r.shape != (img.height, img.width):
        raise ValueError(f"roi has shape {r.shape}, the image is {img.height} x {img.width} (H x W)")
    return np.ascontiguousarray(r, np.uint8)


class Context:
    """Owns a dic_ctx (device, stream, HBM budget)."""

    def __init__(self, device: int = 0, hbm_budget_bytes: int = 0):
        self.lib = load_library()
        # everything created on this context (handles, buffers, batches,
        # servers): released before the context itself, whatever order the
        # garbage collector finds them in
        self._dependents = weakref.WeakSet()
        h = _vp()
        rc = self.lib.dic_ctx_create(int(device), int(hbm_budget_bytes), ctypes.byref(h))
        if rc != DIC_OK:
            raise DicError(rc, "dic_ctx_create failed (no HIP device?)")
        self.handle = h

    def _own(self, obj):
        self._dependents.add(obj)
        return obj

    def check(self, rc: int):
        if rc != DIC_OK:
            msg = self.lib.dic_last_error(self.handle)
            raise DicError(rc, msg.decode() if msg else "")

    def close(self):
        if getattr(self, "handle", None):
            # servers and batches first (their work runs on this context), then
            # the device handles and buffers, then the context
            deps = list(getattr(self, "_dependents", ()))
            for kind in (Server, Batch, _Handle, DeviceBuffer):
                for d in deps:
                    if isinstance(d, kind):
                        d._release()
            self.lib.dic_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_stream(self, stream_handle: int | None):
        self.check(self.lib.dic_ctx_set_stream(self.handle, stream_handle or None))

    def stream(self) -> int:
        return self.lib.dic_ctx_stream(self.handle) or 0

    def synchronize(self):
        self.check(self.lib.dic_ctx_synchronize(self.handle))

    def bytes_in_use(self) -> int:
        return int(self.lib.dic_ctx_bytes_in_use(self.handle))

    # ---- device buffers and events ----
    def alloc(self, nbytes: int) -> "DeviceBuffer":
        p = _vp()
        self.check(self.lib.dic_buffer_alloc(self.handle, int(nbytes), ctypes.byref(p)))
        return DeviceBuffer(self, p.value, int(nbytes))

    def to_device(self, arr: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(arr)
        buf = self.alloc(a.nbytes)
        self.check(self.lib.dic_buffer_upload(self.handle, buf.ptr, _ptr(a), a.nbytes))
        return buf

    def copy(self, dst: "DeviceBuffer", src: "DeviceBuffer", nbytes: int | None = None):
        n = min(dst.nbytes, src.nbytes) if nbytes is None else int(nbytes)
        self.check(self.lib.dic_buffer_copy(self.handle, dst.ptr, src.ptr, n))

    def record_event(self) -> int:
        e = _vp()
        self.check(self.lib.dic_event_record(self.handle, ctypes.byref(e)))
        return e.value

    def elapsed_ms(self, start: int, end: int) -> float:
        ms = ctypes.c_float()
        rc = self.lib.dic_event_elapsed_ms(start, end, ctypes.byref(ms))
        if rc != DIC_OK:
            raise DicError(rc, "event timing failed")
        return float(ms.value)

    def destroy_event(self, e: int):
        self.lib.dic_event_destroy(e)

    # ---- handles ----
    def upload_image(self, img: np.ndarray) -> "DeviceImage":
        a = np.ascontiguousarray(img, dtype=np.float32)
        if a.ndim != 2:
            raise ValueError("Image must be a 2D numpy array.")
        h = _vp()
        self.check(self.lib.dic_image_upload(self.handle, _ptr(a), a.shape[1], a.shape[0], ctypes.byref(h)))
        return DeviceImage(self, h, a.shape[1], a.shape[0])

    def wrap_image(self, device_ptr: int, width: int, height: int) -> "DeviceImage":
        h = _vp()
        self.check(self.lib.dic_image_wrap_device(self.handle, device_ptr, width, height, ctypes.byref(h)))
        return DeviceImage(self, h, width, height)

    def prefilter(self, img: "DeviceImage", degree: int = 5, border: int = 3, exact: bool = True,
                  precision: int = PRECISION_FP64) -> "Coefficients":
        h = _vp()
        self.check(self.lib.dic_bspline_prefilter(self.handle, img.handle, degree, border, int(exact), precision,
                                                  ctypes.byref(h)))
        return Coefficients(self, h)

    def prefilter_into(self, img: "DeviceImage", coeffs: "Coefficients"):
        self.check(self.lib.dic_bspline_prefilter_into(self.handle, img.handle, coeffs.handle))

    def ref_gradient_into(self, img: "DeviceImage", grad: "Gradient"):
        self.check(self.lib.dic_ref_gradient_into(self.handle, img.handle, grad.handle))

    def ref_gradient(self, img: "DeviceImage", mode: int = GRADIENT_CENTRAL, degree: int = 5, border: int = 3,
                     precision: int = PRECISION_FP64) -> "Gradient":
        h = _vp()
        self.check(self.lib.dic_ref_gradient(self.handle, img.handle, mode, degree, border, precision,
                                             ctypes.byref(h)))
        return Gradient(self, h, img.width, img.height)

    def sample(self, coeffs: "Coefficients", xy: np.ndarray):
        xy = np.ascontiguousarray(xy, dtype=np.float64).reshape(-1, 2)
        n = xy.shape[0]
        val = np.empty(n, np.float64)
        grad = np.empty((n, 2), np.float64)
        self.check(self.lib.dic_bspline_sample(self.handle, coeffs.handle, _ptr(xy), n, _ptr(val), _ptr(grad)))
        return val, grad

    def solve(self, params: IcgnParams, ref: "DeviceImage", grad: "Gradient", coeffs: "Coefficients",
              poi_xy: np.ndarray, init: np.ndarray, roi: np.ndarray | None = None) -> np.ndarray:
        xy = np.ascontiguousarray(poi_xy, dtype=np.float64).reshape(-1, 2)
        ini = np.ascontiguousarray(init, dtype=INITIAL_DTYPE)
        n = xy.shape[0]
        if ini.shape[0] != n:
            raise ValueError("one initial guess per POI is required")
        out = np.zeros(n, RESULT_DTYPE)
        roi_p = None
        if roi is not None:
            r = _checked_roi(roi, ref)
            roi_p = _ptr(r)
        self.check(self.lib.dic_icgn_solve_batch(self.handle, ctypes.byref(params), ref.handle, grad.handle,
                                                 coeffs.handle, roi_p, _ptr(xy), _ptr(ini), n, _ptr(out)))
        return out

    def propagate(self, params: IcgnParams, spacing: int, max_znssd: float, ref: "DeviceImage", grad,
                  coeffs: "Coefficients", roi: np.ndarray, seeds: np.ndarray, order: str = "heap"):
        """dic_propagate: returns (grid results (gh, gw), points_computed).
        order: "heap" (the reference's priority-queue order) or "frontier"."""
        if order not in PROPAGATION_ORDERS:
            raise ValueError(f"propagation order must be one of {sorted(PROPAGATION_ORDERS)}, not {order!r}")
        r = _checked_roi(roi, ref)
        sd = np.ascontiguousarray(seeds, SEED_DTYPE)
        step = max(int(spacing), 0) + 1
        W, H = ref.width, ref.height
        gw, gh = (W + step - 1) // step, (H + step - 1) // step
        out = np.zeros(gw * gh, RESULT_DTYPE)
        pp = PropagationParams(int(spacing), PROPAGATION_ORDERS[order], float(max_znssd))
        a, b, n = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self.check(self.lib.dic_propagate(self.handle, ctypes.byref(params), ctypes.byref(pp), ref.handle,
                                          grad.handle if grad is not None else None, coeffs.handle, _ptr(r),
                                          _ptr(sd), sd.shape[0], _ptr(out), ctypes.byref(a), ctypes.byref(b),
                                          ctypes.byref(n)))
        return out.reshape(gh, gw), n.value

    def integer_search(self, params: IntegerSearchParams, ref: "DeviceImage", deformed: "DeviceImage",
                       points_xy: np.ndarray, prior_uv: np.ndarray | None = None,
                       roi: np.ndarray | None = None) -> np.ndarray:
        """dic_integer_search over host points; ESTIMATE_DTYPE records."""
        xy = np.ascontiguousarray(points_xy, np.float64).reshape(-1, 2)
        pr = np.ascontiguousarray(prior_uv, np.float64).reshape(-1, 2) if prior_uv is not None else None
        r = _checked_roi(roi, ref) if roi is not None else None
        out = np.zeros(xy.shape[0], ESTIMATE_DTYPE)
        self.check(self.lib.dic_integer_search(self.handle, ctypes.byref(params), ref.handle, deformed.handle,
                                               _ptr(r) if r is not None else None, _ptr(xy),
                                               _ptr(pr) if pr is not None else None, xy.shape[0], _ptr(out)))
        return out

    def seed_candidates(self, params: SeedParams, ref: "DeviceImage", roi: np.ndarray) -> np.ndarray:
        r = _checked_roi(roi, ref)
        out = np.zeros((max(1, params.seed_count), 2))
        n = ctypes.c_int(0)
        self.check(self.lib.dic_seed_candidates(self.handle, ctypes.byref(params), ref.handle, _ptr(r), _ptr(out),
                                                out.shape[0], ctypes.byref(n)))
        return out[:n.value]

    def select_seed(self, params: SeedParams, ref: "DeviceImage", deformed: "DeviceImage", roi: np.ndarray):
        """dic_select_seed: (SEED_EVAL_DTYPE evaluations, index of the best seed or -1)."""
        r = _checked_roi(roi, ref)
        out = np.zeros(max(1, params.seed_count), SEED_EVAL_DTYPE)
        n, best = ctypes.c_int(0), ctypes.c_int(-1)
        self.check(self.lib.dic_select_seed(self.handle, ctypes.byref(params), ref.handle, deformed.handle, _ptr(r),
                                            _ptr(out), out.shape[0], ctypes.byref(n), ctypes.byref(best)))
        return out[:n.value], best.value

    def download_image(self, img: "DeviceImage") -> np.ndarray:
        out = np.empty((img.height, img.width), np.float32)
        self.check(self.lib.dic_image_download(self.handle, img.handle, _ptr(out)))
        return out

    def mirror_pad(self, img: "DeviceImage", pad: int) -> "DeviceImage":
        h = _vp()
        self.check(self.lib.dic_image_mirror_pad(self.handle, img.handle, int(pad), ctypes.byref(h)))
        return DeviceImage(self, h, img.width + 2 * max(0, pad), img.height + 2 * max(0, pad))

    def solve_device(self, params: IcgnParams, ref: "DeviceImage", grad: "Gradient", coeffs: "Coefficients",
                     xy_ptr: int, init_ptr: int, n: int, out_ptr: int, roi_ptr: int | None = None,
                     stream: int | None = None):
        self.check(self.lib.dic_icgn_solve_batch_device(self.handle, ctypes.byref(params), ref.handle, grad.handle,
                                                        coeffs.handle, roi_ptr, xy_ptr, init_ptr, int(n), out_ptr,
                                                        stream))


    def solve_device_noalloc(self, params: IcgnParams, ref: "DeviceImage", grad: "Gradient", coeffs: "Coefficients",
                             xy_ptr: int, init_ptr: int, n: int, out_ptr: int, roi_ptr: int | None = None,
                             stream: int | None = None):
        """dic_icgn_solve_batch_device_noalloc: solve_device without any device
        allocation (radius <= NOALLOC_MAX_RADIUS)."""
        self.check(self.lib.dic_icgn_solve_batch_device_noalloc(
            self.handle, ctypes.byref(params), ref.handle, grad.handle if grad is not None else None, coeffs.handle,
            roi_ptr, xy_ptr, init_ptr, int(n), out_ptr, stream))

    def batch(self, items, objective: int = OBJECTIVE_ZNSSD) -> "Batch":
        """items: sequence of dicts with keys reference, gradient, deformed (handles),
        poi_xy, init, out (DeviceBuffer), n, subset_radius, shape_order,
        max_iterations, convergence_threshold, optimizer (OPTIMIZER_ICGN by default)."""
        arr = (BatchItem * max(1, len(items)))()
        keep = []
        for k, it in enumerate(items):
            arr[k].reference = it["reference"].handle
            arr[k].gradient = it["gradient"].handle if it.get("gradient") is not None else None
            arr[k].deformed = it["deformed"].handle
            arr[k].poi_xy = it["poi_xy"].ptr
            arr[k].init = it["init"].ptr
            arr[k].out = it["out"].ptr
            arr[k].n = int(it["n"])
            arr[k].subset_radius = int(it["subset_radius"])
            arr[k].shape_order = int(it["shape_order"])
            arr[k].max_iterations = int(it.get("max_iterations", 30))
            arr[k].convergence_threshold = float(it.get("convergence_threshold", 1e-3))
            arr[k].optimizer = int(it.get("optimizer", OPTIMIZER_ICGN))
            keep.append(it)
        h = _vp()
        self.check(self.lib.dic_batch_create(self.handle, arr, len(items), int(objective), ctypes.byref(h)))
        return Batch(self, h, keep)


    def host_array(self, shape, dtype=np.float32) -> np.ndarray:
        """A numpy array in pinned host memory (dic_host_alloc), freed with the array."""
        dt = np.dtype(dtype)
        nbytes = int(np.prod(shape)) * dt.itemsize
        p = _vp()
        self.check(self.lib.dic_host_alloc(self.handle, max(1, nbytes), ctypes.byref(p)))
        buf = (ctypes.c_char * max(1, nbytes)).from_address(p.value)
        arr = np.frombuffer(buf, dtype=np.uint8, count=nbytes).view(dt).reshape(shape)
        _PINNED[id(buf)] = p.value
        import weakref
        weakref.finalize(buf, _free_pinned, self.lib, id(buf))
        return arr

    def server(self, radius=15, shape_order=1, degree=5, precision=PRECISION_MIXED, objective=OBJECTIVE_ZNSSD,
               max_active=64, max_width=4096, max_height=4096, hbm_budget=0, max_iterations=30,
               convergence_threshold=1e-3, optimizer=OPTIMIZER_ICGN, guess_lag=1, tick_pois=0,
               guess_extrapolate=False) -> "Server":
        """guess_lag K: frame f starts from frame f - K's results (frames below
        K from the request's guesses), and up to K frames of a request share a
        tick while the tick holds fewer than tick_pois POIs (0: the library's
        default).  guess_extrapolate: frames f >= 2K start from 2 r(f - K) -
        r(f - 2K) where both are Success."""
        cfg = ServerConfig(int(hbm_budget), int(max_active), int(max_width), int(max_height), int(radius),
                           int(shape_order), int(max_iterations), float(convergence_threshold), int(degree),
                           int(precision), int(objective), int(optimizer), int(guess_lag), int(tick_pois),
                           int(bool(guess_extrapolate)), 0)
        h = _vp()
        self.check(self.lib.dic_server_create(self.handle, ctypes.byref(cfg), ctypes.byref(h)))
        return Server(self, h)


_PINNED = {}


def _free_pinned(lib, key):
    p = _PINNED.pop(key, None)
    if p:
        lib.dic_host_free(p)


class ServedRequest:
    """A submitted request: results[f, i] fill in as frames finish."""

    def __init__(self, server, rid, arrival, results, done_s, keep):
        self.server, self.id, self.arrival, self.results, self.frame_done_s = server, rid, arrival, results, done_s
        self._keep = keep

    def wait(self, timeout_s: float = -1.0):
        rc = self.server.ctx.lib.dic_server_wait(self.server.handle, self.id, float(timeout_s))
        if rc != DIC_OK:
            raise DicError(rc, self.server.error())
        return self.results

    def frames_done(self) -> int:
        f = ctypes.c_int()
        self.server.ctx.check(self.server.ctx.lib.dic_server_frames_done(self.server.handle, self.id, ctypes.byref(f)))
        return f.value


class Server:
    """dic_server: continuous-batching request loop on one device."""

    def __init__(self, ctx, handle):
        self.ctx, self.handle = ctx, handle
        ctx._own(self)

    def submit(self, reference: np.ndarray, frames, poi_xy: np.ndarray, init: np.ndarray) -> ServedRequest:
        ref = np.ascontiguousarray(reference, np.float32)
        fr = [np.ascontiguousarray(f, np.float32) for f in frames]
        for f in fr:
            if f.shape != ref.shape:
                raise ValueError("every frame must have the reference's shape")
        xy = np.ascontiguousarray(poi_xy, np.float64).reshape(-1, 2)
        ini = np.ascontiguousarray(init, INITIAL_DTYPE)
        n = xy.shape[0]
        if ini.shape[0] != n:
            raise ValueError("one initial guess per POI is required")
        ptrs = (ctypes.c_void_p * len(fr))(*[f.ctypes.data for f in fr])
        results = np.zeros((len(fr), n), RESULT_DTYPE)
        done = np.full(len(fr), np.nan)
        req = Request(ref.ctypes.data, ref.shape[1], ref.shape[0], len(fr), n, ctypes.cast(ptrs, ctypes.c_void_p),
                      xy.ctypes.data, ini.ctypes.data, results.ctypes.data, done.ctypes.data)
        rid, arr = ctypes.c_int64(), ctypes.c_double()
        rc = self.ctx.lib.dic_server_submit(self.handle, ctypes.byref(req), ctypes.byref(rid), ctypes.byref(arr))
        if rc != DIC_OK:
            raise DicError(rc, self.error() or "request refused")
        return ServedRequest(self, rid.value, arr.value, results, done, (ref, fr, ptrs, xy, ini))

    def now(self) -> float:
        return self.ctx.lib.dic_server_now(self.handle)

    def stats(self) -> dict:
        st = ServerStats()
        self.ctx.check(self.ctx.lib.dic_server_stats_get(self.handle, ctypes.byref(st)))
        return st.as_dict()

    def error(self) -> str:
        e = self.ctx.lib.dic_server_error(self.handle)
        return e.decode() if e else ""

    def close(self):
        self._release()

    def _release(self):
        if self.handle:
            self.ctx.lib.dic_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """dic_batch: one hipGraph-replayed launch over the POIs of many items."""

    def __init__(self, ctx: Context, handle, items):
        self.ctx, self.handle, self.items = ctx, handle, items
        ctx._own(self)

    def set_tuning(self, mapping: int, waves_per_cu: int = 8, chunk: int = 1, order: int = BATCH_ORDER_COST):
        t = BatchTuning(mapping, waves_per_cu, chunk, order)
        self.ctx.check(self.ctx.lib.dic_batch_set_tuning(self.handle, ctypes.byref(t)))

    def tuning(self) -> BatchTuning:
        t = BatchTuning()
        self.ctx.check(self.ctx.lib.dic_batch_get_tuning(self.handle, ctypes.byref(t)))
        return t

    def info(self):
        n, r = ctypes.c_int64(), ctypes.c_int64()
        self.ctx.check(self.ctx.lib.dic_batch_info(self.handle, ctypes.byref(n), ctypes.byref(r)))
        return n.value, r.value

    def run(self, stream: int | None = None):
        self.ctx.check(self.ctx.lib.dic_batch_run(self.handle, stream))

    def autotune(self, candidates=None, reps: int = 3):
        best, ms = BatchTuning(), ctypes.c_float()
        if candidates:
            arr = (BatchTuning * len(candidates))(*[BatchTuning(*c) for c in candidates])
            n = len(candidates)
        else:
            arr, n = None, 0
        self.ctx.check(self.ctx.lib.dic_batch_autotune(self.handle, arr, n, int(reps), ctypes.byref(best),
                                                       ctypes.byref(ms)))
        return best, ms.value

    def destroy(self):
        self._release()

    def _release(self):
        if self.handle:
            self.ctx.lib.dic_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class DeviceBuffer:
    """Raw HBM allocation of a context (dic_buffer_alloc)."""

    def __init__(self, ctx: Context, ptr: int, nbytes: int):
        self.ctx, self.ptr, self.nbytes = ctx, ptr, nbytes
        ctx._own(self)

    def download(self, dtype, count: int | None = None) -> np.ndarray:
        dt = np.dtype(dtype)
        n = self.nbytes // dt.itemsize if count is None else count
        out = np.empty(n, dt)
        self.ctx.check(self.ctx.lib.dic_buffer_download(self.ctx.handle, _ptr(out), self.ptr, out.nbytes))
        return out

    def upload(self, arr: np.ndarray):
        a = np.ascontiguousarray(arr)
        assert a.nbytes <= self.nbytes
        self.ctx.check(self.ctx.lib.dic_buffer_upload(self.ctx.handle, self.ptr, _ptr(a), a.nbytes))

    def free(self):
        self._release()

    def _release(self):
        if self.ptr and self.ctx.handle:
            self.ctx.lib.dic_buffer_free(self.ctx.handle, self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class _Handle:
    _free = None

    def __init__(self, ctx: Context, handle):
        self.ctx = ctx
        self.handle = handle
        ctx._own(self)

    def free(self):
        self._release()

    def _release(self):
        if self.handle and self.ctx.handle:
            getattr(self.ctx.lib, self._free)(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceImage(_Handle):
    _free = "dic_image_free"

    def __init__(self, ctx, handle, width, height):
        super().__init__(ctx, handle)
        self.width, self.height = width, height


class Gradient(_Handle):
    _free = "dic_grad_free"

    def __init__(self, ctx, handle, width, height):
        super().__init__(ctx, handle)
        self.width, self.height = width, height

    def download(self):
        gx = np.empty((self.height, self.width), np.float64)
        gy = np.empty_like(gx)
        self.ctx.check(self.ctx.lib.dic_grad_download(self.ctx.handle, self.handle, _ptr(gx), _ptr(gy)))
        return gx, gy


class Coefficients(_Handle):
    _free = "dic_coeffs_free"

    def info(self):
        v = [ctypes.c_int() for _ in range(5)]
        self.ctx.check(self.ctx.lib.dic_coeffs_info(self.handle, *[ctypes.byref(x) for x in v]))
        w, h, d, b, p = (x.value for x in v)
        return {"width": w, "height": h, "degree": d, "border": b, "precision": p}

    def download(self) -> np.ndarray:
        i = self.info()
        out = np.empty((i["height"] + 2 * i["border"], i["width"] + 2 * i["border"]), np.float64)
        self.ctx.check(self.ctx.lib.dic_coeffs_download(self.ctx.handle, self.handle, _ptr(out)))
        return out
