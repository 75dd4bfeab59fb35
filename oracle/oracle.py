"""ctypes loader of the CPU FP64 oracle (libdic_oracle.so, built from dic_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker and the timed CPU baseline.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# DIC_ORACLE_LIB selects another build of dic_oracle.c, e.g. the sanitizer
# build of `make -C oracle sanitize`
LIB_PATH = os.environ.get("DIC_ORACLE_LIB") or os.path.join(HERE, "libdic_oracle.so")

RESULT_DTYPE = np.dtype([
    ("x", "<f8"), ("y", "<f8"), ("p", "<f8", (12,)), ("correlation", "<f8"),
    ("status", "<i4"), ("valid", "<i4"), ("iterations", "<i4"), ("reserved", "<i4"),
])
INITIAL_DTYPE = np.dtype([
    ("u", "<f8"), ("v", "<f8"), ("du_dx", "<f8"), ("du_dy", "<f8"),
    ("dv_dx", "<f8"), ("dv_dy", "<f8"), ("confidence", "<f8"),
])


class OrcConfig(ctypes.Structure):
    _fields_ = [("subset_radius", ctypes.c_int32), ("max_iterations", ctypes.c_int32),
                ("convergence_threshold", ctypes.c_double), ("second_order", ctypes.c_int32),
                ("objective", ctypes.c_int32), ("min_valid_samples", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("min_valid_sample_ratio", ctypes.c_double)]


class OrcPrecomputed(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int32), ("height", ctypes.c_int32), ("degree", ctypes.c_int32),
                ("border", ctypes.c_int32), ("n", ctypes.c_int32), ("has_blocks", ctypes.c_int32),
                ("qk", ctypes.c_double * 36), ("coeffs", ctypes.c_void_p), ("blocks", ctypes.c_void_p),
                ("grad_x", ctypes.c_void_p), ("grad_y", ctypes.c_void_p)]


ESTIMATE_DTYPE = np.dtype([
    ("u", "<f8"), ("v", "<f8"), ("du_dx", "<f8"), ("du_dy", "<f8"), ("dv_dx", "<f8"), ("dv_dy", "<f8"),
    ("confidence", "<f8"), ("valid", "<i4"), ("reserved", "<i4"), ("zncc", "<f8"), ("znssd", "<f8"),
])
SEED_EVAL_DTYPE = np.dtype([
    ("x", "<f8"), ("y", "<f8"), ("displacement", RESULT_DTYPE), ("quality", "<f8"),
    ("displacement_norm", "<f8"), ("quality_passed", "<i4"), ("valid", "<i4"),
])


class OrcSearchParams(ctypes.Structure):
    _fields_ = [("subset_radius", ctypes.c_int32), ("search_radius", ctypes.c_int32),
                ("pyramid_enabled", ctypes.c_int32), ("pyramid_scale", ctypes.c_int32),
                ("pyramid_refinement_radius", ctypes.c_int32), ("around_prior", ctypes.c_int32),
                ("masked", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class OrcSeedParams(ctypes.Structure):
    _fields_ = [("search", OrcSearchParams), ("subpixel_enabled", ctypes.c_int32),
                ("subpixel_shape_order", ctypes.c_int32), ("subpixel_optimizer", ctypes.c_int32),
                ("subpixel_objective", ctypes.c_int32), ("subpixel_radius", ctypes.c_int32),
                ("subpixel_max_iterations", ctypes.c_int32), ("subpixel_convergence_threshold", ctypes.c_double),
                ("seed_count", ctypes.c_int32), ("quality_metric", ctypes.c_int32), ("min_zncc", ctypes.c_double),
                ("max_znssd", ctypes.c_double), ("max_ssd", ctypes.c_double),
                ("min_displacement_norm", ctypes.c_double), ("min_texture_std", ctypes.c_double),
                ("kmeans_iterations", ctypes.c_int32), ("kmeans_sample_limit", ctypes.c_int32),
                ("truncate_roi_subsets", ctypes.c_int32), ("min_valid_samples", ctypes.c_int32),
                ("min_valid_sample_ratio", ctypes.c_double), ("degree", ctypes.c_int32), ("border", ctypes.c_int32)]


# Defaults of SeedIntegerSearchConfig / SeedSubpixelRefinementConfig / SeedSelectionConfig
# (seed_config.hpp:33-69) and the SubsetConfig fields the selector reads (subset_config.hpp:33-50).
SEARCH_DEFAULTS = dict(subset_radius=10, search_radius=30, pyramid_enabled=1, pyramid_scale=4,
                       pyramid_refinement_radius=4, around_prior=0, masked=0)
SEED_DEFAULTS = dict(subpixel_enabled=1, subpixel_shape_order=1, subpixel_optimizer=0, subpixel_objective=0,
                     subpixel_radius=15, subpixel_max_iterations=30, subpixel_convergence_threshold=1e-3,
                     seed_count=256, quality_metric=1, min_zncc=0.7, max_znssd=0.2, max_ssd=0.05,
                     min_displacement_norm=0.0, min_texture_std=0.0, kmeans_iterations=20,
                     kmeans_sample_limit=20000, truncate_roi_subsets=0, min_valid_samples=12,
                     min_valid_sample_ratio=0.5, degree=5, border=3)


def search_params(**kw) -> OrcSearchParams:
    d = dict(SEARCH_DEFAULTS)
    d.update(kw)
    return OrcSearchParams(**d)


def seed_params(search: dict | None = None, **kw) -> OrcSeedParams:
    d = dict(SEED_DEFAULTS)
    d.update(kw)
    return OrcSeedParams(search=search_params(**(search or {})), **d)


_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `make -C oracle`")
    lib = ctypes.CDLL(LIB_PATH)
    vp, ci, cd = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    lib.orc_basis.argtypes = [cd, ci, ci]
    lib.orc_basis.restype = cd
    lib.orc_build_qk.argtypes = [ci, vp]
    lib.orc_symmetric_index.argtypes = [ci, ci]
    lib.orc_symmetric_index.restype = ci
    lib.orc_deconvolve_periodic.argtypes = [vp, ci, ci]
    lib.orc_form_coefficients.argtypes = [vp, ci, ci, ci, ci, ci, vp]
    lib.orc_form_coefficients.restype = ci
    lib.orc_precompute.argtypes = [vp, ci, ci, ci, ci, ci, ci, ctypes.POINTER(OrcPrecomputed)]
    lib.orc_precompute.restype = ci
    lib.orc_precomputed_free.argtypes = [ctypes.POINTER(OrcPrecomputed)]
    lib.orc_set_threads.argtypes = [ci]
    lib.orc_value.argtypes = [ctypes.POINTER(OrcPrecomputed), cd, cd]
    lib.orc_value.restype = cd
    lib.orc_value_batch.argtypes = [ctypes.POINTER(OrcPrecomputed), vp, ci, vp, vp]
    lib.orc_reference_gradient_cache.argtypes = [vp, ci, ci, vp, vp]
    lib.orc_ldlt_factor.argtypes = [vp, ci, vp, ctypes.POINTER(ci)]
    lib.orc_ldlt_factor.restype = ci
    lib.orc_ldlt_solve.argtypes = [vp, ci, vp, vp, vp]
    lib.orc_fullpivlu_inverse6.argtypes = [vp, vp]
    lib.orc_fullpivlu_inverse6.restype = ci
    lib.orc_ic_affine_update.argtypes = [vp, vp, vp]
    lib.orc_second_order_warp_matrix.argtypes = [vp, vp]
    lib.orc_ic_second_order_update.argtypes = [vp, vp, vp]
    lib.orc_icgn_solve_batch.argtypes = [ctypes.POINTER(OrcConfig), vp, ci, ci, vp, vp,
                                         ctypes.POINTER(OrcPrecomputed), vp, vp, vp, ci, vp, ci]
    lib.orc_fgn_compose.argtypes = [vp, vp, ci, vp]
    lib.orc_fgn_solve_batch.argtypes = [ctypes.POINTER(OrcConfig), vp, ci, ci, ctypes.POINTER(OrcPrecomputed), vp,
                                        vp, vp, ci, vp, ci]
    lib.orc_propagate.argtypes = [ctypes.POINTER(OrcConfig), ci, ci, ctypes.c_double, ci, vp, ci, ci, vp, vp,
                                  ctypes.POINTER(OrcPrecomputed), vp, vp, ci, vp]
    lib.orc_propagate.restype = ci
    lib.orc_poi_grid.argtypes = [ci, ci, ci, vp]
    lib.orc_poi_grid.restype = ci
    lib.orc_integer_search.argtypes = [ctypes.POINTER(OrcSearchParams), vp, vp, ci, ci, vp, vp, vp, ci, vp, ci]
    lib.orc_subset_zncc.argtypes = [vp, vp, ci, ci, ci, ci, ci, ci, ci]
    lib.orc_subset_zncc.restype = cd
    lib.orc_texture_std.argtypes = [vp, ci, ci, ci, ci, ci]
    lib.orc_texture_std.restype = cd
    lib.orc_deterministic_sample.argtypes = [ci, ci, vp]
    lib.orc_deterministic_sample.restype = ci
    lib.orc_kmeans.argtypes = [vp, ci, ci, ci, vp]
    lib.orc_kmeans.restype = ci
    lib.orc_seed_candidates.argtypes = [ctypes.POINTER(OrcSeedParams), vp, ci, ci, vp, vp, ci]
    lib.orc_seed_candidates.restype = ci
    lib.orc_select_seed.argtypes = [ctypes.POINTER(OrcSeedParams), vp, vp, ci, ci, vp, vp, ci,
                                    ctypes.POINTER(ci)]
    lib.orc_select_seed.restype = ci
    _lib = lib
    return lib


def _p(a):
    return a.ctypes.data if a is not None else None


def basis(x: float, m: int, degree: int) -> float:
    return load().orc_basis(float(x), int(m), int(degree))


def build_qk(degree: int) -> np.ndarray:
    n = degree + 1
    out = np.zeros(36)
    load().orc_build_qk(degree, _p(out))
    return out[: n * n].reshape(n, n).copy()


def symmetric_index(v: int, size: int) -> int:
    return load().orc_symmetric_index(int(v), int(size))


def deconvolve_periodic(line: np.ndarray, degree: int) -> np.ndarray:
    a = np.ascontiguousarray(line, np.float64).copy()
    load().orc_deconvolve_periodic(_p(a), a.size, int(degree))
    return a


def form_coefficients(img: np.ndarray, degree: int = 5, border: int = 3, exact: bool = True) -> np.ndarray:
    img = np.ascontiguousarray(img, np.float32)
    H, W = img.shape
    b = max(3, border)
    out = np.empty((H + 2 * b, W + 2 * b))
    load().orc_form_coefficients(_p(img), W, H, degree, border, int(exact), _p(out))
    return out


class Precomputed:
    """BSplineImagePreprocessor::compute (with_blocks) or compute_lazy."""

    def __init__(self, img: np.ndarray, degree: int = 5, border: int = 3, exact: bool = True,
                 with_blocks: bool = True):
        self._img = np.ascontiguousarray(img, np.float32)
        H, W = self._img.shape
        self.s = OrcPrecomputed()
        rc = load().orc_precompute(_p(self._img), W, H, degree, border, int(exact), int(with_blocks),
                                   ctypes.byref(self.s))
        if rc != 0:
            raise ValueError(f"orc_precompute failed ({rc})")
        self.width, self.height = W, H

    def __del__(self):
        try:
            load().orc_precomputed_free(ctypes.byref(self.s))
        except Exception:
            pass

    def coefficients(self) -> np.ndarray:
        b = self.s.border
        n = (self.height + 2 * b) * (self.width + 2 * b)
        buf = (ctypes.c_double * n).from_address(self.s.coeffs)
        return np.ctypeslib.as_array(buf).reshape(self.height + 2 * b, self.width + 2 * b).copy()

    def gradients(self):
        n = self.width * self.height
        gx = np.ctypeslib.as_array((ctypes.c_double * n).from_address(self.s.grad_x)).reshape(self.height, self.width)
        gy = np.ctypeslib.as_array((ctypes.c_double * n).from_address(self.s.grad_y)).reshape(self.height, self.width)
        return gx.copy(), gy.copy()

    def value(self, x: float, y: float) -> float:
        return load().orc_value(ctypes.byref(self.s), float(x), float(y))

    def sample(self, xy: np.ndarray):
        xy = np.ascontiguousarray(xy, np.float64).reshape(-1, 2)
        n = xy.shape[0]
        val = np.empty(n)
        grad = np.empty((n, 2))
        load().orc_value_batch(ctypes.byref(self.s), _p(xy), n, _p(val), _p(grad))
        return val, grad


def set_threads(n: int) -> None:
    """Threads of the per-pixel block build (Precomputed) and value_batch
    regions started from this thread; n <= 0 restores the default."""
    load().orc_set_threads(int(n))


def reference_gradient_cache(img: np.ndarray):
    img = np.ascontiguousarray(img, np.float32)
    H, W = img.shape
    gx = np.empty((H, W))
    gy = np.empty((H, W))
    load().orc_reference_gradient_cache(_p(img), W, H, _p(gx), _p(gy))
    return gx, gy


def ldlt(a: np.ndarray):
    m = np.ascontiguousarray(a, np.float64).copy()
    n = m.shape[0]
    tr = np.zeros(n, np.int32)
    pos = ctypes.c_int(0)
    info = load().orc_ldlt_factor(_p(m), n, _p(tr), ctypes.byref(pos))
    return m, tr, info, bool(pos.value)


def ldlt_solve(factor: np.ndarray, tr: np.ndarray, rhs: np.ndarray) -> np.ndarray:
    n = factor.shape[0]
    x = np.empty(n)
    r = np.ascontiguousarray(rhs, np.float64)
    load().orc_ldlt_solve(_p(np.ascontiguousarray(factor)), n, _p(np.ascontiguousarray(tr, np.int32)), _p(r), _p(x))
    return x


def fullpivlu_inverse6(m: np.ndarray):
    a = np.ascontiguousarray(m, np.float64)
    inv = np.zeros((6, 6))
    ok = load().orc_fullpivlu_inverse6(_p(a), _p(inv))
    return bool(ok), inv


def ic_affine_update(p, dp):
    p = np.ascontiguousarray(p, np.float64)
    dp = np.ascontiguousarray(dp, np.float64)
    out = np.empty(6)
    load().orc_ic_affine_update(_p(p), _p(dp), _p(out))
    return out


def second_order_warp_matrix(p):
    p = np.ascontiguousarray(p, np.float64)
    g = np.empty((6, 6))
    load().orc_second_order_warp_matrix(_p(p), _p(g))
    return g


def ic_second_order_update(p, dp):
    p = np.ascontiguousarray(p, np.float64)
    dp = np.ascontiguousarray(dp, np.float64)
    out = np.empty(12)
    load().orc_ic_second_order_update(_p(p), _p(dp), _p(out))
    return out


def icgn_solve(reference: np.ndarray, pre: Precomputed, grad_x: np.ndarray, grad_y: np.ndarray,
               poi_xy: np.ndarray, init: np.ndarray, radius: int = 15, max_iterations: int = 30,
               threshold: float = 1e-3, shape_order: int = 1, objective: int = 0, roi: np.ndarray | None = None,
               min_valid_samples: int = 12, min_valid_sample_ratio: float = 0.5, threads: int = 0) -> np.ndarray:
    ref = np.ascontiguousarray(reference, np.float32)
    H, W = ref.shape
    cfg = OrcConfig(int(radius), int(max_iterations), float(threshold), 1 if shape_order == 2 else 0, int(objective),
                    int(min_valid_samples), 0, float(min_valid_sample_ratio))
    xy = np.ascontiguousarray(poi_xy, np.float64).reshape(-1, 2)
    ini = np.ascontiguousarray(init, INITIAL_DTYPE)
    gx = np.ascontiguousarray(grad_x, np.float64)
    gy = np.ascontiguousarray(grad_y, np.float64)
    r = np.ascontiguousarray(roi, np.uint8) if roi is not None else None
    out = np.zeros(xy.shape[0], RESULT_DTYPE)
    load().orc_icgn_solve_batch(ctypes.byref(cfg), _p(ref), W, H, _p(gx), _p(gy), ctypes.byref(pre.s), _p(r),
                                _p(xy), _p(ini), xy.shape[0], _p(out), int(threads))
    return out


def fgn_compose(p: np.ndarray, dp: np.ndarray) -> np.ndarray:
    """compose_warp of forward_gauss_newton.cpp (:64-76 first order, :1025-1053 second order)."""
    P = len(p)
    a = np.ascontiguousarray(p, np.float64)
    b = np.ascontiguousarray(dp, np.float64)
    out = np.zeros(P)
    load().orc_fgn_compose(_p(a), _p(b), P, _p(out))
    return out


def fgn_solve(reference: np.ndarray, pre: Precomputed, poi_xy: np.ndarray, init: np.ndarray, radius: int = 15,
              max_iterations: int = 30, threshold: float = 1e-3, shape_order: int = 1, objective: int = 0,
              roi: np.ndarray | None = None, min_valid_samples: int = 12, min_valid_sample_ratio: float = 0.5,
              threads: int = 0) -> np.ndarray:
    """ForwardGaussNewtonSolver (forward_gauss_newton.cpp) over POIs; `pre` is the deformed image's precompute."""
    ref = np.ascontiguousarray(reference, np.float32)
    H, W = ref.shape
    cfg = OrcConfig(int(radius), int(max_iterations), float(threshold), 1 if shape_order == 2 else 0, int(objective),
                    int(min_valid_samples), 0, float(min_valid_sample_ratio))
    xy = np.ascontiguousarray(poi_xy, np.float64).reshape(-1, 2)
    ini = np.ascontiguousarray(init, INITIAL_DTYPE)
    r = np.ascontiguousarray(roi, np.uint8) if roi is not None else None
    out = np.zeros(xy.shape[0], RESULT_DTYPE)
    load().orc_fgn_solve_batch(ctypes.byref(cfg), _p(ref), W, H, ctypes.byref(pre.s), _p(r), _p(xy), _p(ini),
                               xy.shape[0], _p(out), int(threads))
    return out


def propagate(reference: np.ndarray, pre: Precomputed, grad_x, grad_y, roi: np.ndarray, seeds: np.ndarray,
              radius: int = 15, max_iterations: int = 30, threshold: float = 1e-3, shape_order: int = 1,
              objective: int = 0, optimizer: int = 0, spacing: int = 5, max_znssd: float = 0.5,
              truncate_roi: bool = False, min_valid_samples: int = 12, min_valid_sample_ratio: float = 0.5):
    """ReliabilityPropagation::propagate (reliability_propagation.cpp:130-394).
    seeds: (n, 9) x, y, u, v, du_dx, du_dy, dv_dx, dv_dy, correlation.
    Returns (grid results (gh, gw), points_computed)."""
    ref = np.ascontiguousarray(reference, np.float32)
    H, W = ref.shape
    step = max(spacing, 0) + 1
    gw, gh = (W + step - 1) // step, (H + step - 1) // step
    cfg = OrcConfig(int(radius), int(max_iterations), float(threshold), 1 if shape_order == 2 else 0, int(objective),
                    int(min_valid_samples), 0, float(min_valid_sample_ratio))
    gx = np.ascontiguousarray(grad_x, np.float64) if grad_x is not None else None
    gy = np.ascontiguousarray(grad_y, np.float64) if grad_y is not None else None
    r = np.ascontiguousarray(roi, np.uint8)
    sd = np.ascontiguousarray(seeds, np.float64).reshape(-1, 9)
    out = np.zeros(gw * gh, RESULT_DTYPE)
    n = load().orc_propagate(ctypes.byref(cfg), int(optimizer), int(spacing), float(max_znssd), int(truncate_roi),
                             _p(ref), W, H, _p(gx), _p(gy), ctypes.byref(pre.s), _p(r), _p(sd), sd.shape[0], _p(out))
    return out.reshape(gh, gw), n


def integer_search(reference: np.ndarray, deformed: np.ndarray, xy: np.ndarray, params: OrcSearchParams | None = None,
                   roi: np.ndarray | None = None, prior_uv: np.ndarray | None = None, threads: int = 0) -> np.ndarray:
    """IntegerSearchInitializer over points (integer_search.cpp:303-645); ESTIMATE_DTYPE records."""
    params = params or search_params()
    ref = np.ascontiguousarray(reference, np.float32)
    dfm = np.ascontiguousarray(deformed, np.float32)
    H, W = ref.shape
    pts = np.ascontiguousarray(xy, np.float64).reshape(-1, 2)
    pr = np.ascontiguousarray(prior_uv, np.float64).reshape(-1, 2) if prior_uv is not None else None
    r = np.ascontiguousarray(roi, np.uint8) if roi is not None else None
    out = np.zeros(pts.shape[0], ESTIMATE_DTYPE)
    load().orc_integer_search(ctypes.byref(params), _p(ref), _p(dfm), W, H, _p(r), _p(pts), _p(pr), pts.shape[0],
                              _p(out), int(threads))
    return out


def subset_zncc(reference: np.ndarray, deformed: np.ndarray, radius: int, cx: int, cy: int, u: int, v: int) -> float:
    ref = np.ascontiguousarray(reference, np.float32)
    dfm = np.ascontiguousarray(deformed, np.float32)
    H, W = ref.shape
    return load().orc_subset_zncc(_p(ref), _p(dfm), W, H, radius, cx, cy, u, v)


def texture_std(reference: np.ndarray, x: int, y: int, radius: int) -> float:
    ref = np.ascontiguousarray(reference, np.float32)
    H, W = ref.shape
    return load().orc_texture_std(_p(ref), W, H, x, y, radius)


def deterministic_sample(n: int, limit: int) -> np.ndarray:
    idx = np.zeros(max(1, min(n, max(limit, 1)) if limit > 0 else n), np.int32)
    k = load().orc_deterministic_sample(n, limit, _p(idx))
    return idx[:k]


def kmeans(points: np.ndarray, requested: int, iterations: int) -> np.ndarray:
    pts = np.ascontiguousarray(points, np.int32).reshape(-1, 2)
    c = np.zeros((max(1, min(requested, pts.shape[0])), 2))
    k = load().orc_kmeans(_p(pts), pts.shape[0], requested, iterations, _p(c))
    return c[:k]


def seed_candidates(reference: np.ndarray, roi: np.ndarray, params: OrcSeedParams | None = None) -> np.ndarray:
    params = params or seed_params()
    ref = np.ascontiguousarray(reference, np.float32)
    H, W = ref.shape
    r = np.ascontiguousarray(roi, np.uint8)
    cap = max(1, params.seed_count)
    out = np.zeros((cap, 2))
    k = load().orc_seed_candidates(ctypes.byref(params), _p(ref), W, H, _p(r), _p(out), cap)
    return out[:k]


def select_seed(reference: np.ndarray, deformed: np.ndarray, roi: np.ndarray, params: OrcSeedParams | None = None):
    """SeedSelector::select_best_seed (seed_selector.cpp:146-206).  Returns
    (evaluations SEED_EVAL_DTYPE, index of the best seed or -1)."""
    params = params or seed_params()
    ref = np.ascontiguousarray(reference, np.float32)
    dfm = np.ascontiguousarray(deformed, np.float32)
    H, W = ref.shape
    r = np.ascontiguousarray(roi, np.uint8)
    cap = max(1, params.seed_count)
    out = np.zeros(cap, SEED_EVAL_DTYPE)
    n = ctypes.c_int(0)
    best = load().orc_select_seed(ctypes.byref(params), _p(ref), _p(dfm), W, H, _p(r), _p(out), cap,
                                  ctypes.byref(n))
    return out[:n.value], best
