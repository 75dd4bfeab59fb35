"""Host-side mirror of Traditional-DIC's subset solver surface on the HIP library.

Reference interfaces mirrored (paths inside the Traditional-DIC tree):
  SubsetConfig                  include/dic/subset/subset_config.hpp:33-50
  config_from_dict              bindings/python/bind_subset.cpp:158-248
  ICGNSolver::solve / solve_with_interpolators / solve_with_mask
                                src/subset/solver/icgn.cpp:363-432
  results_to_dict keys          bindings/python/bind_subset.cpp:102-153
  POI grid + margin defaults    src/subset/seed/reliability_propagation.cpp:154-176, :318-323

The numerics run in libdic_hip.so; nothing here computes a correlation.
"""
from __future__ import annotations

import dataclasses
import threading

from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import INITIAL_DTYPE, RESULT_DTYPE, IcgnParams


@dataclass
class BSplinePrecomputeConfig:
    """include/dic/interpolation/bspline.hpp:42-67 (defaults: quintic, border 3, exact prefilter)."""
    degree: int = 5
    border: int = 3
    use_exact_prefilter: bool = True


@dataclass
class SeedIntegerSearchConfig:
    """include/dic/initialization/seed_config.hpp:33-40."""
    subset_radius: int = 10
    search_radius: int = 30
    sift_enabled: bool = False
    pyramid_enabled: bool = True
    pyramid_scale: int = 4
    pyramid_refinement_radius: int = 4


@dataclass
class SeedSubpixelRefinementConfig:
    """seed_config.hpp:42-50."""
    enabled: bool = True
    shape_function: str = "first_order"   # first_order | second_order
    optimizer: str = "icgn"               # icgn | forward_gauss_newton
    objective: str = "znssd"              # znssd | ssd
    subset_radius: int = 15
    max_iterations: int = 30
    convergence_threshold: float = 1e-3


@dataclass
class SeedInitializationConfig:
    """seed_config.hpp:52-56 (the integer-search method; SIFT is not part of this library)."""
    method: str = "integer_search"
    integer_search: SeedIntegerSearchConfig = field(default_factory=SeedIntegerSearchConfig)
    subpixel: SeedSubpixelRefinementConfig = field(default_factory=SeedSubpixelRefinementConfig)


@dataclass
class SeedSelectionConfig:
    """seed_config.hpp:58-69."""
    seed_count: int = 256
    threads: int = 1
    quality_metric: str = "znssd"         # zncc | znssd | ssd
    min_zncc: float = 0.7
    max_znssd: float = 0.2
    max_ssd: float = 0.05
    min_displacement_norm: float = 0.0
    min_texture_std: float = 0.0
    kmeans_iterations: int = 20
    kmeans_sample_limit: int = 20000


@dataclass
class SubsetConfig:
    """Fields of dic::SubsetConfig (defaults as in subset_config.hpp:33-50)."""
    subset_radius: int = 15
    max_iterations: int = 30
    convergence_threshold: float = 1e-3
    use_second_order: bool = False
    search_radius: int = 20
    propagation_spacing: int = 5
    propagation_max_znssd: float = 0.5
    truncate_roi_subsets: bool = False
    # walk order of the propagation (not a field of the reference's config):
    # "heap" follows its priority queue exactly, "frontier" expands whole
    # frontiers per launch (faster; a strict max_znssd can reject other points)
    propagation_order: str = "heap"
    min_valid_sample_ratio: float = 0.5
    min_valid_samples: int = 12
    image_precompute: BSplinePrecomputeConfig = field(default_factory=BSplinePrecomputeConfig)
    seed_initialization: SeedInitializationConfig = field(default_factory=SeedInitializationConfig)
    seed_selection: SeedSelectionConfig = field(default_factory=SeedSelectionConfig)
    shape_function: str = "first_order"   # first_order | second_order
    optimizer: str = "icgn"               # icgn | forward_gauss_newton
    objective: str = "znssd"              # znssd | ssd

    @property
    def shape_order(self) -> int:
        return 2 if (self.shape_function == "second_order" or self.use_second_order) else 1


def config_from_dict(d: dict | None) -> SubsetConfig:
    """Restates config_from_dict (bind_subset.cpp:158-248) with the same
    validation messages and the same reach: interpolation degree maps only 1
    and 3 (any other value keeps the quintic default, :195-200); of the seed
    sub-pixel refinement only enabled / subset_radius / max_iterations /
    convergence_threshold are read (its shape function and optimizer keep
    their defaults, its objective follows correlation.criterion, :190, :203-221);
    subset.truncate_roi_subsets is not read."""
    cfg = SubsetConfig()
    if d is None:
        return cfg
    if not isinstance(d, dict):
        raise RuntimeError("config must be a dict (from YAML) or None.")
    s = d.get("subset")
    if isinstance(s, dict) and "radius" in s:
        cfg.subset_radius = int(s["radius"])
    o = d.get("optimization")
    if isinstance(o, dict):
        cfg.max_iterations = int(o.get("max_iterations", cfg.max_iterations))
        cfg.convergence_threshold = float(o.get("convergence_threshold", cfg.convergence_threshold))
        if "method" in o:
            m = str(o["method"])
            if m == "icgn":
                cfg.optimizer = "icgn"
            elif m in ("forward_gauss_newton", "fgn"):
                cfg.optimizer = "forward_gauss_newton"
            else:
                raise RuntimeError("subset optimization.method must be 'icgn', 'fgn', or 'forward_gauss_newton'.")
    c = d.get("correlation")
    if isinstance(c, dict) and "criterion" in c:
        crit = str(c["criterion"])
        if crit not in ("ssd", "znssd"):
            raise RuntimeError("subset correlation.criterion must be 'ssd' or 'znssd'.")
        cfg.objective = crit
        cfg.seed_initialization.subpixel.objective = crit
    it = d.get("interpolation")
    if isinstance(it, dict) and "degree" in it:
        deg = int(it["degree"])
        if deg in (1, 3):
            cfg.image_precompute.degree = deg
    ini = d.get("initialization")
    if isinstance(ini, dict):
        if "method" in ini:
            if str(ini["method"]) != "integer_search":
                raise RuntimeError("subset initialization.method must be integer_search")
            cfg.seed_initialization.method = "integer_search"
        isr = ini.get("integer_search")
        if isinstance(isr, dict):
            c = cfg.seed_initialization.integer_search
            c.subset_radius = int(isr.get("subset_radius", c.subset_radius))
            c.search_radius = int(isr.get("search_radius", c.search_radius))
            if "sift_enabled" in isr:
                c.sift_enabled = bool(isr["sift_enabled"])
            if "pyramid_enabled" in isr:
                c.pyramid_enabled = bool(isr["pyramid_enabled"])
            c.pyramid_scale = int(isr.get("pyramid_scale", c.pyramid_scale))
            c.pyramid_refinement_radius = int(isr.get("pyramid_refinement_radius", c.pyramid_refinement_radius))
        spr = ini.get("subpixel_refinement")
        if isinstance(spr, dict):
            c = cfg.seed_initialization.subpixel
            if "enabled" in spr:
                c.enabled = bool(spr["enabled"])
            c.subset_radius = int(spr.get("subset_radius", c.subset_radius))
            c.max_iterations = int(spr.get("max_iterations", c.max_iterations))
            c.convergence_threshold = float(spr.get("convergence_threshold", c.convergence_threshold))
    ss = d.get("seed_selection")
    if isinstance(ss, dict):
        c = cfg.seed_selection
        c.seed_count = int(ss.get("seed_count", c.seed_count))
        c.max_znssd = float(ss.get("max_znssd", c.max_znssd))
        c.min_displacement_norm = float(ss.get("min_displacement_norm", c.min_displacement_norm))
        c.min_texture_std = float(ss.get("min_texture_std", c.min_texture_std))
        c.kmeans_iterations = int(ss.get("kmeans_iterations", c.kmeans_iterations))
        c.kmeans_sample_limit = int(ss.get("kmeans_sample_limit", c.kmeans_sample_limit))
    rp = d.get("reliability_propagation")
    if isinstance(rp, dict):
        cfg.propagation_spacing = int(rp.get("spacing", cfg.propagation_spacing))
        cfg.propagation_max_znssd = float(rp.get("max_znssd", cfg.propagation_max_znssd))
        cfg.propagation_order = str(rp.get("order", cfg.propagation_order))
    sf = d.get("shape_function")
    if isinstance(sf, dict) and "order" in sf:
        order = str(sf["order"])
        if order in ("first_order", "1"):
            cfg.shape_function = "first_order"
        elif order in ("second_order", "2"):
            cfg.shape_function = "second_order"
        else:
            raise RuntimeError("subset shape_function.order must be 'first_order'/'1' or 'second_order'/'2'.")
    return cfg


def make_params(cfg: SubsetConfig, precision: int, masked: bool = False) -> IcgnParams:
    p = IcgnParams()
    p.subset_radius = int(cfg.subset_radius)
    p.max_iterations = int(cfg.max_iterations)
    p.convergence_threshold = float(cfg.convergence_threshold)
    p.shape_order = cfg.shape_order
    p.objective = _lib.OBJECTIVE_SSD if cfg.objective == "ssd" else _lib.OBJECTIVE_ZNSSD
    p.optimizer = _lib.OPTIMIZER_FGN if cfg.optimizer == "forward_gauss_newton" else _lib.OPTIMIZER_ICGN
    p.precision = int(precision)
    p.masked = 1 if masked else 0
    p.min_valid_samples = int(cfg.min_valid_samples)
    p.min_valid_sample_ratio = float(cfg.min_valid_sample_ratio)
    return p


def initial_guesses(u, v, du_dx=0.0, du_dy=0.0, dv_dx=0.0, dv_dy=0.0, confidence=0.0) -> np.ndarray:
    """Array of InitialDisplacement records (initializer.hpp:33-44)."""
    u = np.atleast_1d(np.asarray(u, np.float64))
    out = np.zeros(u.shape[0], INITIAL_DTYPE)
    out["u"] = u
    out["v"] = v
    out["du_dx"] = du_dx
    out["du_dy"] = du_dy
    out["dv_dx"] = dv_dx
    out["dv_dy"] = dv_dy
    out["confidence"] = confidence
    return out


class PreparedPair:
    """Device-resident state of one (reference, deformed) pair: the float32
    images, the reference gradient planes and the deformed image's B-spline
    coefficients (the reference's gradient cache + BSplinePrecomputedImage)."""

    def __init__(self, ctx: _lib.Context, reference, deformed, cfg: SubsetConfig,
                 precision: int, gradient_mode: int = _lib.GRADIENT_CENTRAL):
        """reference / deformed: host arrays (uploaded here) or DeviceImage handles (borrowed)."""
        self.ctx = ctx
        self._owned = not isinstance(reference, _lib.DeviceImage)
        if self._owned:
            ref = np.ascontiguousarray(reference, np.float32)
            dfm = np.ascontiguousarray(deformed, np.float32)
            self.ref = ctx.upload_image(ref)
            self.deformed = ctx.upload_image(dfm)
        else:
            self.ref, self.deformed = reference, deformed
        self.width, self.height = self.ref.width, self.ref.height
        self.precision = precision
        ip = cfg.image_precompute
        self.grad = ctx.ref_gradient(self.ref, gradient_mode, ip.degree, ip.border, precision)
        self.coeffs = ctx.prefilter(self.deformed, ip.degree, ip.border, ip.use_exact_prefilter, precision)

    def free(self):
        for h in (self.coeffs, self.grad) + ((self.deformed, self.ref) if self._owned else ()):
            h.free()


class _HostPairWorkspace:
    """Device buffers for a solver's calls on host images of one size and one
    precompute configuration.  Every call uploads both images into them and
    recomputes the central-difference gradient cache and the coefficients, so
    a run of same-size calls allocates once; nothing computed is carried from
    one call to the next."""

    def __init__(self, ctx: _lib.Context, key: tuple):
        self.ctx, self.key = ctx, key
        w, h = key[0], key[1]
        self.width, self.height = w, h
        self.bufs = [ctx.alloc(4 * w * h), ctx.alloc(4 * w * h)]
        self.ref = ctx.wrap_image(self.bufs[0].ptr, w, h)
        self.deformed = ctx.wrap_image(self.bufs[1].ptr, w, h)
        self.grad = self.coeffs = None

    def fill(self, reference, deformed, cfg: SubsetConfig, precision: int):
        self.bufs[0].upload(np.ascontiguousarray(reference, np.float32))
        self.bufs[1].upload(np.ascontiguousarray(deformed, np.float32))
        ip = cfg.image_precompute
        if self.grad is None:
            self.grad = self.ctx.ref_gradient(self.ref, _lib.GRADIENT_CENTRAL, ip.degree, ip.border, precision)
            self.coeffs = self.ctx.prefilter(self.deformed, ip.degree, ip.border, ip.use_exact_prefilter, precision)
        else:
            self.ctx.ref_gradient_into(self.ref, self.grad)
            self.ctx.prefilter_into(self.deformed, self.coeffs)

    def free(self):
        for h in (self.coeffs, self.grad, self.deformed, self.ref):
            if h is not None:
                h.free()
        for b in self.bufs:
            b.free()


class ICGNSolver:
    """IC-GN on the GPU behind the reference's ICGNSolver call pattern.

    precision: "fp64" (plain drop-in, the reference's arithmetic) or "mixed"
    (FP32 sampling and compensated sums, FP64 solve/update).
    gradient: "central" (pipeline convention, reliability_propagation.cpp:84-112)
    or "bspline" (standalone solve(), bspline.cpp:289-294).

    Nothing computed is cached between calls: solve() / solve_batch() on host
    images upload both images and build the pair's gradient cache and B-spline
    coefficients for that call, as the reference's solve() builds its
    interpolators per call (icgn.cpp:363-382).  With central-difference
    gradients the device buffers they go into are kept for the next call of
    the same size and configuration (release() frees them); with B-spline
    gradients the pair is prepared and freed per call.  A caller that solves one pair
    many times prepares it once with prepare() and hands the PreparedPair to
    solve_with_interpolators / solve_with_mask / solve_batch(pair=...), as the
    reference's callers hand over their interpolators
    (reliability_propagation.cpp:182-185, :341-354)."""

    def __init__(self, config: SubsetConfig | dict | None = None, precision: str = "fp64",
                 gradient: str = "central", device: int = 0, context: _lib.Context | None = None):
        self.config = config if isinstance(config, SubsetConfig) else config_from_dict(config)
        if precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        self.precision = _lib.PRECISION_FP64 if precision == "fp64" else _lib.PRECISION_MIXED
        self.gradient_mode = _lib.GRADIENT_CENTRAL if gradient == "central" else _lib.GRADIENT_BSPLINE_HALF
        self.ctx = context if context is not None else _lib.Context(device)
        self._ws = None
        self._ws_lock = threading.Lock()

    def release(self):
        """Frees the device buffers kept for calls on host images."""
        with self._ws_lock:
            if self._ws is not None:
                self._ws.free()
                self._ws = None

    def _workspace(self, w: int, h: int) -> _HostPairWorkspace:
        ip = self.config.image_precompute
        key = (w, h, ip.degree, ip.border, bool(ip.use_exact_prefilter), self.precision)
        if self._ws is None or self._ws.key != key:
            if self._ws is not None:
                self._ws.free()
                self._ws = None
            self._ws = _HostPairWorkspace(self.ctx, key)
        return self._ws

    def prepare(self, reference, deformed) -> PreparedPair:
        """A new device-resident pair (uploaded host arrays or borrowed
        DeviceImage handles) with its gradient cache and deformed coefficients;
        the caller owns it and frees it with free()."""
        return PreparedPair(self.ctx, reference, deformed, self.config, self.precision, self.gradient_mode)

    def solve_batch(self, reference, deformed, points, initial, roi=None, pair: PreparedPair | None = None) -> np.ndarray:
        """Many POIs at once; returns RESULT_DTYPE records (Displacement2D + iterations).
        With roi the masked variants run (solve_with_mask).  pair: a PreparedPair
        of (reference, deformed) to use instead of preparing one for this call;
        it must have their size."""
        xy = np.asarray(points, np.float64).reshape(-1, 2)
        ini = np.ascontiguousarray(initial, INITIAL_DTYPE)
        rw, rh = _image_size(reference)
        dw, dh = _image_size(deformed)
        if (rw, rh) != (dw, dh) or rw <= 0 or rh <= 0:
            return invalid_results(xy, ini)  # icgn.cpp:464-470
        if roi is not None:
            roi = np.asarray(roi)
            if roi.shape != (rh, rw):
                return invalid_results(xy, ini)  # icgn.cpp:648-652
            roi = np.ascontiguousarray(roi, np.uint8)
        host_images = not isinstance(reference, _lib.DeviceImage) and not isinstance(deformed, _lib.DeviceImage)
        if pair is None and host_images and self.gradient_mode == _lib.GRADIENT_CENTRAL:
            with self._ws_lock:
                ws = self._workspace(rw, rh)
                ws.fill(reference, deformed, self.config, self.precision)
                params = make_params(self.config, self.precision, roi is not None)
                return self.ctx.solve(params, ws.ref, ws.grad, ws.coeffs, xy, ini, roi)
        owned = pair is None
        if owned:
            pair = self.prepare(reference, deformed)
        elif (pair.width, pair.height) != (rw, rh):
            raise ValueError(f"prepared pair is {pair.width}x{pair.height}, the images are {rw}x{rh}")
        try:
            params = make_params(self.config, self.precision, roi is not None)
            return self.ctx.solve(params, pair.ref, pair.grad, pair.coeffs, xy, ini, roi)
        finally:
            if owned:
                pair.free()

    def solve(self, reference, deformed, point, initial) -> np.void:
        return self.solve_batch(reference, deformed, np.asarray(point, np.float64).reshape(1, 2),
                                np.atleast_1d(np.asarray(initial, INITIAL_DTYPE)))[0]

    def solve_with_interpolators(self, reference, deformed, point, initial, pair: PreparedPair) -> np.void:
        """subset_solver.hpp:41-51: the pair's gradient cache and coefficients
        play the reference and deformed interpolators."""
        return self.solve_batch(reference, deformed, np.asarray(point, np.float64).reshape(1, 2),
                                np.atleast_1d(np.asarray(initial, INITIAL_DTYPE)), pair=pair)[0]

    def solve_with_mask(self, reference, deformed, roi, point, initial, pair: PreparedPair | None = None) -> np.void:
        return self.solve_batch(reference, deformed, np.asarray(point, np.float64).reshape(1, 2),
                                np.atleast_1d(np.asarray(initial, INITIAL_DTYPE)), roi=roi, pair=pair)[0]


def _image_size(img) -> tuple[int, int]:
    """(width, height) of a host array or a DeviceImage."""
    if isinstance(img, _lib.DeviceImage):
        return img.width, img.height
    a = np.asarray(img)
    if a.ndim != 2:
        raise ValueError("Image must be a 2D numpy array.")
    return a.shape[1], a.shape[0]


def invalid_results(xy: np.ndarray, ini: np.ndarray) -> np.ndarray:
    """Every POI InvalidInput with the initial u, v and correlation = confidence
    (the reference's early exits, icgn.cpp:455-470, :639-652)."""
    out = np.zeros(xy.shape[0], RESULT_DTYPE)
    out["x"], out["y"] = xy[:, 0], xy[:, 1]
    out["p"][:, 0], out["p"][:, 1] = ini["u"], ini["v"]
    out["correlation"] = ini["confidence"]
    out["status"] = _lib.INVALID_INPUT
    return out


class ForwardGaussNewtonSolver(ICGNSolver):
    """ForwardGaussNewtonSolver (include/dic/subset/solver/forward_gauss_newton.hpp:11-107)
    on the GPU, same call pattern as ICGNSolver: the deformed image's B-spline
    gradient at the warped points drives the steepest descent, H is rebuilt
    every iteration, dp = +H^-1 G is capped and composed forward."""

    def __init__(self, config: SubsetConfig | dict | None = None, precision: str = "fp64",
                 gradient: str = "central", device: int = 0, context: _lib.Context | None = None):
        super().__init__(config, precision, gradient, device, context)
        self.config = dataclasses.replace(self.config, optimizer="forward_gauss_newton")


def make_solver(config: SubsetConfig | dict | None = None, precision: str = "fp64", **kw) -> ICGNSolver:
    """The solver the configuration names (reliability_propagation.cpp:191-196)."""
    cfg = config if isinstance(config, SubsetConfig) else config_from_dict(config)
    if cfg.optimizer == "forward_gauss_newton":
        return ForwardGaussNewtonSolver(cfg, precision, **kw)
    return ICGNSolver(cfg, precision, **kw)


def propagate(reference: np.ndarray, deformed: np.ndarray, roi: np.ndarray, seeds: np.ndarray,
              config: SubsetConfig | dict | None = None, precision: str = "fp64",
              solver: ICGNSolver | None = None, pair: PreparedPair | None = None):
    """ReliabilityPropagation::propagate (reliability_propagation.cpp:130-394) on
    the GPU from caller-supplied seeds (SEED_DTYPE: x, y, displacement).
    Returns (results_to_dict of the grid, grid results (gh, gw), points_computed).
    pair: a PreparedPair of (reference, deformed) to reuse (otherwise one is
    prepared for this call and freed).  Images of different sizes, an ROI of
    another size, a subset radius below 1, a negative spacing or no seeds give
    the reference's empty result (:140-146)."""
    solver = solver or make_solver(config, precision)
    cfg = solver.config
    rw, rh = _image_size(reference)
    dw, dh = _image_size(deformed)
    roi = np.asarray(roi)
    if ((rw, rh) != (dw, dh) or rw <= 0 or rh <= 0 or roi.shape != (rh, rw) or cfg.subset_radius < 1
            or cfg.propagation_spacing < 0 or len(seeds) == 0):
        empty = np.zeros((0, 0), RESULT_DTYPE)
        return results_to_dict(empty.ravel()), empty, 0
    owned = pair is None
    if owned:
        pair = solver.prepare(reference, deformed)
    elif (pair.width, pair.height) != (rw, rh):
        raise ValueError(f"prepared pair is {pair.width}x{pair.height}, the images are {rw}x{rh}")
    try:
        params = make_params(cfg, solver.precision, masked=bool(cfg.truncate_roi_subsets))
        grid, n = solver.ctx.propagate(params, cfg.propagation_spacing, cfg.propagation_max_znssd, pair.ref,
                                       pair.grad, pair.coeffs, np.ascontiguousarray(roi, np.uint8), seeds,
                                       order=cfg.propagation_order)
    finally:
        if owned:
            pair.free()
    return results_to_dict(grid.ravel()), grid, n


def results_to_dict(results: np.ndarray) -> dict:
    """Same keys as bind_subset.cpp:102-153 (first-order parameters only)."""
    return {
        "x": results["x"].copy(), "y": results["y"].copy(),
        "u": results["p"][:, 0].copy(), "v": results["p"][:, 1].copy(),
        "du_dx": results["p"][:, 2].copy(), "du_dy": results["p"][:, 3].copy(),
        "dv_dx": results["p"][:, 4].copy(), "dv_dy": results["p"][:, 5].copy(),
        "correlation": results["correlation"].copy(), "valid": results["valid"].astype(bool),
    }


def compute_grid(reference: np.ndarray, deformed: np.ndarray, initial_u: np.ndarray, initial_v: np.ndarray,
                 config: dict | SubsetConfig | None = None, precision: str = "fp64",
                 solver: ICGNSolver | None = None) -> dict:
    """Grid solve with caller-supplied initial guesses on the propagation grid
    (step = spacing + 1, row-major, reliability_propagation.cpp:154-176).
    POIs whose subset crosses the image margin keep the reference defaults
    (valid = False, u = v = 0; :318-323).  initial_u / initial_v have one
    entry per grid point."""
    solver = solver or ICGNSolver(config, precision)
    cfg = solver.config
    H, W = reference.shape
    step = cfg.propagation_spacing + 1
    gw = (W + step - 1) // step
    gh = (H + step - 1) // step
    i = np.arange(gw * gh)
    xy = np.stack([(i % gw) * step, (i // gw) * step], axis=1).astype(np.float64)
    r = cfg.subset_radius
    inside = (xy[:, 0] - r >= 0) & (xy[:, 1] - r >= 0) & (xy[:, 0] + r < W) & (xy[:, 1] + r < H)
    res = np.zeros(xy.shape[0], RESULT_DTYPE)
    res["x"], res["y"] = xy[:, 0], xy[:, 1]
    res["status"] = _lib.INVALID_INPUT
    idx = np.nonzero(inside)[0]
    if idx.size:
        ini = initial_guesses(np.asarray(initial_u, np.float64).ravel()[idx],
                              np.asarray(initial_v, np.float64).ravel()[idx])
        res[idx] = solver.solve_batch(reference, deformed, xy[idx], ini)
    return results_to_dict(res)
