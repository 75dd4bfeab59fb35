"""Seeding and the Subset-DIC pipeline on the HIP library.

Reference interfaces mirrored (paths inside the Traditional-DIC tree):
  IntegerSearchInitializer     include/dic/initialization/integer_search.hpp:32-71,
                               src/initialization/integer_search.cpp:303-645
  SeedSelector                 include/dic/subset/seed/seed_selector.hpp:20-64,
                               src/subset/seed/seed_selector.cpp:146-539
  SubsetDIC::compute           src/subset/subset_dic.cpp:24-105
  recommended_subset_padding   src/subset/padding.cpp:28-36
  subset() (Python facade)     python/traditional_dic/subset.py:16-164

The search, candidate generation, refinement and propagation run in
libdic_hip.so; this module assembles parameters and records.
"""
from __future__ import annotations

import warnings
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from . import _lib
from ._lib import ESTIMATE_DTYPE, RESULT_DTYPE, SEED_DTYPE, SEED_EVAL_DTYPE
from .subset import (SeedInitializationConfig, SubsetConfig, config_from_dict, make_params, PreparedPair,
                     results_to_dict)

_QUALITY = {"zncc": _lib.QUALITY_ZNCC, "znssd": _lib.QUALITY_ZNSSD, "ssd": _lib.QUALITY_SSD}


def _config(config) -> SubsetConfig:
    return config if isinstance(config, SubsetConfig) else config_from_dict(config)


def search_params(init: SeedInitializationConfig, mode: int = _lib.SEARCH_GLOBAL,
                  masked: bool = False) -> _lib.IntegerSearchParams:
    """IntegerSearchInitializer's constructor clamps (integer_search.cpp:293-301)."""
    c = init.integer_search
    p = _lib.IntegerSearchParams()
    p.subset_radius = max(1, int(c.subset_radius))
    p.search_radius = max(0, int(c.search_radius))
    p.pyramid_enabled = 1 if c.pyramid_enabled else 0
    p.pyramid_scale = max(1, int(c.pyramid_scale))
    p.pyramid_refinement_radius = max(1, int(c.pyramid_refinement_radius))
    p.mode = int(mode)
    p.masked = 1 if masked else 0
    return p


def seed_params(cfg: SubsetConfig) -> _lib.SeedParams:
    """SeedSelector / SubsetInitializer configuration (seed_selector.cpp:138-144,
    subset_initializer.cpp:28-39)."""
    si, sel = cfg.seed_initialization, cfg.seed_selection
    p = _lib.SeedParams()
    p.search = search_params(si)
    sp = si.subpixel
    p.subpixel_enabled = 1 if sp.enabled else 0
    p.subpixel_shape_order = 2 if sp.shape_function in ("second_order", "2", 2) else 1
    p.subpixel_optimizer = _lib.OPTIMIZER_FGN if sp.optimizer in ("forward_gauss_newton", "fgn") \
        else _lib.OPTIMIZER_ICGN
    p.subpixel_objective = _lib.OBJECTIVE_SSD if sp.objective == "ssd" else _lib.OBJECTIVE_ZNSSD
    p.subpixel_radius = max(1, int(sp.subset_radius))
    p.subpixel_max_iterations = max(0, int(sp.max_iterations))
    p.subpixel_convergence_threshold = float(sp.convergence_threshold)
    p.seed_count = max(1, int(sel.seed_count))
    p.quality_metric = _QUALITY[sel.quality_metric]
    p.min_zncc, p.max_znssd, p.max_ssd = float(sel.min_zncc), float(sel.max_znssd), float(sel.max_ssd)
    p.min_displacement_norm = float(sel.min_displacement_norm)
    p.min_texture_std = float(sel.min_texture_std)
    p.kmeans_iterations = max(0, int(sel.kmeans_iterations))
    p.kmeans_sample_limit = max(1, int(sel.kmeans_sample_limit))
    p.truncate_roi_subsets = 1 if cfg.truncate_roi_subsets else 0
    p.min_valid_samples = int(cfg.min_valid_samples)
    p.min_valid_sample_ratio = float(cfg.min_valid_sample_ratio)
    p.degree = int(cfg.image_precompute.degree)
    p.border = int(cfg.image_precompute.border)
    return p


class _Images:
    """Host arrays uploaded for one call, or DeviceImage handles passed through."""

    def __init__(self, ctx: _lib.Context, *imgs):
        self.owned = []
        self.handles = []
        for im in imgs:
            if isinstance(im, _lib.DeviceImage):
                self.handles.append(im)
            else:
                h = ctx.upload_image(np.ascontiguousarray(im, np.float32))
                self.owned.append(h)
                self.handles.append(h)

    def free(self):
        for h in self.owned:
            h.free()


def _shape(img):
    return (img.height, img.width) if isinstance(img, _lib.DeviceImage) else np.shape(img)


class IntegerSearchInitializer:
    """IntegerSearchInitializer on the GPU; each method takes many points at once
    (points: (n, 2) x, y) and returns ESTIMATE_DTYPE records (InitialDisplacement
    with valid, zncc, znssd)."""

    def __init__(self, config: SeedInitializationConfig | SubsetConfig | dict | None = None, device: int = 0,
                 context: _lib.Context | None = None):
        if isinstance(config, SeedInitializationConfig):
            self.config = config
        else:
            self.config = _config(config).seed_initialization
        self.ctx = context if context is not None else _lib.Context(device)

    def _run(self, reference, deformed, points, mode, roi=None, prior=None):
        pts = np.asarray(points, np.float64).reshape(-1, 2)
        if _shape(reference) != _shape(deformed):
            return np.zeros(pts.shape[0], ESTIMATE_DTYPE)  # :309-314
        imgs = _Images(self.ctx, reference, deformed)
        try:
            return self.ctx.integer_search(search_params(self.config, mode, roi is not None), imgs.handles[0],
                                           imgs.handles[1], pts, prior, roi)
        finally:
            imgs.free()

    def estimate(self, reference, deformed, points):
        return self._run(reference, deformed, points, _lib.SEARCH_GLOBAL)

    estimate_with_interpolators = estimate

    def estimate_with_mask(self, reference, deformed, roi, points):
        return self._run(reference, deformed, points, _lib.SEARCH_GLOBAL, roi=np.asarray(roi, np.uint8))

    def estimate_around_displacement(self, reference, deformed, points, initial_u, initial_v):
        prior = np.stack([np.broadcast_to(np.asarray(initial_u, np.float64), (len(points),)),
                          np.broadcast_to(np.asarray(initial_v, np.float64), (len(points),))], 1)
        return self._run(reference, deformed, points, _lib.SEARCH_AROUND_PRIOR, prior=prior)

    def estimate_with_mask_around_displacement(self, reference, deformed, roi, points, initial_u, initial_v):
        if roi is None or np.size(roi) == 0:  # :590-592
            return self.estimate_around_displacement(reference, deformed, points, initial_u, initial_v)
        prior = np.stack([np.broadcast_to(np.asarray(initial_u, np.float64), (len(points),)),
                          np.broadcast_to(np.asarray(initial_v, np.float64), (len(points),))], 1)
        return self._run(reference, deformed, points, _lib.SEARCH_AROUND_PRIOR, roi=np.asarray(roi, np.uint8),
                         prior=prior)


@dataclass
class SeedSelectionResult:
    """seed_selector.hpp:29-33."""
    candidates: np.ndarray = field(default_factory=lambda: np.zeros(0, SEED_EVAL_DTYPE))
    best_seed: np.void | None = None
    found: bool = False


class SeedSelector:
    """SeedSelector (integer-search path) on the GPU."""

    def __init__(self, config: SubsetConfig | dict | None = None, device: int = 0,
                 context: _lib.Context | None = None):
        self.config = _config(config)
        if self.config.seed_initialization.integer_search.sift_enabled:
            warnings.warn("SIFT seeding (OpenCV) is not part of this library; the integer-pixel ZNCC search "
                          "seeds instead", RuntimeWarning, stacklevel=2)
        self.ctx = context if context is not None else _lib.Context(device)

    def generate_candidates(self, reference, roi) -> np.ndarray:
        imgs = _Images(self.ctx, reference)
        try:
            return self.ctx.seed_candidates(seed_params(self.config), imgs.handles[0], np.asarray(roi, np.uint8))
        finally:
            imgs.free()

    def select_best_seed(self, reference, deformed, roi) -> SeedSelectionResult:
        if _shape(reference) != _shape(deformed) or np.shape(roi) != _shape(reference):
            return SeedSelectionResult()  # :151-157
        imgs = _Images(self.ctx, reference, deformed)
        try:
            ev, best = self.ctx.select_seed(seed_params(self.config), imgs.handles[0], imgs.handles[1],
                                            np.asarray(roi, np.uint8))
        finally:
            imgs.free()
        return SeedSelectionResult(ev, ev[best] if best >= 0 else None, best >= 0)


def recommended_subset_padding(cfg: SubsetConfig) -> int:
    """padding.cpp:28-36."""
    si = cfg.seed_initialization
    radius = max(cfg.subset_radius, max(1, si.integer_search.subset_radius), max(1, si.subpixel.subset_radius))
    return max(0, si.integer_search.search_radius) + radius + max(0, cfg.image_precompute.border)


class SubsetDIC:
    """SubsetDIC::compute (subset_dic.cpp:24-105): seed selection, then
    reliability-guided propagation from the best seed, on the GPU.  Returns
    every grid point inside the image (valid or not) as RESULT_DTYPE records;
    an empty array where the reference returns an empty vector."""

    def __init__(self, config: SubsetConfig | dict | None = None, precision: str = "fp64", device: int = 0,
                 context: _lib.Context | None = None):
        self.config = _config(config)
        if precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        self.precision = _lib.PRECISION_FP64 if precision == "fp64" else _lib.PRECISION_MIXED
        self.ctx = context if context is not None else _lib.Context(device)
        self.selector = SeedSelector(self.config, context=self.ctx)
        self.last_seed: SeedSelectionResult | None = None
        self.points_computed = 0

    def compute(self, reference: np.ndarray, deformed: np.ndarray, roi: np.ndarray | None = None) -> np.ndarray:
        empty = np.zeros(0, RESULT_DTYPE)
        ref = np.ascontiguousarray(reference, np.float32)
        dfm = np.ascontiguousarray(deformed, np.float32)
        if ref.size == 0 or ref.shape != dfm.shape:
            return empty
        H, W = ref.shape
        mask = np.ones((H, W), np.uint8) if roi is None else np.ascontiguousarray(roi, np.uint8)
        if mask.shape != (H, W):
            return empty
        cfg = self.config
        pad = recommended_subset_padding(cfg) if cfg.truncate_roi_subsets else 0
        r0, d0 = self.ctx.upload_image(ref), self.ctx.upload_image(dfm)
        handles = [r0, d0]
        try:
            if pad > 0:
                rp, dp = self.ctx.mirror_pad(r0, pad), self.ctx.mirror_pad(d0, pad)
                handles += [rp, dp]
                mask = np.pad(mask, pad)  # zero_pad_mask, padding.cpp:61-87
            else:
                rp, dp = r0, d0
            seed = self.selector.select_best_seed(rp, dp, mask)
            self.last_seed = seed
            if not seed.found:
                return empty
            b = seed.best_seed
            sd = np.zeros(1, SEED_DTYPE)
            sd["x"], sd["y"] = b["x"], b["y"]
            d = sd["displacement"]
            for j, k in enumerate(("u", "v", "du_dx", "du_dy", "dv_dx", "dv_dy")):
                d[k] = b["displacement"]["p"][j]
            d["confidence"] = b["displacement"]["correlation"]
            pair = PreparedPair(self.ctx, rp, dp, cfg, self.precision)
            try:
                params = make_params(cfg, self.precision, masked=bool(cfg.truncate_roi_subsets))
                grid, n = self.ctx.propagate(params, cfg.propagation_spacing, cfg.propagation_max_znssd, pair.ref,
                                             pair.grad, pair.coeffs, mask, sd, order=cfg.propagation_order)
            finally:
                pair.free()
        finally:
            for h in handles:
                h.free()
        self.points_computed = n
        if n == 0:
            return empty
        pts = grid.ravel().copy()
        pts["x"] -= pad
        pts["y"] -= pad
        keep = (pts["x"] >= 0) & (pts["y"] >= 0) & (pts["x"] < W) & (pts["y"] < H)
        return pts[keep]


def subset(reference: np.ndarray, deformed: np.ndarray, config: str | Path | dict | None = None,
           roi: np.ndarray | None = None, *, radius: int = 15, max_iterations: int = 30,
           convergence_threshold: float = 1e-3, bspline_degree: int = 5, search_radius: int = 20,
           seed_subset_radius: int = 10, sift_enabled: bool = False, pyramid_enabled: bool = True,
           subpixel_refinement: bool = True, seed_count: int = 64, max_znssd: float = 0.3,
           min_displacement_norm: float = 0.0, min_texture_std: float = 0.02, propagation_spacing: int = 5,
           propagation_max_znssd: float = 0.5, precision: str = "fp64", device: int = 0,
           propagation_order: str = "heap") -> dict:
    """Run 2D Subset-DIC with reliability-guided propagation
    (python/traditional_dic/subset.py:16-164, same arguments and defaults):
    a YAML path or dict config takes precedence over the keyword arguments.
    precision, device and propagation_order ("heap": the reference's walk
    order; "frontier": whole frontiers per launch) are this build's own.
    Returns the dict of bind_subset.cpp:102-153 (x, y, u, v, du_dx, du_dy,
    dv_dx, dv_dy, correlation, valid)."""
    if not isinstance(reference, np.ndarray) or reference.ndim != 2:
        raise ValueError("reference must be a 2D numpy array")
    if not isinstance(deformed, np.ndarray) or deformed.ndim != 2:
        raise ValueError("deformed must be a 2D numpy array")
    if reference.shape != deformed.shape:
        raise ValueError("reference and deformed must have the same shape")
    if config is not None:
        if isinstance(config, (str, Path)):
            import yaml
            with open(config) as f:
                config_dict = yaml.safe_load(f) or {}
        elif isinstance(config, dict):
            config_dict = dict(config)
        else:
            raise TypeError(f"config must be str, Path, dict, or None, got {type(config)}")
    else:
        config_dict = {
            "subset": {"radius": radius},
            "optimization": {"max_iterations": max_iterations, "convergence_threshold": convergence_threshold},
            "interpolation": {"degree": bspline_degree},
            "initialization": {
                "method": "integer_search",
                "integer_search": {"subset_radius": seed_subset_radius, "search_radius": search_radius,
                                   "sift_enabled": sift_enabled, "pyramid_enabled": pyramid_enabled},
                "subpixel_refinement": {"enabled": subpixel_refinement, "subset_radius": radius,
                                        "max_iterations": max_iterations,
                                        "convergence_threshold": convergence_threshold},
            },
            "seed_selection": {"seed_count": seed_count, "max_znssd": max_znssd,
                               "min_displacement_norm": min_displacement_norm, "min_texture_std": min_texture_std},
            "reliability_propagation": {"spacing": propagation_spacing, "max_znssd": propagation_max_znssd,
                                        "order": propagation_order},
        }
    if roi is not None and (not isinstance(roi, np.ndarray) or roi.dtype not in (np.bool_, np.uint8)):
        raise RuntimeError("roi must be a 2D numpy array (bool or uint8) or None.")
    dic = SubsetDIC(config_from_dict(config_dict), precision=precision, device=device)
    try:
        return results_to_dict(dic.compute(reference, deformed, roi))
    finally:
        dic.ctx.close()
