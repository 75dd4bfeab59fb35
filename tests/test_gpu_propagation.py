"""Reliability-guided propagation on the GPU (dic_propagate) against the
reference's golden fields of the stereo-plate case (radius 41, quintic,
first order, 50 iterations, ||dp|| < 1e-6, spacing 3, max ZNSSD 2.0, the
reference's ROI), seeded at one golden POI: the valid set must equal the
golden one exactly; u, v within 1e-5 px for >= 97% of the POIs and within
1e-4 px for all (the golden run used other seeds, so guesses come from other
parents; see tests/test_oracle_propagation.py).  On a window the GPU walk is
also compared with the oracle's best-first walk from the same seed."""
import numpy as np
import pytest

from conftest import gradient_init, read_gray
from test_oracle_propagation import crop_case

pytestmark = pytest.mark.gpu


def _report(msg):
    print(msg, flush=True)


def _cfg(pkg):
    return pkg.config_from_dict({
        "subset": {"radius": 41}, "shape_function": {"order": 1},
        "optimization": {"method": "icgn", "max_iterations": 50, "convergence_threshold": 1e-6},
        "correlation": {"criterion": "znssd"}, "interpolation": {"degree": 5},
        "reliability_propagation": {"spacing": 3, "max_znssd": 2.0}})


def _seed(pkg, x, y, u, v, g, corr):
    s = np.zeros(1, pkg.SEED_DTYPE)
    s["x"], s["y"] = x, y
    d = s["displacement"]
    d["u"], d["v"], d["du_dx"], d["du_dy"], d["dv_dx"], d["dv_dy"], d["confidence"] = u, v, *g, corr
    return s


def _compare(got_grid, golden):
    flat = got_grid.ravel()
    valid = flat[flat["valid"] == 1]
    got = {(int(r["x"]), int(r["y"])): r for r in valid}
    want = {(int(r[0]), int(r[1])): r for r in golden}
    assert set(got) == set(want), (len(got), len(want), len(set(got) ^ set(want)))
    keys = list(want)
    du = np.array([max(abs(got[k]["p"][0] - want[k][2]), abs(got[k]["p"][1] - want[k][3])) for k in keys])
    return du, keys


# The golden value at this POI comes from the golden run's seed, solved on the
# seed path's non-prefiltered interpolant (seed_selector.cpp:166-173), not by
# propagation; every solver here lands ~3e-3 px away from it.
GOLDEN_SEED_ARTIFACTS = {"left_temporal": {(852, 624)}}


@pytest.mark.parametrize("name", ["left_temporal", "deformed_disparity", "reference_disparity"])
def test_gpu_propagation_reproduces_golden_field(pkg, ctx, stereo, name):
    rows = stereo["field_" + name]
    a, b = stereo["pair_" + name]
    ref, dfm = read_gray(stereo["img_" + a]), read_gray(stereo["img_" + b])
    g = gradient_init(rows)
    c = rows[:, :2].mean(0)
    k = int(np.argmin(np.hypot(rows[:, 0] - c[0], rows[:, 1] - c[1])))
    seed = _seed(pkg, rows[k, 0], rows[k, 1], rows[k, 2], rows[k, 3], g[k], rows[k, 4])
    solver = pkg.ICGNSolver(_cfg(pkg), precision="fp64", context=ctx)
    fields, grid, n = pkg.propagate(ref, dfm, stereo["roi"], seed, solver=solver)
    assert grid.shape == (270, 360) and n == rows.shape[0]
    du, keys = _compare(grid, rows)
    far = {keys[i] for i in np.nonzero(du > 1e-4)[0]}
    _report(f"gpu propagation {name}: {n} valid = golden, outside 1e-5: {(du > 1e-5).sum()}, "
            f"beyond 1e-4: {sorted(far)} (max {du.max():.2e})")
    assert (du > 1e-5).sum() <= 0.03 * du.size
    assert far <= GOLDEN_SEED_ARTIFACTS.get(name, set())
    assert fields["valid"].sum() == n and np.all(fields["u"][~fields["valid"]] == 0)


def test_gpu_propagation_matches_oracle_window(pkg, ctx, orc, stereo):
    ref, dfm, roi, seed_row, golden = crop_case(stereo, "left_temporal")
    pre = orc.Precomputed(dfm, degree=5)
    gx, gy = orc.reference_gradient_cache(ref)
    cpu, ncpu = orc.propagate(ref, pre, gx, gy, roi, seed_row, radius=41, max_iterations=50, threshold=1e-6,
                              spacing=3, max_znssd=2.0)
    s = seed_row[0]
    seed = _seed(pkg, s[0], s[1], s[2], s[3], s[4:8], s[8])
    solver = pkg.ICGNSolver(_cfg(pkg), precision="fp64", context=ctx)
    _, gpu, ngpu = pkg.propagate(ref, dfm, roi, seed, solver=solver)
    assert ngpu == ncpu
    assert np.array_equal(gpu["valid"], cpu["valid"])
    v = cpu["valid"] == 1
    d = np.abs(gpu["p"][v][:, :2] - cpu["p"][v][:, :2]).max(axis=1)  # noqa: E501
    _report(f"gpu vs oracle propagation: {ngpu} valid each, max |du| {d.max():.2e}, outside 1e-5 {(d > 1e-5).sum()}")
    assert (d > 1e-5).sum() <= 0.03 * d.size and d.max() <= 1e-4


@pytest.mark.parametrize("method,order,criterion", [("icgn", 2, "ssd"), ("forward_gauss_newton", 1, "znssd")])
def test_gpu_propagation_other_solvers_match_oracle_window(pkg, ctx, orc, stereo, method, order, criterion):
    """The walk over the other solver variants the reference's factory can
    build (reliability_propagation.cpp:191-196): second-order SSD IC-GN and
    first-order ZNSSD FGN, radius 21, on the same window and seed as the
    oracle's best-first walk.  The acceptance limit (max ZNSSD 2.0) accepts
    every solved point, so the valid sets must agree; values may come from
    other parents than in the oracle's walk."""
    ref, dfm, roi, seed_row, _ = crop_case(stereo, "left_temporal")
    pre = orc.Precomputed(dfm, degree=5)
    gx, gy = orc.reference_gradient_cache(ref)
    cpu, ncpu = orc.propagate(ref, pre, gx, gy, roi, seed_row, radius=21, max_iterations=50, threshold=1e-6,
                              shape_order=order, objective=1 if criterion == "ssd" else 0,
                              optimizer=1 if method != "icgn" else 0, spacing=3, max_znssd=2.0)
    s = seed_row[0]
    seed = _seed(pkg, s[0], s[1], s[2], s[3], s[4:8], s[8])
    cfg = pkg.config_from_dict({
        "subset": {"radius": 21}, "shape_function": {"order": order},
        "optimization": {"method": method, "max_iterations": 50, "convergence_threshold": 1e-6},
        "correlation": {"criterion": criterion}, "interpolation": {"degree": 5},
        "reliability_propagation": {"spacing": 3, "max_znssd": 2.0}})
    solver = pkg.make_solver(cfg, "fp64", context=ctx)
    _, gpu, ngpu = pkg.propagate(ref, dfm, roi, seed, solver=solver)
    v = cpu["valid"] == 1
    d = np.abs(gpu["p"][v][:, :2] - cpu["p"][v][:, :2]).max(axis=1) if v.any() else np.zeros(1)
    _report(f"{method} order {order} {criterion}: gpu {ngpu} / oracle {ncpu} points, valid {int(v.sum())}, "
            f"max |du| {d.max():.2e}, outside 1e-5 {(d > 1e-5).sum()}")
    assert ngpu == ncpu and v.sum() > 1000
    assert np.array_equal(gpu["valid"], cpu["valid"])
    assert (d > 1e-5).sum() <= 0.03 * d.size and d.max() <= 1e-4


def test_gpu_propagation_mixed_precision_follows_fp64(pkg, ctx, stereo):
    """The golden configuration walked in mixed precision: the same valid set as
    the FP64 walk from the same seed and u, v within 1e-3 px of it."""
    ref, dfm, roi, seed_row, _ = crop_case(stereo, "left_temporal")
    s = seed_row[0]
    seed = _seed(pkg, s[0], s[1], s[2], s[3], s[4:8], s[8])
    out = {}
    for prec in ("fp64", "mixed"):
        solver = pkg.ICGNSolver(_cfg(pkg), precision=prec, context=ctx)
        _, out[prec], n = pkg.propagate(ref, dfm, roi, seed, solver=solver)
    v = out["fp64"]["valid"] == 1
    d = np.abs(out["mixed"]["p"][v][:, :2] - out["fp64"]["p"][v][:, :2]).max(axis=1)
    _report(f"mixed vs fp64 walk: {int(v.sum())} valid, max |du| {d.max():.2e}")
    assert v.sum() > 1000 and np.array_equal(out["mixed"]["valid"], out["fp64"]["valid"])
    assert d.max() <= 1e-3


def test_gpu_propagation_strict_threshold_and_bad_seed(pkg, ctx, stereo):
    rows = stereo["field_left_temporal"]
    ref, dfm = read_gray(stereo["img_00_L"]), read_gray(stereo["img_04_L"])
    cfg = _cfg(pkg)
    cfg.propagation_max_znssd = 0.0  # nothing passes correlation < 0
    solver = pkg.ICGNSolver(cfg, precision="fp64", context=ctx)
    k = 100
    seed = _seed(pkg, rows[k, 0], rows[k, 1], rows[k, 2], rows[k, 3], (0, 0, 0, 0), rows[k, 4])
    _, grid, n = pkg.propagate(ref, dfm, stereo["roi"], seed, solver=solver)
    assert n == 1 and grid["valid"].sum() == 1
    outside = _seed(pkg, 5.0, 5.0, 0.0, 0.0, (0, 0, 0, 0), 0.0)  # outside the ROI
    _, grid, n = pkg.propagate(ref, dfm, stereo["roi"], outside, solver=solver)
    assert n == 0 and grid["valid"].sum() == 0 and np.all(grid["status"] == 2)


def _window_walks(pkg, ctx, orc, stereo, max_znssd, order):
    ref, dfm, roi, seed_row, golden = crop_case(stereo, "left_temporal")
    pre = orc.Precomputed(dfm, degree=5)
    gx, gy = orc.reference_gradient_cache(ref)
    cpu, ncpu = orc.propagate(ref, pre, gx, gy, roi, seed_row, radius=41, max_iterations=50, threshold=1e-6,
                              spacing=3, max_znssd=max_znssd)
    s = seed_row[0]
    seed = _seed(pkg, s[0], s[1], s[2], s[3], s[4:8], s[8])
    cfg = _cfg(pkg)
    cfg.propagation_max_znssd = max_znssd
    cfg.propagation_order = order
    solver = pkg.ICGNSolver(cfg, precision="fp64", context=ctx)
    _, gpu, ngpu = pkg.propagate(ref, dfm, roi, seed, solver=solver)
    return cpu, ncpu, gpu, ngpu, len(golden)


@pytest.mark.parametrize("max_znssd", [0.5, 0.05, 0.0082, 0.0077])
def test_gpu_heap_order_walk_equals_oracle_walk(pkg, ctx, orc, stereo, max_znssd):
    """The heap order replays the reference's priority-queue walk
    (reliability_propagation.cpp:277-389): at any acceptance limit, including
    limits that reject part of the window (its golden ZNSSD spans 0.006 to
    0.0098, median 0.0077), every point is solved from the parent the oracle's best-first walk
    uses, so the valid set, the count and u, v (1e-5 px) equal the oracle's."""
    cpu, ncpu, gpu, ngpu, lenient = _window_walks(pkg, ctx, orc, stereo, max_znssd, "heap")
    v = cpu["valid"] == 1
    d = np.abs(gpu["p"][v][:, :2] - cpu["p"][v][:, :2]).max(axis=1) if v.any() else np.zeros(1)
    _report(f"heap walk, max ZNSSD {max_znssd}: gpu {ngpu} / oracle {ncpu} valid ({lenient} at the golden "
            f"limit), max |du| {d.max():.2e}")
    assert ngpu == ncpu and np.array_equal(gpu["valid"], cpu["valid"])
    assert d.max() <= 1e-5
    if max_znssd < 0.01:
        assert ncpu < lenient  # the limit rejects points, so the walk order decides the valid set


def test_gpu_frontier_order_accepts_what_heap_order_accepts_at_lenient_limit(pkg, ctx, orc, stereo):
    """The frontier order at the golden limit (2.0: every solved point passes)
    gives the heap order's valid set; values may come from other parents."""
    _, _, heap, nh, _ = _window_walks(pkg, ctx, orc, stereo, 2.0, "heap")
    _, _, front, nf, _ = _window_walks(pkg, ctx, orc, stereo, 2.0, "frontier")
    assert nh == nf and np.array_equal(heap["valid"], front["valid"])
    v = heap["valid"] == 1
    assert np.abs(heap["p"][v][:, :2] - front["p"][v][:, :2]).max() <= 1e-4


@pytest.mark.parametrize("order,criterion,masked,limit", [(2, "znssd", False, 0.0082), (1, "ssd", False, 1.3),
                                                          (1, "znssd", True, 0.0082), (2, "ssd", True, 1.3)])
def test_gpu_heap_order_walk_other_variants_equal_oracle_walk(pkg, ctx, orc, stereo, order, criterion, masked,
                                                              limit):
    """The heap-order walk solves a few POIs per launch, each on a workgroup of
    several waves that share its samples; over the solver variants the
    reference's factory builds (first / second order, ZNSSD / SSD, with and
    without ROI-truncated subsets), at acceptance limits that reject part of
    the window (near the median correlation: ZNSSD 0.0082, SSD 1.3), the valid
    set, the count and u, v (1e-5 px) equal the oracle's best-first walk."""
    ref, dfm, roi, seed_row, _ = crop_case(stereo, "left_temporal")
    pre = orc.Precomputed(dfm, degree=5)
    gx, gy = orc.reference_gradient_cache(ref)
    obj = 1 if criterion == "ssd" else 0
    cpu, ncpu = orc.propagate(ref, pre, gx, gy, roi, seed_row, radius=41, max_iterations=50, threshold=1e-6,
                              shape_order=order, objective=obj, spacing=3, max_znssd=limit, truncate_roi=masked)
    s = seed_row[0]
    seed = _seed(pkg, s[0], s[1], s[2], s[3], s[4:8], s[8])
    cfg = pkg.config_from_dict({
        "subset": {"radius": 41}, "shape_function": {"order": order},
        "optimization": {"method": "icgn", "max_iterations": 50, "convergence_threshold": 1e-6},
        "correlation": {"criterion": criterion}, "interpolation": {"degree": 5},
        "reliability_propagation": {"spacing": 3, "max_znssd": limit}})
    cfg.truncate_roi_subsets = masked
    solver = pkg.ICGNSolver(cfg, precision="fp64", context=ctx)
    _, gpu, ngpu = pkg.propagate(ref, dfm, roi, seed, solver=solver)
    v = cpu["valid"] == 1
    d = np.abs(gpu["p"][v][:, :2] - cpu["p"][v][:, :2]).max(axis=1) if v.any() else np.zeros(1)
    _report(f"heap walk order {order} {criterion} masked={masked}: gpu {ngpu} / oracle {ncpu} valid, "
            f"max |du| {d.max():.2e}")
    assert ngpu == ncpu and ncpu > 100 and np.array_equal(gpu["valid"], cpu["valid"])
    assert d.max() <= 1e-5


def test_gpu_heap_order_walk_second_order_mixed_follows_fp64(pkg, ctx, orc, stereo):
    """Second-order ZNSSD walks in mixed precision and FP64 at the golden
    limit (2.0, every solved point accepted): the same valid set, u, v within
    1e-3 px."""
    ref, dfm, roi, seed_row, _ = crop_case(stereo, "left_temporal")
    s = seed_row[0]
    seed = _seed(pkg, s[0], s[1], s[2], s[3], s[4:8], s[8])
    cfg = pkg.config_from_dict({
        "subset": {"radius": 41}, "shape_function": {"order": 2},
        "optimization": {"method": "icgn", "max_iterations": 50, "convergence_threshold": 1e-6},
        "interpolation": {"degree": 5}, "reliability_propagation": {"spacing": 3, "max_znssd": 2.0}})
    out = {}
    for prec in ("fp64", "mixed"):
        solver = pkg.ICGNSolver(cfg, precision=prec, context=ctx)
        _, out[prec], n = pkg.propagate(ref, dfm, roi, seed, solver=solver)
    v = out["fp64"]["valid"] == 1
    d = np.abs(out["mixed"]["p"][v][:, :2] - out["fp64"]["p"][v][:, :2]).max(axis=1)
    _report(f"second-order heap walk mixed vs fp64: {int(v.sum())} valid, max |du| {d.max():.2e}")
    assert v.sum() > 1000 and np.array_equal(out["mixed"]["valid"], out["fp64"]["valid"])
    assert d.max() <= 1e-3


def test_gpu_propagation_at_the_largest_fp64_radius(pkg, ctx, orc, synth):
    """FP64 ZNSSD IC-GN at radius 79, the largest subset whose per-sample
    values fit one workgroup's LDS beside the one-wave kernel's block.  The
    walk's launches are small, but the cooperative kernel's per-wave slots do
    not fit beside those values, so they keep one wave per POI; the walk
    equals the oracle's (valid set, count, u, v within 1e-5 px)."""
    f = synth.affine(0.6, -0.4, 2e-3, 1e-3, -1e-3, 1.5e-3, 99.5, 99.5)
    ref = synth.quantize_unit(synth.render(200, 200, 620, 17))
    dfm = synth.quantize_unit(synth.render(200, 200, 620, 17, f))
    roi = np.ones((200, 200), np.uint8)
    u, v, _ = f.displacement(np.array([100.0]), np.array([100.0]))
    su, sv = float(np.round(u[0])), float(np.round(v[0]))
    pre = orc.Precomputed(dfm, degree=5)
    gx, gy = orc.reference_gradient_cache(ref)
    seed_row = np.array([[100.0, 100.0, su, sv, 0.0, 0.0, 0.0, 0.0, 0.0]])
    cpu, ncpu = orc.propagate(ref, pre, gx, gy, roi, seed_row, radius=79, max_iterations=30, threshold=1e-3,
                              spacing=3, max_znssd=2.0)
    cfg = pkg.config_from_dict({
        "subset": {"radius": 79}, "shape_function": {"order": 1},
        "optimization": {"method": "icgn", "max_iterations": 30, "convergence_threshold": 1e-3},
        "correlation": {"criterion": "znssd"}, "interpolation": {"degree": 5},
        "reliability_propagation": {"spacing": 3, "max_znssd": 2.0}})
    solver = pkg.ICGNSolver(cfg, precision="fp64", context=ctx)
    seed = _seed(pkg, 100.0, 100.0, su, sv, (0.0, 0.0, 0.0, 0.0), 0.0)
    _, gpu, ngpu = pkg.propagate(ref, dfm, roi, seed, solver=solver)
    valid = cpu["valid"] == 1
    d = np.abs(gpu["p"][valid][:, :2] - cpu["p"][valid][:, :2]).max()
    _report(f"radius 79 walk: gpu {ngpu} / oracle {ncpu} points, max |du| {d:.2e}")
    assert ncpu > 50 and ngpu == ncpu
    assert np.array_equal(gpu["valid"], cpu["valid"])
    assert d <= 1e-5
