"""Pins of the oracle's IC-GN part (icgn.cpp restated in oracle/dic_oracle.c):
Eigen-semantics LDLT / FullPivLU against numpy, warp updates against warp
composition, solver invariants, failure statuses and accuracy vs truth."""
import numpy as np
import pytest


def _spd(rng, n):
    a = rng.standard_normal((n, n))
    return a @ a.T + n * np.eye(n)


@pytest.mark.parametrize("n", [6, 12])
def test_ldlt_solves_spd(orc, n):
    rng = np.random.default_rng(n)
    for _ in range(5):
        h = _spd(rng, n)
        f, tr, info, pos = orc.ldlt(h)
        assert info == 0 and pos
        b = rng.standard_normal(n)
        x = orc.ldlt_solve(f, tr, b)
        assert np.allclose(x, np.linalg.solve(h, b), rtol=1e-12, atol=1e-12)


def test_ldlt_pivots_on_largest_diagonal_and_sign(orc):
    h = np.diag([1.0, 5.0, 3.0])
    f, tr, info, pos = orc.ldlt(h)
    assert list(tr) == [1, 2, 2] and info == 0 and pos
    _, _, info, pos = orc.ldlt(-np.eye(3))
    assert info == 0 and not pos
    _, _, info, pos = orc.ldlt(np.diag([2.0, -1.0, 1.0]))
    assert not pos


def test_ldlt_semidefinite_uses_pseudo_inverse_of_d(orc):
    h = np.array([[1.0, 1.0, 0.0], [1.0, 1.0, 0.0], [0.0, 0.0, 0.0]])
    f, tr, info, pos = orc.ldlt(h)
    assert info == 0 and pos
    x = orc.ldlt_solve(f, tr, np.array([1.0, 1.0, 0.0]))
    assert np.allclose(h @ x, [1.0, 1.0, 0.0])
    assert np.allclose(x, [1.0, 0.0, 0.0])


def test_fullpivlu_inverse(orc):
    rng = np.random.default_rng(7)
    m = rng.standard_normal((6, 6))
    ok, inv = orc.fullpivlu_inverse6(m)
    assert ok and np.allclose(inv, np.linalg.inv(m), rtol=1e-12, atol=1e-12)
    m[3] = m[1]
    ok, _ = orc.fullpivlu_inverse6(m)
    assert not ok


def _affine_matrix(p):
    return np.array([[1 + p[2], p[3], p[0]], [p[4], 1 + p[5], p[1]], [0, 0, 1.0]])


def test_affine_ic_update_is_composition_with_inverse(orc):
    rng = np.random.default_rng(1)
    for _ in range(10):
        p = rng.normal(0, [1, 1, .01, .01, .01, .01])
        dp = rng.normal(0, [.1, .1, .001, .001, .001, .001])
        m = _affine_matrix(p) @ np.linalg.inv(_affine_matrix(dp))
        out = orc.ic_affine_update(p, dp)
        expect = [m[0, 2], m[1, 2], m[0, 0] - 1, m[0, 1], m[1, 0], m[1, 1] - 1]
        assert np.allclose(out, expect, rtol=1e-12, atol=1e-13)
    p = rng.normal(size=6) * 0.01
    assert np.allclose(orc.ic_affine_update(p, np.zeros(6)), p, atol=1e-15)


def _warp2(p, x, y):
    return (x + p[0] + p[2] * x + p[3] * y + 0.5 * p[6] * x * x + p[7] * x * y + 0.5 * p[8] * y * y,
            y + p[1] + p[4] * x + p[5] * y + 0.5 * p[9] * x * x + p[10] * x * y + 0.5 * p[11] * y * y)


def test_second_order_update_identity_and_affine_restriction(orc):
    rng = np.random.default_rng(2)
    p = np.concatenate([rng.normal(0, [.5, .5, .01, .01, .01, .01]), rng.normal(0, 1e-4, 6)])
    assert np.allclose(orc.ic_second_order_update(p, np.zeros(12)), p, atol=1e-14)
    pa = np.concatenate([p[:6], np.zeros(6)])
    dpa = np.concatenate([rng.normal(0, [.05, .05, 1e-3, 1e-3, 1e-3, 1e-3]), np.zeros(6)])
    out = orc.ic_second_order_update(pa, dpa)
    assert np.allclose(out[:6], orc.ic_affine_update(pa[:6], dpa[:6]), rtol=1e-10, atol=1e-12)
    assert np.abs(out[6:]).max() < 1e-12


def test_second_order_update_composes_warps_to_second_order(orc):
    """W(p_new)(x) = W(p)(W(dp)^-1(x)) + O(|x|^3) (Bai et al. 2017)."""
    rng = np.random.default_rng(3)
    p = np.concatenate([rng.normal(0, [.3, .3, .02, .02, .02, .02]), rng.normal(0, 2e-3, 6)])
    dp = np.concatenate([rng.normal(0, [.05, .05, 5e-3, 5e-3, 5e-3, 5e-3]), rng.normal(0, 5e-4, 6)])
    pn = orc.ic_second_order_update(p, dp)
    for eps in (1e-2, 5e-3):
        errs = []
        for ang in np.linspace(0, 2 * np.pi, 7):
            x, y = eps * np.cos(ang), eps * np.sin(ang)
            # invert W(dp) by Newton, starting from the affine inverse
            u, v = x, y
            for _ in range(50):
                wx, wy = _warp2(dp, u, v)
                J = np.array([[1 + dp[2] + dp[6] * u + dp[7] * v, dp[3] + dp[7] * u + dp[8] * v],
                              [dp[4] + dp[9] * u + dp[10] * v, 1 + dp[5] + dp[10] * u + dp[11] * v]])
                du = np.linalg.solve(J, [wx - x, wy - y])
                u, v = u - du[0], v - du[1]
            a = np.array(_warp2(p, u, v))
            b = np.array(_warp2(pn, x, y))
            errs.append(np.abs(a - b).max())
        # third-order remainder; the value at x = 0 must agree to rounding
        assert max(errs) < 50 * eps ** 3 + 1e-12
    u, v = 0.0, 0.0
    for _ in range(50):
        wx, wy = _warp2(dp, u, v)
        J = np.array([[1 + dp[2] + dp[6] * u + dp[7] * v, dp[3] + dp[7] * u + dp[8] * v],
                      [dp[4] + dp[9] * u + dp[10] * v, 1 + dp[5] + dp[10] * u + dp[11] * v]])
        d = np.linalg.solve(J, [wx, wy])
        u, v = u - d[0], v - d[1]
    assert np.allclose(_warp2(p, u, v), _warp2(pn, 0.0, 0.0), atol=1e-12)


# ---------------------------------------------------------------- solver ---

@pytest.fixture(scope="module")
def speckle(synth):
    ref = synth.quantize_unit(synth.render(128, 128, 260, 5))
    return ref


def _solve(orc, ref, dfm, xy, ini, **kw):
    kw.setdefault("radius", 15)
    deg = kw.pop("degree", 3)
    pre = orc.Precomputed(dfm, degree=deg)
    gx, gy = orc.reference_gradient_cache(ref)
    return orc.icgn_solve(ref, pre, gx, gy, np.asarray(xy, float), ini, **kw)


def _ini(orc, u, v, conf=0.0):
    a = np.zeros(np.size(u), orc.INITIAL_DTYPE)
    a["u"], a["v"], a["confidence"] = u, v, conf
    return a


@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("objective", [0, 1])
def test_identity_pair_converges_to_zero(orc, speckle, order, objective):
    xy = [[40, 50], [64, 64], [90, 70]]
    r = _solve(orc, speckle, speckle, xy, _ini(orc, [0, 0, 0], [0, 0, 0]), shape_order=order, objective=objective)
    assert (r["status"] == 0).all() and (r["valid"] == 1).all()
    assert np.abs(r["p"]).max() < 1e-12 and np.abs(r["correlation"]).max() < 1e-20
    assert (r["iterations"] == 1).all()


@pytest.mark.parametrize("order", [1, 2])
def test_integer_shift_is_recovered_exactly(orc, speckle, order):
    t = (3, -2)
    # def(x) = ref(x - t); B-spline interpolation is exact at integer pixels
    dfm = np.roll(np.roll(speckle, t[0], axis=1), t[1], axis=0)
    xy = [[50, 50], [64, 70]]
    r = _solve(orc, speckle, dfm, xy, _ini(orc, [t[0]] * 2, [t[1]] * 2), shape_order=order)
    assert (r["status"] == 0).all()
    assert np.abs(r["p"][:, 0] - t[0]).max() < 1e-9 and np.abs(r["p"][:, 1] - t[1]).max() < 1e-9


def test_failure_statuses(orc, speckle):
    """icgn.cpp:455-479, :524-527, :549, :563-575, :620-626 (SolverStatus order)."""
    # margin POI -> InvalidInput, initial u, v and confidence kept
    r = _solve(orc, speckle, speckle, [[5, 64]], _ini(orc, 0.5, -0.25, 0.7))
    assert r["status"][0] == 2 and r["valid"][0] == 0
    assert r["p"][0, 0] == 0.5 and r["p"][0, 1] == -0.25 and r["correlation"][0] == 0.7
    # constant reference -> NumericalFailure
    flat = np.full_like(speckle, 0.5)
    r = _solve(orc, flat, speckle, [[64, 64]], _ini(orc, 0, 0))
    assert r["status"][0] == 3
    # max_iterations = 0 -> NotConverged, invalid, correlation = inf
    r = _solve(orc, speckle, speckle, [[64, 64]], _ini(orc, 0.25, 0), max_iterations=0)
    assert r["status"][0] == 1 and r["valid"][0] == 0 and np.isinf(r["correlation"][0])
    assert r["p"][0, 0] == 0.25
    # huge initial guess -> warped point out of the image -> InvalidInput
    r = _solve(orc, speckle, speckle, [[64, 64]], _ini(orc, 200.0, 0, 0.3))
    assert r["status"][0] == 2 and r["p"][0, 0] == 200.0 and r["correlation"][0] == 0.3
    # radius 1 holds 5 samples < 6 -> InvalidInput
    r = _solve(orc, speckle, speckle, [[64, 64]], _ini(orc, 0, 0), radius=1)
    assert r["status"][0] == 2
    # running out of iterations still reports Success (icgn.cpp:620-622)
    r = _solve(orc, speckle, np.roll(speckle, 1, axis=1), [[64, 64]], _ini(orc, 0.3, 0), max_iterations=2,
               threshold=0.0)
    assert r["status"][0] == 0 and r["valid"][0] == 1 and r["iterations"][0] == 2


def test_masked_full_roi_equals_unmasked_and_min_samples(orc, speckle):
    dfm = np.roll(speckle, 2, axis=1)
    xy = [[50, 60], [70, 64]]
    ini = _ini(orc, [2, 2], [0, 0])
    a = _solve(orc, speckle, dfm, xy, ini)
    roi = np.ones_like(speckle, dtype=np.uint8)
    b = _solve(orc, speckle, dfm, xy, ini, roi=roi)
    assert np.array_equal(a["status"], b["status"]) and np.allclose(a["p"], b["p"], atol=0, rtol=0)
    # keep fewer than ceil(0.5 * full) samples -> InvalidInput
    roi2 = np.zeros_like(roi)
    roi2[:, 67:] = 1  # dx >= 1: 339 of 709 samples < ceil(0.5 * 709)
    r = _solve(orc, speckle, dfm, [[66, 64]], _ini(orc, 2, 0), roi=roi2)
    assert r["status"][0] == 2
    # a subset truncated to a bit more than half still solves
    roi3 = np.zeros_like(roi)
    roi3[:, 62:] = 1
    r = _solve(orc, speckle, dfm, [[66, 64]], _ini(orc, 2, 0), roi=roi3)
    assert r["status"][0] == 0 and abs(r["p"][0, 0] - 2) < 1e-6


def test_plain_512_accuracy_against_truth(orc, synth):
    wl = synth.plain_512()
    r = _solve(orc, wl.reference, wl.deformed, wl.poi_xy, _ini(orc, wl.init["u"], wl.init["v"]), degree=3,
               threads=0)
    u, v, g = wl.truth()
    assert (r["status"] == 0).all()
    assert np.abs(r["p"][:, 0] - u).max() < 0.03 and np.abs(r["p"][:, 1] - v).max() < 0.03
    assert np.sqrt(np.mean((r["p"][:, 0] - u) ** 2)) < 0.005
    assert np.abs(r["p"][:, 2] - g[:, 0]).max() < 0.005


def test_second_order_fits_a_quadratic_field_better(orc, synth):
    W = H = 200
    f = synth.Field(synth.SYNTH_AFFINE_SINUSOID, u0=0.3, v0=-0.2, xc=100, yc=100, au=1.2, lu=90.0, av=0.8, lv=110.0)
    ref = synth.render(W, H, 700, 21)
    dfm = synth.render(W, H, 700, 21, f)
    xy = synth.poi_grid(W, H, 20)
    xy = xy[synth.solvable_mask(xy, W, H, 20)]
    u, v, _ = f.displacement(xy[:, 0], xy[:, 1])
    ini = _ini(orc, np.round(u), np.round(v))
    e = {}
    for order in (1, 2):
        r = _solve(orc, ref, dfm, xy, ini, radius=20, shape_order=order, degree=5, threshold=1e-4)
        assert (r["status"] == 0).all()
        e[order] = np.sqrt(np.mean((r["p"][:, 0] - u) ** 2 + (r["p"][:, 1] - v) ** 2))
    assert e[2] < 0.2 * e[1] and e[2] < 0.05
