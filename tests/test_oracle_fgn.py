"""Pins of the oracle's forward-compositional Gauss-Newton part
(forward_gauss_newton.cpp restated in oracle/dic_oracle.c):

* compose_warp: identity, the affine group law (first order, exact) and the
  exact composition of second-order warps truncated to second order in the
  subset coordinates and first order in dp;
* solver: identity pair, integer shift, convergence to the IC-GN minimum of
  the same objective (the IC-GN oracle is pinned by the reference's golden
  fields), the per-variant step caps and which norm the convergence test
  reads, and the Levenberg-Marquardt term on 3 of the 4 second-order
  variants (one step against an independent numpy assembly of H and G from
  the pinned B-spline value/gradient)."""
import numpy as np
import pytest


def _affine(p):
    return np.array([[1 + p[2], p[3], p[0]], [p[4], 1 + p[5], p[1]], [0, 0, 1.0]])


def test_compose_identity(orc):
    rng = np.random.default_rng(3)
    for P in (6, 12):
        p = rng.normal(0, 0.05, P)
        assert np.allclose(orc.fgn_compose(p, np.zeros(P)), p, atol=0, rtol=0)
        assert np.allclose(orc.fgn_compose(np.zeros(P), p), p, atol=0, rtol=0)


def test_compose_first_order_is_affine_group_law(orc):
    rng = np.random.default_rng(4)
    for _ in range(5):
        p, dp = rng.normal(0, 0.3, 6), rng.normal(0, 0.3, 6)
        want = _affine(p) @ _affine(dp)  # W(W(z; dp); p)
        got = _affine(orc.fgn_compose(p, dp))
        assert np.allclose(got, want, rtol=1e-14, atol=1e-14)


def _poly_warp(p):
    """Second-order warp as 2D polynomials {(i, j): coef} of (x, y)."""
    u = {(0, 0): p[0], (1, 0): 1 + p[2], (0, 1): p[3], (2, 0): p[6] / 2, (1, 1): p[7], (0, 2): p[8] / 2}
    v = {(0, 0): p[1], (1, 0): p[4], (0, 1): 1 + p[5], (2, 0): p[9] / 2, (1, 1): p[10], (0, 2): p[11] / 2}
    return u, v


def _pmul(a, b):
    out = {}
    for (i, j), c in a.items():
        for (k, l), d in b.items():
            out[(i + k, j + l)] = out.get((i + k, j + l), 0.0) + c * d
    return out


def _padd(*ps):
    out = {}
    for sc, p in ps:
        for k, c in p.items():
            out[k] = out.get(k, 0.0) + sc * c
    return out


def _compose_exact(p, dp):
    """Coefficients of degree <= 2 of W(W(z; dp); p)."""
    X, Y = _poly_warp(dp)
    one = {(0, 0): 1.0}
    pu, pv = _poly_warp(p)
    terms_u = _padd((pu[(0, 0)], one), (pu[(1, 0)], X), (pu[(0, 1)], Y), (pu[(2, 0)], _pmul(X, X)),
                    (pu[(1, 1)], _pmul(X, Y)), (pu[(0, 2)], _pmul(Y, Y)))
    terms_v = _padd((pv[(0, 0)], one), (pv[(1, 0)], X), (pv[(0, 1)], Y), (pv[(2, 0)], _pmul(X, X)),
                    (pv[(1, 1)], _pmul(X, Y)), (pv[(0, 2)], _pmul(Y, Y)))
    u, v = terms_u, terms_v
    g = lambda d, k: d.get(k, 0.0)  # noqa: E731
    return np.array([g(u, (0, 0)), g(v, (0, 0)), g(u, (1, 0)) - 1, g(u, (0, 1)), g(v, (1, 0)), g(v, (0, 1)) - 1,
                     2 * g(u, (2, 0)), g(u, (1, 1)), 2 * g(u, (0, 2)), 2 * g(v, (2, 0)), g(v, (1, 1)),
                     2 * g(v, (0, 2))])


def test_compose_second_order_is_first_order_in_dp(orc):
    """The reference's second-order compose_warp (:1025-1053) is the exact
    composition to first order in dp, except that the d2u/dy2 and d2v/dy2
    components carry p7 dp3 and p10 dp3 where the exact expansion has
    2 p7 dp3 and 2 p10 dp3 (kept: parity follows the reference)."""
    rng = np.random.default_rng(5)
    p = rng.normal(0, 0.05, 12)
    d = rng.normal(0, 1.0, 12)
    errs = []
    for eps in (1e-3, 1e-4):
        dp = eps * d
        want = _compose_exact(p, dp)
        want[8] -= p[7] * dp[3]
        want[11] -= p[10] * dp[3]
        errs.append(np.abs(orc.fgn_compose(p, dp) - want).max())
    # the remaining dropped terms are products dp_i dp_j: 100x smaller per 10x in eps
    assert errs[1] < 1e-8 and errs[0] / errs[1] > 50, errs


@pytest.fixture(scope="module")
def speckle(synth):
    return synth.quantize_unit(synth.render(128, 128, 260, 5))


def _ini(orc, u, v, n=1, conf=0.0):
    ini = np.zeros(n, orc.INITIAL_DTYPE)
    ini["u"], ini["v"], ini["confidence"] = u, v, conf
    return ini


@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("objective", [0, 1])
def test_identity_and_integer_shift(orc, speckle, order, objective):
    pre = orc.Precomputed(speckle, degree=3)
    r = orc.fgn_solve(speckle, pre, np.array([[64.0, 64.0]]), _ini(orc, 0, 0), radius=15, shape_order=order,
                      objective=objective)
    assert r["status"][0] == 0 and np.abs(r["p"][0, :2]).max() < 1e-9
    shifted = np.roll(np.roll(speckle, 3, axis=1), -2, axis=0)
    pre = orc.Precomputed(shifted, degree=3)
    r = orc.fgn_solve(speckle, pre, np.array([[64.0, 64.0]]), _ini(orc, 3, -2), radius=15, shape_order=order,
                      objective=objective)
    assert r["status"][0] == 0 and np.abs(r["p"][0, :2] - [3, -2]).max() < 1e-9


@pytest.mark.parametrize("order", [1, 2])
def test_converges_to_the_icgn_minimum(orc, synth, order):
    """Both solvers minimise the same objective over the same warps; run to a
    tight threshold they reach the same parameters."""
    f = synth.affine(0.35, -0.2, 1e-3, 0.0, 0.0, -5e-4, 63.5, 63.5)
    ref = synth.quantize_unit(synth.render(128, 128, 260, 7))
    dfm = synth.quantize_unit(synth.render(128, 128, 260, 7, f))
    xy = np.array([[40.0, 40.0], [64.0, 64.0], [88.0, 70.0]])
    ini = _ini(orc, 0, 0, n=3)
    pre = orc.Precomputed(dfm, degree=5)
    fg = orc.fgn_solve(ref, pre, xy, ini, radius=15, shape_order=order, threshold=1e-10, max_iterations=200)
    gx, gy = orc.reference_gradient_cache(ref)
    ic = orc.icgn_solve(ref, pre, gx, gy, xy, ini, radius=15, shape_order=order, threshold=1e-10,
                        max_iterations=200)
    assert np.all(fg["status"] == 0) and np.all(ic["status"] == 0)
    assert np.abs(fg["p"][:, :2] - ic["p"][:, :2]).max() < 5e-3
    u, v, _ = f.displacement(xy[:, 0], xy[:, 1])
    assert np.abs(fg["p"][:, 0] - u).max() < 0.05 and np.abs(fg["p"][:, 1] - v).max() < 0.05


# (order, objective, masked) -> (cap, convergence test on the norm before the cap)
CAPS = {(1, 0, False): (0.5, False), (1, 0, True): (0.5, False), (1, 1, False): (0.35, True),
        (1, 1, True): (0.35, False), (2, 0, False): (0.5, False), (2, 0, True): (0.3, True),
        (2, 1, False): (0.35, False), (2, 1, True): (0.35, False)}


@pytest.mark.parametrize("key", sorted(CAPS))
def test_step_caps_per_variant(orc, speckle, key):
    order, objective, masked = key
    cap, precap = CAPS[key]
    shifted = np.roll(speckle, 2, axis=1)
    pre = orc.Precomputed(shifted, degree=3)
    roi = np.ones_like(speckle, dtype=np.uint8) if masked else None
    xy = np.array([[64.0, 64.0]])
    # one step from a guess 2 px off: the full step is capped, so the first
    # step's length (translation-only guess: p1 - p0 = dp) is the cap
    r = orc.fgn_solve(speckle, pre, xy, _ini(orc, 0, 0), radius=15, shape_order=order, objective=objective,
                      roi=roi, max_iterations=1)
    assert r["status"][0] == 0
    P = 12 if order == 2 else 6
    assert abs(np.linalg.norm(r["p"][0, :P]) - cap) < 1e-12
    # with a threshold above the cap, the variant that tests the capped norm
    # stops after one step; the one that tests the norm before the cap goes on
    r = orc.fgn_solve(speckle, pre, xy, _ini(orc, 0, 0), radius=15, shape_order=order, objective=objective,
                      roi=roi, max_iterations=3, threshold=cap + 0.02)
    assert r["iterations"][0] == (3 if precap else 1)


def _numpy_one_step(orc, ref, pre, x0, y0, p, radius, objective, lam):
    """One FGN step from p assembled in numpy from the oracle's (pinned)
    B-spline value and gradient, with H + lam * max(diag H) I."""
    pts = [(dx, dy) for dy in range(-radius, radius + 1) for dx in range(-radius, radius + 1)
           if dx * dx + dy * dy <= radius * radius]
    lx = np.array([q[0] for q in pts], float)
    ly = np.array([q[1] for q in pts], float)
    f = ref[(y0 + ly).astype(int), (x0 + lx).astype(int)].astype(float)
    q = np.stack([lx * lx / 2, lx * ly, ly * ly / 2])
    wx = x0 + lx + p[0] + p[2] * lx + p[3] * ly + p[6:9] @ q
    wy = y0 + ly + p[1] + p[4] * lx + p[5] * ly + p[9:12] @ q
    val, grad = pre.sample(np.stack([wx, wy], 1))
    a00 = 1 + p[2] + p[6] * lx + p[7] * ly
    a01 = p[3] + p[7] * lx + p[8] * ly
    a10 = p[4] + p[9] * lx + p[10] * ly
    a11 = 1 + p[5] + p[10] * lx + p[11] * ly
    ge = grad[:, 0] * a00 + grad[:, 1] * a10
    he = grad[:, 0] * a01 + grad[:, 1] * a11
    sd = np.stack([ge, he, ge * lx, ge * ly, he * lx, he * ly, ge * q[0], ge * q[1], ge * q[2], he * q[0],
                   he * q[1], he * q[2]], 1)
    if objective == 0:
        fn = (f - f.mean()) / np.linalg.norm(f - f.mean())
        gn = (val - val.mean()) / np.linalg.norm(val - val.mean())
        e = fn - gn
        fnorm = np.linalg.norm(f - f.mean())
        H = sd.T @ sd * 2 / fnorm ** 2
        G = sd.T @ e * 2 / fnorm
    else:
        e = f - val
        H, G = sd.T @ sd, sd.T @ e
    H = H + lam * np.max(np.diag(H)) * np.eye(12)
    return np.linalg.solve(H, G)


@pytest.mark.parametrize("objective", [0, 1])
@pytest.mark.parametrize("masked", [False, True])
def test_levenberg_marquardt_on_three_second_order_variants(orc, speckle, objective, masked):
    f = np.roll(speckle, 1, axis=0)
    pre = orc.Precomputed(f, degree=3)
    p0 = np.zeros(12)
    p0[0], p0[1], p0[2], p0[5] = 0.2, 0.9, 0.004, -0.003
    ini = _ini(orc, p0[0], p0[1])
    ini["du_dx"], ini["dv_dy"] = p0[2], p0[5]
    roi = np.ones_like(speckle, dtype=np.uint8) if masked else None
    r = orc.fgn_solve(speckle, pre, np.array([[64.0, 64.0]]), ini, radius=15, shape_order=2, objective=objective,
                      roi=roi, max_iterations=1, threshold=0.0)
    lm = not (objective == 0 and masked)
    dp_lm = _numpy_one_step(orc, speckle, pre, 64, 64, p0, 15, objective, 1e-4)
    dp_no = _numpy_one_step(orc, speckle, pre, 64, 64, p0, 15, objective, 0.0)
    assert np.linalg.norm(dp_lm) < {(0, False): 0.5, (0, True): 0.3}.get((objective, masked), 0.35)
    want = orc.fgn_compose(p0, dp_lm if lm else dp_no)
    other = orc.fgn_compose(p0, dp_no if lm else dp_lm)
    got = r["p"][0]
    assert np.abs(got - want).max() < 1e-9 * max(1.0, np.abs(want).max())
    assert np.abs(other - want).max() > 1e3 * np.abs(got - want).max()
