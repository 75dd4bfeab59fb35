"""Pins of the oracle's B-spline part (bspline.cpp restated in oracle/dic_oracle.c)
against closed forms, an FFT deconvolution (the reference's own method), a
dense circulant solve, the interpolation identity and the tensor-product form."""
import numpy as np
import pytest
import scipy.linalg


@pytest.mark.parametrize("d,vals", [
    (1, {0: 1.0, 1: 0.0}),
    (3, {0: 2 / 3, 1: 1 / 6, 2: 0.0}),
    (5, {0: 11 / 20, 1: 13 / 60, 2: 1 / 120, 3: 0.0}),
])
def test_basis_closed_forms(orc, d, vals):
    for x, v in vals.items():
        assert orc.basis(x, 0, d) == pytest.approx(v, abs=1e-15)
        assert orc.basis(-x, 0, d) == pytest.approx(v, abs=1e-15)


@pytest.mark.parametrize("d", [1, 3, 5])
def test_partition_of_unity_and_derivative_sum(orc, d):
    for t in np.linspace(-0.9, 0.9, 13):
        s = sum(orc.basis(t - k, 0, d) for k in range(-4, 5))
        assert s == pytest.approx(1.0, abs=1e-12)
        if d > 1:
            ds = sum(orc.basis(t - k, 1, d) for k in range(-4, 5))
            assert ds == pytest.approx(0.0, abs=1e-13)


@pytest.mark.parametrize("d", [1, 3, 5])
def test_qk_is_taylor_expansion(orc, d):
    """sum_m qk(m, c) t^m == beta(c - d/2 - t) on t in [0, 1) (left-limit convention)."""
    qk = orc.build_qk(d)
    n, off = d + 1, d // 2
    for t in (0.0, 0.1, 0.5, 0.77, 0.999):
        for c in range(n):
            lhs = sum(qk[m, c] * t ** m for m in range(n))
            assert lhs == pytest.approx(orc.basis(c - off - t, 0, d), abs=1e-13)


def test_symmetric_index(orc):
    assert [orc.symmetric_index(v, 4) for v in range(-5, 10)] == [3, 3, 2, 1, 0, 0, 1, 2, 3, 3, 2, 1, 0, 0, 1]


def _kernel(orc, L, d):
    k = np.zeros(L)
    k[0] = orc.basis(0, 0, d)
    for o in range(1, d // 2 + 1):
        k[o] = k[L - o] = orc.basis(o, 0, d)
    return k


@pytest.mark.parametrize("d", [3, 5])
@pytest.mark.parametrize("L", [7, 8, 13, 64, 518])
def test_deconvolution_matches_fft_and_circulant(orc, d, L):
    rng = np.random.default_rng(L * 10 + d)
    s = rng.random(L)
    k = _kernel(orc, L, d)
    fft = np.real(np.fft.ifft(np.fft.fft(s) / np.fft.fft(k)))  # bspline.cpp:191-214
    mine = orc.deconvolve_periodic(s, d)
    assert np.abs(mine - fft).max() < 1e-13
    if L <= 64:
        dense = scipy.linalg.solve(scipy.linalg.circulant(k), s)
        assert np.abs(mine - dense).max() < 1e-13


def _numpy_form_coefficients(orc, img, d, b):
    """form_coefficients by numpy: half-sample symmetric pad, FFT rows then columns."""
    p = np.pad(img.astype(np.float64), b, mode="symmetric")
    for axis in (1, 0):
        L = p.shape[axis]
        k = np.fft.fft(_kernel(orc, L, d))
        p = np.real(np.fft.ifft(np.fft.fft(p, axis=axis) / (k if axis == 1 else k[:, None]), axis=axis))
    return p


@pytest.mark.parametrize("d", [1, 3, 5])
def test_form_coefficients_matches_fft(orc, d):
    rng = np.random.default_rng(d)
    img = rng.random((23, 31)).astype(np.float32)
    c = orc.form_coefficients(img, d, 3, True)
    assert c.shape == (29, 37)
    assert np.abs(c - _numpy_form_coefficients(orc, img, d, 3)).max() < 1e-13


def test_border_is_raised_to_three_and_clamp_without_prefilter(orc):
    rng = np.random.default_rng(3)
    img = rng.random((9, 11)).astype(np.float32)
    c = orc.form_coefficients(img, 3, 1, False)  # border 1 -> 3 (bspline.cpp:261-268)
    assert c.shape == (15, 17)
    assert np.array_equal(c, np.pad(img.astype(np.float64), 3, mode="edge"))


@pytest.mark.parametrize("d", [1, 3, 5])
def test_interpolation_identity_at_every_pixel(orc, d):
    rng = np.random.default_rng(10 + d)
    img = rng.random((17, 26)).astype(np.float32)
    pre = orc.Precomputed(img, degree=d)
    yy, xx = np.mgrid[0:17, 0:26]
    val, _ = pre.sample(np.stack([xx.ravel(), yy.ravel()], 1).astype(float))
    assert np.abs(val - img.ravel()).max() < 1e-12


@pytest.mark.parametrize("d", [3, 5])
def test_value_equals_tensor_product(orc, d):
    rng = np.random.default_rng(20 + d)
    img = rng.random((20, 24)).astype(np.float32)
    pre = orc.Precomputed(img, degree=d)
    C = pre.coefficients()
    b, off, n = 3, d // 2, d + 1
    pts = np.column_stack([rng.uniform(0, 23, 50), rng.uniform(0, 19, 50)])
    val, grad = pre.sample(pts)
    for (x, y), v, g in zip(pts, val, grad):
        ix, iy = int(np.floor(x)), int(np.floor(y))
        tx, ty = x - ix, y - iy
        wx = np.array([orc.basis(tx + off - k, 0, d) for k in range(n)])
        wy = np.array([orc.basis(ty + off - j, 0, d) for j in range(n)])
        dwx = np.array([orc.basis(tx + off - k, 1, d) for k in range(n)])
        blk = C[iy + b - off: iy + b - off + n, ix + b - off: ix + b - off + n]
        assert v == pytest.approx(wy @ blk @ wx, abs=1e-12)
        assert g[0] == pytest.approx(wy @ blk @ dwx, abs=1e-11)


def test_polynomial_reproduction_in_the_interior(orc):
    """A cubic/quintic B-spline reproduces a linear ramp away from the periodic seam."""
    yy, xx = np.mgrid[0:64, 0:64]
    img = (0.25 + 0.003 * xx + 0.002 * yy).astype(np.float32)
    for d in (3, 5):
        pre = orc.Precomputed(img, degree=d)
        pts = np.array([[30.3, 31.7], [25.5, 40.25], [33.9, 28.1]])
        val, grad = pre.sample(pts)
        expect = 0.25 + 0.003 * pts[:, 0] + 0.002 * pts[:, 1]
        assert np.abs(val - expect).max() < 1e-6
        assert np.abs(grad[:, 0] - 0.003).max() < 1e-6


def test_reference_gradient_cache_closed_form(orc):
    rng = np.random.default_rng(4)
    img = rng.random((7, 9)).astype(np.float32)
    gx, gy = orc.reference_gradient_cache(img)
    f = img.astype(np.float64)
    assert np.array_equal(gx[:, 1:-1], (f[:, 2:] - f[:, :-2]) / 2.0)
    assert np.array_equal(gx[:, 0], f[:, 1] - f[:, 0])
    assert np.array_equal(gx[:, -1], f[:, -1] - f[:, -2])
    assert np.array_equal(gy[1:-1], (f[2:] - f[:-2]) / 2.0)
    assert np.array_equal(gy[0], f[1] - f[0])
