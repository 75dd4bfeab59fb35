"""The mesh solver's use of the B-spline module, on the GPU entry points.

src/mesh/mesh_dic.cpp:476-496 precomputes both images with one
BSplineImagePreprocessor (exact prefilter), takes the reference image's
gradients from that precompute (the local polynomial's derivative at
(+0.5, +0.5), bspline.cpp:285-294), and samples the deformed image's
interpolator at the warped pixels of every element: values
(fem_assembler.cpp:94) and gradients (fem_assembler.cpp:652).  Here the same
sequence runs through dic_bspline_prefilter, dic_ref_gradient(BSPLINE_HALF)
and dic_bspline_sample on a speckle pair with a Q4 mesh whose nodes carry an
analytic displacement, against the oracle's precompute and interpolator.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _q4_warped_pixels(W, H, step, field):
    """Pixel centres inside each Q4 element of a regular mesh, moved by the
    bilinear interpolation of the nodal displacements."""
    xs, ys = np.arange(step, W - step, step, dtype=float), np.arange(step, H - step, step, dtype=float)
    pts = []
    for y0, y1 in zip(ys[:-1], ys[1:]):
        for x0, x1 in zip(xs[:-1], xs[1:]):
            nx, ny = np.array([x0, x1, x1, x0]), np.array([y0, y0, y1, y1])
            nu, nv, _ = field.displacement(nx, ny)
            py, px = np.mgrid[int(y0):int(y1), int(x0):int(x1)].astype(float)
            xi, eta = (px - x0) / (x1 - x0), (py - y0) / (y1 - y0)
            N = np.stack([(1 - xi) * (1 - eta), xi * (1 - eta), xi * eta, (1 - xi) * eta])
            u = np.tensordot(nu, N, axes=1)
            v = np.tensordot(nv, N, axes=1)
            pts.append(np.column_stack([(px + u).ravel(), (py + v).ravel()]))
    return np.concatenate(pts)


@pytest.mark.parametrize("degree", [3, 5])
def test_mesh_bspline_call_pattern(pkg, ctx, orc, synth, degree):
    W, H = 224, 160
    f = synth.Field(synth.SYNTH_AFFINE_SINUSOID, u0=0.8, v0=-0.45, ux=3e-3, uy=-1e-3, vx=2e-3, vy=1.5e-3,
                    xc=W / 2, yc=H / 2, au=0.3, lu=90.0, av=0.25, lv=110.0)
    ref = synth.quantize_unit(synth.render(W, H, 600, 23))
    dfm = synth.quantize_unit(synth.render(W, H, 600, 23, f))
    pts = _q4_warped_pixels(W, H, 16, f)
    print(f"mesh pattern: {W}x{H}, degree {degree}, {pts.shape[0]} warped element pixels", flush=True)

    # preproc.compute(reference), preproc.compute(deformed) with one configuration
    dref, ddef = ctx.upload_image(ref), ctx.upload_image(dfm)
    cref = ctx.prefilter(dref, degree, 3, True, pkg.PRECISION_FP64)
    cdef = ctx.prefilter(ddef, degree, 3, True, pkg.PRECISION_FP64)
    gx, gy = ctx.ref_gradient(dref, pkg.GRADIENT_BSPLINE_HALF, degree, 3, pkg.PRECISION_FP64).download()
    oref, odef = orc.Precomputed(ref, degree=degree), orc.Precomputed(dfm, degree=degree)
    ox, oy = oref.gradients()
    assert np.abs(gx - ox).max() <= 1e-11 and np.abs(gy - oy).max() <= 1e-11

    # the deformed interpolator at the warped pixels: value and gradient
    val, grad = ctx.sample(cdef, pts)
    ov, og = odef.sample(pts)
    assert np.abs(val - ov).max() <= 1e-12 and np.abs(grad - og).max() <= 1e-11
    # the reference interpolator reproduces the image at pixel centres
    cen = np.column_stack([np.tile(np.arange(W, dtype=float), H), np.repeat(np.arange(H, dtype=float), W)])
    rv, _ = ctx.sample(cref, cen)
    assert np.abs(rv.reshape(H, W) - ref).max() <= 1e-9
    # mixed-precision storage of the same precompute: within FP32 rounding
    c32 = ctx.prefilter(ddef, degree, 3, True, pkg.PRECISION_MIXED)
    v32, _ = ctx.sample(c32, pts)
    assert np.abs(v32 - ov).max() <= 2e-6
