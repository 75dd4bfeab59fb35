"""Cross-camera subset matching of a stereo pair (second order, radius 25,
POI step 8): left view = reference, right view = deformed, guesses from the
flat-plane epipolar prediction.  FP64 GPU vs the oracle on a POI sample
(identical statuses, |du|, |dv| <= 1e-5 px), mixed vs FP64 (<= 1e-3 px), and
both against the geometric truth of the bulged surface."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _report(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def case(synth):
    return synth.stereo_pair()


def _solve(pkg, ctx, c, precision):
    cfg = pkg.SubsetConfig(subset_radius=c.radius, shape_function="second_order")
    dref, ddef = ctx.upload_image(c.left), ctx.upload_image(c.right)
    grad = ctx.ref_gradient(dref, 0, 5, 3, precision)
    co = ctx.prefilter(ddef, 5, 3, True, precision)
    return ctx.solve(pkg.make_params(cfg, precision), dref, grad, co, c.poi_xy, c.init)


def test_stereo_fp64_oracle_mixed_and_truth(pkg, ctx, orc, case):
    c = case
    fp = _solve(pkg, ctx, c, pkg.PRECISION_FP64)
    mx = _solve(pkg, ctx, c, pkg.PRECISION_MIXED)
    sel = np.arange(0, c.poi_xy.shape[0], 97)
    pre = orc.Precomputed(c.right, degree=5)
    gx, gy = orc.reference_gradient_cache(c.left)
    cpu = orc.icgn_solve(c.left, pre, gx, gy, c.poi_xy[sel], c.init[sel], radius=c.radius, max_iterations=30,
                         threshold=1e-3, shape_order=2)
    assert np.array_equal(fp["status"][sel], cpu["status"])
    okc = cpu["status"] == 0
    d_or = np.abs(fp["p"][sel][okc, :2] - cpu["p"][okc, :2]).max(initial=0)
    ok = (fp["status"] == 0) & (mx["status"] == 0)
    d_mx = np.abs(mx["p"][ok, :2] - fp["p"][ok, :2]).max(axis=1)
    err = np.hypot(fp["p"][ok, 0] + c.poi_xy[ok, 0] - c.truth_xy[ok, 0],
                   fp["p"][ok, 1] + c.poi_xy[ok, 1] - c.truth_xy[ok, 1])
    _report(f"stereo: {c.poi_xy.shape[0]} POIs, {np.mean(fp['status'] == 0):.4f} Success, "
            f"|fp64-oracle| {d_or:.2e} on {sel.size}, |mixed-fp64| max {d_mx.max():.2e}, "
            f"error vs truth median {np.median(err):.4f} p99 {np.percentile(err, 99):.4f} px")
    assert d_or <= 1e-5
    assert np.mean(fp["status"] == 0) > 0.98
    assert np.mean(d_mx <= 1e-3) > 0.995
    assert np.median(err) < 0.02
