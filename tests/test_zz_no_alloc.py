"""The no-allocation entry (dic_icgn_solve_batch_device_noalloc) against
dic_icgn_solve_batch_device, through the C-ABI: same kernels, so the same
bits, on two calls in a row (the second with fewer POIs and a smaller
subset, so the rebuilt table is shorter than the one left behind), in both
precisions, second- and first-order, masked included; radii beyond the
array are refused before any launch."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pair(synth):
    W = H = 224
    f = synth.affine(0.8, -1.3, 2e-3, -1e-3, 1.5e-3, 1e-3, 112.0, 112.0)
    ref = synth.quantize_unit(synth.render(W, H, 900, 41))
    dfm = synth.quantize_unit(synth.render(W, H, 900, 41, f))
    return ref, dfm, f


def _inputs(pkg, ctx, synth, f, W, H, step, radius):
    xy = synth.poi_grid(W, H, step)
    xy = np.ascontiguousarray(xy[synth.solvable_mask(xy, W, H, radius)])
    u, v, _ = f.displacement(xy[:, 0], xy[:, 1])
    ini = pkg.initial_guesses(np.round(u), np.round(v), confidence=0.25)
    return xy, ini, ctx.to_device(xy), ctx.to_device(ini)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("masked", [False, True])
def test_noalloc_entry_equals_device_entry(pkg, ctx, synth, pair, precision, masked):
    ref, dfm, f = pair
    H, W = ref.shape
    prec = pkg.PRECISION_FP64 if precision == "fp64" else pkg.PRECISION_MIXED
    dref, ddfm = ctx.upload_image(ref), ctx.upload_image(dfm)
    grad = ctx.ref_gradient(dref, pkg.GRADIENT_CENTRAL, 5, 3, prec)
    co = ctx.prefilter(ddfm, 5, 3, True, prec)
    yy, xx = np.mgrid[0:H, 0:W]
    roi = ((xx - 110) ** 2 + (yy - 118) ** 2 < 95 ** 2).astype(np.uint8)
    droi = ctx.to_device(roi) if masked else None
    # second call: fewer POIs and a smaller subset than the first
    for radius, order, step in ((21, 2, 6), (12, 1, 11)):
        xy, ini, dxy, dini = _inputs(pkg, ctx, synth, f, W, H, step, radius)
        n = xy.shape[0]
        cfg = pkg.SubsetConfig(subset_radius=radius,
                               shape_function="second_order" if order == 2 else "first_order")
        params = pkg.make_params(cfg, prec, masked)
        want_buf = ctx.to_device(np.full(n * pkg.RESULT_DTYPE.itemsize, 0xAB, np.uint8))
        got_buf = ctx.to_device(np.full(n * pkg.RESULT_DTYPE.itemsize, 0xCD, np.uint8))
        rp = droi.ptr if masked else None
        ctx.solve_device(params, dref, grad, co, dxy.ptr, dini.ptr, n, want_buf.ptr, rp)
        ctx.solve_device_noalloc(params, dref, grad, co, dxy.ptr, dini.ptr, n, got_buf.ptr, rp)
        ctx.synchronize()
        want, got = want_buf.download(pkg.RESULT_DTYPE), got_buf.download(pkg.RESULT_DTYPE)
        ok = want["status"] == pkg.SUCCESS
        print(f"noalloc {precision} masked={masked} r={radius} order={order}: {n} POIs, {ok.sum()} Success",
              flush=True)
        assert ok.sum() > 0.5 * n
        assert np.array_equal(got["status"], want["status"]) and np.array_equal(got["valid"], want["valid"])
        assert np.abs(got["p"][ok, :2] - want["p"][ok, :2]).max() <= 1e-5
        assert got.tobytes() == want.tobytes()  # the same kernels: the same bits


def test_noalloc_entry_refuses_radii_beyond_its_table(pkg, ctx, synth, pair):
    ref, dfm, f = pair
    dref, ddfm = ctx.upload_image(ref), ctx.upload_image(dfm)
    grad = ctx.ref_gradient(dref, pkg.GRADIENT_CENTRAL, 5, 3, pkg.PRECISION_MIXED)
    co = ctx.prefilter(ddfm, 5, 3, True, pkg.PRECISION_MIXED)
    xy, ini, dxy, dini = _inputs(pkg, ctx, synth, f, 224, 224, 50, 10)
    out = ctx.alloc(xy.shape[0] * pkg.RESULT_DTYPE.itemsize)
    cfg = pkg.SubsetConfig(subset_radius=pkg._lib.NOALLOC_MAX_RADIUS + 1)
    with pytest.raises(pkg.DicError, match="accepts radii up to 102"):
        ctx.solve_device_noalloc(pkg.make_params(cfg, pkg.PRECISION_MIXED), dref, grad, co, dxy.ptr, dini.ptr,
                                 xy.shape[0], out.ptr)
