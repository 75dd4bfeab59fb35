"""Batched solving (dic_batch_*): POIs of several pairs with uneven radii and
shape orders in one hipGraph-replayed launch.

Each POI's result depends only on its own inputs, so the batch must give the
results of one dic_icgn_solve_batch_device launch per pair under every launch
mapping and dispatch order, and the same bits on every graph replay.  The
grid mapping runs the same code as the per-pair kernels and must match bit
for bit; the persistent mapping is compiled separately (the compiler may fuse
multiplies and adds differently), so it must give identical statuses, valid
flags and iteration counts with parameters within 1e-9 (FP64) / 1e-5 (mixed)
px.  The FP64 batch is also held to the oracle bar (identical statuses,
|du|, |dv| <= 1e-5 px).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _report(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def cases(synth):
    # small uneven batch: 256^2 / 384^2 pairs, radii 10..30, orders 1 and 2
    return synth.batched_pairs(n_pairs=6, sizes=(256, 384), seed0=3100)


def _setup(pkg, ctx, cases, precision, degree=5):
    prec = pkg._lib.PRECISION_FP64 if precision == "fp64" else pkg._lib.PRECISION_MIXED
    items, keep = [], []
    for c in cases:
        ref, dfm = ctx.upload_image(c.reference), ctx.upload_image(c.deformed)
        grad = ctx.ref_gradient(ref, 0, degree, 3, prec)
        co = ctx.prefilter(dfm, degree, 3, True, prec)
        n = c.poi_xy.shape[0]
        xy, ini = ctx.to_device(np.ascontiguousarray(c.poi_xy)), ctx.to_device(c.init)
        out = ctx.alloc(n * pkg.RESULT_DTYPE.itemsize)
        items.append(dict(reference=ref, gradient=grad, deformed=co, poi_xy=xy, init=ini, out=out, n=n,
                          subset_radius=c.radius, shape_order=c.shape_order, max_iterations=30,
                          convergence_threshold=1e-3))
        keep.append((ref, dfm, grad, co, xy, ini, out))
    return prec, items, keep


def _per_pair(pkg, ctx, prec, items):
    """One ordinary launch per pair (the drop-in path)."""
    res = []
    for it in items:
        cfg = pkg.SubsetConfig(subset_radius=it["subset_radius"],
                               shape_function="second_order" if it["shape_order"] == 2 else "first_order",
                               optimizer="forward_gauss_newton" if it.get("optimizer") == 1 else "icgn")
        params = pkg.make_params(cfg, prec)
        out = ctx.alloc(it["n"] * pkg.RESULT_DTYPE.itemsize)
        ctx.solve_device(params, it["reference"], it["gradient"], it["deformed"], it["poi_xy"].ptr,
                         it["init"].ptr, it["n"], out.ptr)
        res.append(out.download(pkg.RESULT_DTYPE))
    return res


def _batch_results(pkg, items):
    return [it["out"].download(pkg.RESULT_DTYPE) for it in items]


def _same_bits(a, b):
    return a.tobytes() == b.tobytes()


def _diff(a, b):
    """Largest differences between two result arrays (for the report)."""
    return dict(status=int((a["status"] != b["status"]).sum()), iters=int((a["iterations"] != b["iterations"]).sum()),
                p=float(np.abs(a["p"] - b["p"]).max(initial=0)),
                corr=float(np.abs(a["correlation"] - b["correlation"]).max(initial=0)))


def _assert_equivalent(g, r, tol):
    assert np.array_equal(g["status"], r["status"]) and np.array_equal(g["valid"], r["valid"])
    assert np.array_equal(g["iterations"], r["iterations"])
    assert np.array_equal(g[["x", "y"]], r[["x", "y"]])
    assert np.abs(g["p"] - r["p"]).max(initial=0) <= tol


TUNINGS = [
    (0, 0, 1, 0), (0, 0, 1, 1), (0, 0, 1, 2), (0, 0, 1, 3),  # grid: natural, cost, interleave, xcd
    (1, 4, 1, 0), (1, 8, 4, 1), (1, 16, 16, 2),    # persistent
    (2, 0, 1, 0), (2, 0, 1, 1), (2, 0, 1, 3),      # tile (first-order FP32 groups): natural, cost, xcd
]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_equals_per_pair_launches(pkg, ctx, cases, precision):
    prec, items, keep = _setup(pkg, ctx, cases, precision)
    ref = _per_pair(pkg, ctx, prec, items)
    orders = sorted({it["shape_order"] for it in items})
    radii = sorted({it["subset_radius"] for it in items})
    _report(f"batch {precision}: {len(items)} pairs, {sum(it['n'] for it in items)} POIs, orders {orders}, radii {radii}")
    assert orders == [1, 2]
    b = ctx.batch(items)
    for t in TUNINGS:
        for it in items:  # clear outputs so a run that skipped a POI shows
            it["out"].upload(np.zeros(it["n"], pkg.RESULT_DTYPE))
        b.set_tuning(*t)
        b.run()
        got = _batch_results(pkg, items)
        same = all(_same_bits(g, r) for g, r in zip(got, ref))
        _report(f"  tuning {b.tuning().as_dict()}: " +
                ("bit-identical to per-pair launches" if same else str([_diff(g, r) for g, r in zip(got, ref)])))
        for k, (g, r) in enumerate(zip(got, ref)):
            if t[0] in (0, 2):  # grid and tile run the per-pair code: same bits
                assert _same_bits(g, r), (t, k, _diff(g, r))
            else:
                _assert_equivalent(g, r, 1e-9 if precision == "fp64" else 1e-5)
    b.destroy()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_fgn_items_equal_per_pair_launches(pkg, ctx, cases, precision):
    """Items that name the forward Gauss-Newton optimizer (ForwardGaussNewtonSolver)
    run in their own launch groups, one POI per workgroup with the per-sample
    LDS buffer; beside IC-GN items, under every mapping, each item gives its
    per-pair launch's results bit for bit."""
    prec, items, keep = _setup(pkg, ctx, cases, precision)
    for k, it in enumerate(items):
        if k % 2 == 0:
            it["optimizer"] = pkg._lib.OPTIMIZER_FGN
    ref = _per_pair(pkg, ctx, prec, items)
    b = ctx.batch(items)
    for t in [(0, 0, 1, 3), (1, 8, 4, 1), (2, 0, 1, 0)]:
        for it in items:
            it["out"].upload(np.zeros(it["n"], pkg.RESULT_DTYPE))
        b.set_tuning(*t)
        b.run()
        got = _batch_results(pkg, items)
        _report(f"fgn items {precision}, tuning {b.tuning().as_dict()}: "
                f"{sum(int((g['status'] == 0).sum()) for g in got)} Success")
        for k, (g, r) in enumerate(zip(got, ref)):
            if items[k].get("optimizer") == 1 or t[0] in (0, 2):
                assert _same_bits(g, r), (t, k, _diff(g, r))
            else:
                _assert_equivalent(g, r, 1e-9 if precision == "fp64" else 1e-5)
    b.destroy()


@pytest.mark.parametrize("degree", [1, 3])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_lower_degrees_equal_per_pair_launches(pkg, ctx, cases, precision, degree):
    """Linear and cubic B-splines: grid and tile mappings give the per-pair
    results bit for bit, the persistent mapping within its tolerance."""
    prec, items, keep = _setup(pkg, ctx, cases, precision, degree=degree)
    ref = _per_pair(pkg, ctx, prec, items)
    b = ctx.batch(items)
    for t in [(0, 0, 1, 3), (1, 8, 4, 1), (2, 0, 1, 0)]:
        for it in items:
            it["out"].upload(np.zeros(it["n"], pkg.RESULT_DTYPE))
        b.set_tuning(*t)
        b.run()
        got = _batch_results(pkg, items)
        for k, (g, r) in enumerate(zip(got, ref)):
            if t[0] in (0, 2):
                assert _same_bits(g, r), (degree, t, k, _diff(g, r))
            else:
                _assert_equivalent(g, r, 1e-9 if precision == "fp64" else 1e-5)
    _report(f"batch degree {degree} {precision}: {sum(int((r['status'] == 0).sum()) for r in ref)} Success")
    b.destroy()


def test_batch_graph_replays_with_eager_work_between(pkg, ctx, cases):
    prec, items, keep = _setup(pkg, ctx, cases, "mixed")
    b = ctx.batch(items)
    b.set_tuning(1, 8, 4, 1)
    b.run()
    first = _batch_results(pkg, items)
    for rep in range(3):
        for it in items:
            it["out"].upload(np.zeros(it["n"], pkg.RESULT_DTYPE))
        _per_pair(pkg, ctx, prec, items[:2])  # eager work and synchronisation between replays
        b.run()
        got = _batch_results(pkg, items)
        assert all(_same_bits(g, f) for g, f in zip(got, first)), rep
    n, runs = b.info()
    assert n == sum(it["n"] for it in items) and runs == 4
    b.destroy()


def test_batch_fp64_matches_oracle(pkg, ctx, orc, cases):
    prec, items, keep = _setup(pkg, ctx, cases[:3], "fp64")
    b = ctx.batch(items)
    b.run()
    got = _batch_results(pkg, items)
    for c, g in zip(cases[:3], got):
        pre = orc.Precomputed(c.deformed, degree=5)
        gx, gy = orc.reference_gradient_cache(c.reference)
        cpu = orc.icgn_solve(c.reference, pre, gx, gy, c.poi_xy, c.init, radius=c.radius, max_iterations=30,
                             threshold=1e-3, shape_order=c.shape_order, objective=0)
        assert np.array_equal(g["status"], cpu["status"])
        ok = cpu["status"] == 0
        du = np.abs(g["p"][ok, :2] - cpu["p"][ok, :2]).max(initial=0)
        assert du <= 1e-5, du
        u, v, _ = c.field.displacement(c.poi_xy[ok, 0], c.poi_xy[ok, 1])
        err = np.abs(g["p"][ok, 0] - u).max(initial=0)
        _report(f"  pair r={c.radius} order={c.shape_order}: {ok.sum()} ok, |gpu-oracle| {du:.2e}, |u-truth| {err:.2e}")
        assert ok.mean() > 0.95
    b.destroy()


def test_batch_autotune_keeps_results(pkg, ctx, cases):
    prec, items, keep = _setup(pkg, ctx, cases, "mixed")
    ref = _per_pair(pkg, ctx, prec, items)
    b = ctx.batch(items)
    best, ms = b.autotune(reps=2)
    _report(f"autotune: best {best.as_dict()} {ms:.3f} ms")
    assert ms > 0
    assert b.tuning().as_dict() == best.as_dict()
    b.run()
    got = _batch_results(pkg, items)
    for g, r in zip(got, ref):
        _assert_equivalent(g, r, 1e-5)
    b.destroy()


def test_batch_rejects_bad_items(pkg, ctx, cases):
    prec, items, keep = _setup(pkg, ctx, cases[:2], "mixed")
    bad = dict(items[0])
    bad["shape_order"] = 3
    with pytest.raises(pkg.DicError, match="shape_order"):
        ctx.batch([items[1], bad])
    # mixing precisions in one batch is refused
    _, items64, keep64 = _setup(pkg, ctx, cases[:1], "fp64")
    with pytest.raises(pkg.DicError, match="precision"):
        ctx.batch([items[0], items64[0]])
    b = ctx.batch(items)
    with pytest.raises(pkg.DicError, match="mapping"):
        b.set_tuning(7)
    b.destroy()


def test_batch_size_mismatch_item_reports_invalid_input(pkg, ctx, cases):
    prec, items, keep = _setup(pkg, ctx, cases[:2], "mixed")
    small = ctx.upload_image(np.zeros((100, 100), np.float32))
    it = dict(items[0])
    other = ctx.prefilter(small, 5, 3, True, prec)
    it["deformed"] = other
    b = ctx.batch([it, items[1]])
    b.run()
    r0 = it["out"].download(pkg.RESULT_DTYPE)
    assert np.all(r0["status"] == pkg._lib.INVALID_INPUT)
    assert np.array_equal(r0["p"][:, :2], np.stack([it["init"].download(pkg.INITIAL_DTYPE)["u"],
                                                    it["init"].download(pkg.INITIAL_DTYPE)["v"]], 1))
    b.destroy()


def test_first_order_tile_and_fallback_match_the_gather(pkg, ctx, synth):
    """First-order FP32 per-pair launches solve four consecutive POIs per
    workgroup from an LDS tile of the coefficients staged around their
    starting subsets, and fall back to the global gather for iterations whose
    warped subset leaves it.  Guesses 3-5 px off the truth walk many POIs out
    of the tile.  The same POIs solved in a random order (the four POIs of a
    workgroup then lie far apart, the tile does not fit, every sample is
    gathered), the batch kernel (one POI per workgroup, gather only) and the
    batch tile mapping (four items per workgroup) must give the same results
    bit for bit; edge POIs and a POI count that is not a
    multiple of four are included."""
    c = synth.batched_pairs(n_pairs=1, sizes=(512,), seed0=3300)[0]
    rng = np.random.default_rng(7)
    n = (c.poi_xy.shape[0] // 4) * 4 - 1
    xy = np.ascontiguousarray(c.poi_xy[:n])
    init = c.init[:n].copy()
    far = rng.random(n) < 0.5
    init["u"][far] += rng.choice([-5.0, -3.5, 3.2, 4.6], far.sum())
    init["v"][far] += rng.choice([-4.1, -3.0, 3.3, 5.0], far.sum())
    perm = rng.permutation(n)
    prec = pkg._lib.PRECISION_MIXED
    ref, dfm = ctx.upload_image(c.reference), ctx.upload_image(c.deformed)
    grad = ctx.ref_gradient(ref, 0, 5, 3, prec)
    co = ctx.prefilter(dfm, 5, 3, True, prec)
    params = pkg.make_params(pkg.SubsetConfig(subset_radius=15, shape_function="first_order"), prec)

    def solve(order):
        dxy, dini = ctx.to_device(np.ascontiguousarray(xy[order])), ctx.to_device(init[order])
        out = ctx.alloc(n * pkg.RESULT_DTYPE.itemsize)
        ctx.solve_device(params, ref, grad, co, dxy.ptr, dini.ptr, n, out.ptr)
        per_pair = out.download(pkg.RESULT_DTYPE)
        out.upload(np.zeros(n, pkg.RESULT_DTYPE))
        b = ctx.batch([dict(reference=ref, gradient=grad, deformed=co, poi_xy=dxy, init=dini, out=out, n=n,
                            subset_radius=15, shape_order=1, max_iterations=30, convergence_threshold=1e-3)])
        b.set_tuning(0, 0, 1, 3)  # grid mapping, XCD order
        b.run()
        batched = out.download(pkg.RESULT_DTYPE)
        out.upload(np.zeros(n, pkg.RESULT_DTYPE))
        b.set_tuning(2, 0, 1, 0)  # tile mapping
        b.run()
        batched_tile = out.download(pkg.RESULT_DTYPE)
        b.destroy()
        res = np.zeros(n, pkg.RESULT_DTYPE), np.zeros(n, pkg.RESULT_DTYPE), np.zeros(n, pkg.RESULT_DTYPE)
        res[0][order], res[1][order], res[2][order] = per_pair, batched, batched_tile
        return res

    tiled_pp, tiled_b, tiled_bt = solve(np.arange(n))
    gathered_pp, gathered_b, gathered_bt = solve(perm)
    moved = int((np.hypot(tiled_pp["p"][:, 0] - init["u"], tiled_pp["p"][:, 1] - init["v"]) > 3.0).sum())
    _report(f"tile/fallback: {n} POIs, {far.sum()} guesses 3-5 px off, {moved} moved > 3 px, "
            f"{int((tiled_pp['status'] == 0).sum())} Success")
    assert moved > 0.2 * n
    assert tiled_pp.tobytes() == gathered_pp.tobytes()
    assert tiled_b.tobytes() == gathered_b.tobytes() == tiled_pp.tobytes()
    assert tiled_bt.tobytes() == gathered_bt.tobytes() == tiled_pp.tobytes()
