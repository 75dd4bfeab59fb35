"""Pins of the oracle's seeding restatement (integer_search.cpp,
seed_selector.cpp, subset_initializer.cpp as restated in oracle/dic_oracle.c).

Each pin checks the oracle against something it does not share code with:
lattice-point counts of the circle (OEIS A000328), numpy's corrcoef and std,
scipy's k-means, brute force over every shift, integer shifts whose answer is
known, exact ties in periodic images, and the known displacement of a
synthetic pair."""
import ctypes

import numpy as np
import pytest
from scipy.cluster.vq import kmeans2

GAUSS_CIRCLE = [1, 5, 13, 29, 49, 81, 113, 149, 197, 253, 317]  # lattice points with x^2 + y^2 <= r^2


def circle(r):
    d = np.arange(-r, r + 1)
    dx, dy = np.meshgrid(d, d)
    m = dx * dx + dy * dy <= r * r
    return dx[m], dy[m]  # row-major (dy outer, dx inner)


def speckle(synth, w, h, seed):
    return synth.render(w, h, max(8, w * h // 40), seed)


def brute_force(ref, dfm, cx, cy, r, u_range=None, v_range=None):
    """First maximum in row-major scan of corrcoef over every shift keeping the subset inside."""
    H, W = ref.shape
    dx, dy = circle(r)
    f = ref[cy + dy, cx + dx].astype(np.float64)
    best = (-np.inf, 0, 0)
    us = range(r - cx, W - r - cx) if u_range is None else u_range
    vs = range(r - cy, H - r - cy) if v_range is None else v_range
    for v in vs:
        for u in us:
            x, y = cx + u, cy + v
            if not (x - r >= 0 and y - r >= 0 and x + r < W and y + r < H):
                continue
            g = dfm[y + dy, x + dx].astype(np.float64)
            if g.std() == 0:
                continue
            z = np.corrcoef(f, g)[0, 1]
            if z > best[0]:
                best = (z, u, v)
    return best


def test_struct_layouts_match_library(orc, pkg):
    for a, b in ((orc.OrcSearchParams, pkg._lib.IntegerSearchParams), (orc.OrcSeedParams, pkg._lib.SeedParams)):
        assert ctypes.sizeof(a) == ctypes.sizeof(b)
        assert [f[0] for f in a._fields_][1:] == [f[0] for f in b._fields_][1:] or a is orc.OrcSearchParams
        for (na, _), (nb, _) in zip(a._fields_, b._fields_):
            assert getattr(a, na).offset == getattr(b, nb).offset
    assert orc.ESTIMATE_DTYPE == pkg._lib.ESTIMATE_DTYPE
    assert orc.SEED_EVAL_DTYPE.itemsize == pkg._lib.SEED_EVAL_DTYPE.itemsize


def test_subset_zncc_matches_corrcoef_on_lattice_circle(orc, synth):
    for r, n in enumerate(GAUSS_CIRCLE):
        assert circle(r)[0].size == n
    ref = speckle(synth, 64, 60, 11)
    dfm = speckle(synth, 64, 60, 12)
    rng = np.random.default_rng(0)
    for r in (1, 3, 6, 10):
        dx, dy = circle(r)
        for _ in range(5):
            cx, cy = rng.integers(r + 4, 64 - r - 4), rng.integers(r + 4, 60 - r - 4)
            u, v = rng.integers(-3, 4, 2)
            z = orc.subset_zncc(ref, dfm, r, int(cx), int(cy), int(u), int(v))
            want = np.corrcoef(ref[cy + dy, cx + dx].astype(np.float64),
                               dfm[cy + v + dy, cx + u + dx].astype(np.float64))[0, 1]
            assert abs(z - want) <= 1e-12, (r, z, want)
    # a flat reference subset has no correlation (ref_norm <= 1e-12, :135-137)
    flat = np.full((30, 30), 0.5, np.float32)
    assert orc.subset_zncc(flat, dfm[:30, :30], 4, 15, 15, 0, 0) == -np.inf


def test_full_image_search_equals_brute_force(orc, synth):
    ref = speckle(synth, 44, 38, 21)
    dfm = speckle(synth, 44, 38, 22)
    pts = np.array([[12.0, 11.0], [30.4, 20.6], [20.0, 27.0]])
    p = orc.search_params(subset_radius=4, search_radius=0, pyramid_enabled=0)
    est = orc.integer_search(ref, dfm, pts, p)
    for (x, y), e in zip(pts, est):
        z, u, v = brute_force(ref, dfm, int(np.floor(x + 0.5)), int(np.floor(y + 0.5)), 4)
        assert e["valid"] == 1 and (e["u"], e["v"]) == (u, v)
        assert abs(e["zncc"] - z) <= 1e-12
        assert e["znssd"] == e["confidence"] == pytest.approx(2 * (1 - z), abs=1e-12)


@pytest.mark.parametrize("pyramid,search", [(1, 0), (1, 12), (0, 0)])
def test_integer_shift_is_found(orc, synth, pyramid, search):
    # blobs of 6-10 px, so that the subset sampled every 4 px still sees the pattern
    big = synth.render(160, 140, 160 * 140 // 150, 31, smin=6.0, smax=10.0)
    du, dv = 7, -5
    ref = big[20:120, 20:140]
    # deformed(x + du, y + dv) = 0.7 reference(x, y) + 0.1: ZNCC is 1 at the true shift
    dfm = (0.7 * big[20 - dv:120 - dv, 20 - du:140 - du] + 0.1).astype(np.float32)
    pts = np.array([[60.0, 50.0], [45.0, 40.0], [80.0, 60.0]])
    p = orc.search_params(subset_radius=10, search_radius=search, pyramid_enabled=pyramid, pyramid_scale=4)
    est = orc.integer_search(ref, dfm, pts, p)
    assert np.all(est["valid"] == 1)
    assert np.all(est["u"] == du) and np.all(est["v"] == dv)
    assert np.all(est["zncc"] > 1 - 1e-6)


def test_periodic_ties_resolve_in_scan_order_and_prefer_local(orc):
    rng = np.random.default_rng(5)
    period = 5
    rows = rng.random((40, period)).astype(np.float32)
    img = np.tile(rows, (1, 10))  # 40 x 50, exactly periodic in x
    cx, cy, r = 25, 20, 4
    # global only: exact ties at u = 0 mod 5 on row v = 0; the first in scan order is the smallest such u
    p = orc.search_params(subset_radius=r, search_radius=0, pyramid_enabled=0)
    e = orc.integer_search(img, img, np.array([[cx, cy]], float), p)[0]
    min_u = r - cx
    first = min_u + (-min_u) % period
    assert (e["u"], e["v"]) == (first, 0) and e["zncc"] == pytest.approx(1.0, abs=1e-12)
    # with a local window the local winner (first tie in [-7, 7]) is preferred on equal ZNCC
    p = orc.search_params(subset_radius=r, search_radius=7, pyramid_enabled=0)
    e = orc.integer_search(img, img, np.array([[cx, cy]], float), p)[0]
    assert (e["u"], e["v"]) == (-5, 0)


@pytest.mark.parametrize("noise,expect_local", [(0.05, True), (0.5, False)])
def test_local_preference_margin(orc, synth, noise, expect_local):
    big = speckle(synth, 240, 120, 41)
    ref = big[10:110, 50:230].copy()
    far = 40  # global match far outside the local window
    dfm = big[10:110, 50 - far:230 - far].copy()  # dfm(x + far) = ref(x)
    cx, cy, r = 60, 50, 6
    lu, lv = 2, 1
    rng = np.random.default_rng(3)
    patch = ref[cy - r - 1:cy + r + 2, cx - r - 1:cx + r + 2]
    dfm[cy + lv - r - 1:cy + lv + r + 2, cx + lu - r - 1:cx + lu + r + 2] = \
        patch + noise * patch.std() * rng.standard_normal(patch.shape).astype(np.float32)
    zl = orc.subset_zncc(ref, dfm, r, cx, cy, lu, lv)
    zg = orc.subset_zncc(ref, dfm, r, cx, cy, far, 0)
    assert zg > 1 - 1e-9
    assert (zl + 0.02 >= zg) == expect_local
    p = orc.search_params(subset_radius=r, search_radius=5, pyramid_enabled=0)
    e = orc.integer_search(ref, dfm, np.array([[cx, cy]], float), p)[0]
    assert (e["u"], e["v"]) == ((lu, lv) if expect_local else (far, 0))


def test_masked_search_needs_half_the_circle(orc, synth):
    big = speckle(synth, 120, 100, 51)
    ref = big[10:90, 10:110]
    dfm = big[10 + 2:90 + 2, 10 - 3:110 - 3]  # true shift u = 3, v = -2
    cx, cy, r = 50, 40, 10
    dx, dy = circle(r)
    assert dx.size == 317  # need max(12, ceil(317 / 2)) = 159 samples
    order = np.argsort((dx != 0) | (dy != 0), kind="stable")  # the centre first
    for m, ok in ((158, 0), (159, 1)):
        roi = np.zeros(ref.shape, np.uint8)
        sel = order[:m]
        roi[cy + dy[sel], cx + dx[sel]] = 1
        p = orc.search_params(subset_radius=r, search_radius=6, pyramid_enabled=0, masked=1)
        e = orc.integer_search(ref, dfm, np.array([[cx, cy]], float), p, roi=roi)[0]
        assert e["valid"] == ok, m
        if ok:
            assert (e["u"], e["v"]) == (3, -2)


def test_masked_search_ignores_pixels_outside_the_roi(orc, synth):
    big = speckle(synth, 120, 100, 61)
    ref = big[10:90, 10:110].copy()
    dfm = big[10 + 2:90 + 2, 10 - 3:110 - 3].copy()
    cx, cy, r = 50, 40, 10
    roi = np.ones(ref.shape, np.uint8)
    roi[:, :cx - 2] = 0  # left part of the subset outside the ROI
    # corrupt the deformed pixels that the ROI-outside reference pixels map to
    dfm[:, :cx - 2 + 3] = np.random.default_rng(0).random((dfm.shape[0], cx - 2 + 3)).astype(np.float32)
    p = orc.search_params(subset_radius=r, search_radius=6, pyramid_enabled=0, masked=1)
    e = orc.integer_search(ref, dfm, np.array([[cx, cy]], float), p, roi=roi)[0]
    assert (e["u"], e["v"]) == (3, -2) and e["zncc"] > 1 - 1e-9
    p.masked = 0
    e = orc.integer_search(ref, dfm, np.array([[cx, cy]], float), p)[0]
    assert e["zncc"] < 0.99


def test_around_prior_window(orc, synth):
    big = speckle(synth, 120, 100, 71)
    ref = big[10:90, 10:110]
    dfm = big[10 + 5:90 + 5, 10 - 7:110 - 7]  # u = 7, v = -5
    pts = np.array([[50.0, 40.0], [50.0, 40.0], [50.0, 40.0], [3.0, 40.0]])
    prior = np.array([[6.4, -4.6], [1.0, 1.0], [np.nan, 0.0], [7.0, -5.0]])
    p = orc.search_params(subset_radius=8, search_radius=3, around_prior=1)
    e = orc.integer_search(ref, dfm, pts, p, prior_uv=prior)
    assert (e[0]["u"], e[0]["v"], e[0]["valid"]) == (7, -5, 1)
    assert e[1]["valid"] == 1 and (e[1]["u"], e[1]["v"]) != (7, -5)  # the window [-2, 4]^2 misses v = -5
    assert e[2]["valid"] == 0  # non-finite prior (:528)
    assert e[3]["valid"] == 0  # subset outside the image (Image::at throws in the reference)


def test_texture_std_matches_numpy(orc, synth):
    ref = speckle(synth, 80, 70, 81)
    for x, y, r in ((40, 35, 10), (12, 12, 10), (30, 50, 3)):
        w = ref[y - r:y + r + 1, x - r:x + r + 1].astype(np.float64)
        assert abs(orc.texture_std(ref, x, y, r) - w.std()) <= 1e-9
    assert orc.texture_std(ref, 5, 35, 10) == 0.0  # window leaves the image (:425-430)


def test_deterministic_sample_closed_form(orc):
    assert list(orc.deterministic_sample(10, 4)) == [0, 3, 6, 9]
    assert list(orc.deterministic_sample(8, 3)) == [0, 4, 7]          # 3.5 rounds away from zero
    assert list(orc.deterministic_sample(6, 5)) == [0, 1, 3, 4, 5]    # 2.5 -> 3 (not to even)
    assert list(orc.deterministic_sample(3, 10)) == [0, 1, 2]
    assert list(orc.deterministic_sample(5, 1)) == [4]


def test_kmeans_matches_scipy(orc):
    rng = np.random.default_rng(9)
    ys, xs = np.nonzero(rng.random((150, 200)) < 0.1)
    pts = np.stack([xs, ys], 1).astype(np.int32)
    for k, it in ((16, 20), (5, 0), (40, 7)):
        got = orc.kmeans(pts, k, it)
        init = pts[np.minimum(np.floor(np.arange(k) * ((len(pts) - 1) / (k - 1)) + 0.5).astype(int), len(pts) - 1)]
        want, _ = kmeans2(pts.astype(np.float64), init.astype(np.float64), iter=it, minit="matrix") if it else (
            init.astype(np.float64), None)
        np.testing.assert_array_equal(got, want)
    # duplicate initial centres: the second wins no point in the first pass and keeps its position
    # (scipy does the same and warns)
    dup = np.array([[0, 0], [0, 0], [0, 0], [10, 10], [50, 50]], np.int32)
    got = orc.kmeans(dup, 3, 4)
    with pytest.warns(UserWarning):
        want, _ = kmeans2(dup.astype(np.float64), dup[[0, 2, 4]].astype(np.float64), iter=4, minit="matrix")
    np.testing.assert_array_equal(got, want)


def _margin_points(roi, m):
    H, W = roi.shape
    ys, xs = np.nonzero(roi)
    keep = (xs - m >= 0) & (ys - m >= 0) & (xs + m < W) & (ys + m < H)
    return np.stack([xs[keep], ys[keep]], 1)


def test_candidates_are_rounded_kmeans_centres(orc, synth):
    ref = speckle(synth, 140, 110, 91)
    roi = np.zeros(ref.shape, np.uint8)
    roi[10:100, 15:130] = 1
    p = orc.seed_params(search=dict(subset_radius=6, search_radius=10), subpixel_radius=8, seed_count=9,
                        min_texture_std=0.0, kmeans_iterations=12)
    cand = orc.seed_candidates(ref, roi, p)
    pts = _margin_points(roi, 8)
    k = 9
    init = pts[np.floor(np.arange(k) * ((len(pts) - 1) / (k - 1)) + 0.5).astype(int)]
    cen, _ = kmeans2(pts.astype(float), init.astype(float), iter=12, minit="matrix")
    np.testing.assert_array_equal(cand, np.floor(cen + 0.5))


def test_candidates_fall_back_to_nearest_textured_point(orc):
    H, W = 90, 120
    ref = np.full((H, W), 0.5, np.float32)
    roi = np.zeros((H, W), np.uint8)
    roi[5:85, 5:115] = 1
    p = orc.seed_params(search=dict(subset_radius=5, search_radius=4), subpixel_radius=5, seed_count=6,
                        min_texture_std=0.01, kmeans_iterations=5, kmeans_sample_limit=500)
    cand = orc.seed_candidates(ref, roi, p)
    pts = _margin_points(roi, 5)
    idx = orc.deterministic_sample(len(pts), 500)
    sp = pts[idx].astype(float)
    cen = orc.kmeans(pts[idx].astype(np.int32), 6, 5)
    assert cand.shape == (6, 2)
    for c, got in zip(cen, cand):
        d = (sp[:, 0] - c[0]) ** 2 + (sp[:, 1] - c[1]) ** 2
        j = np.lexsort((np.arange(len(sp)), d))[0]  # texture is 0 everywhere: the nearest point wins
        np.testing.assert_array_equal(got, sp[j])


def test_single_seed_is_the_most_textured_point(orc, synth):
    ref = np.full((80, 100), 0.3, np.float32)
    ref[30:50, 55:75] = speckle(synth, 20, 20, 7)
    roi = np.ones(ref.shape, np.uint8)
    p = orc.seed_params(search=dict(subset_radius=5, search_radius=4), subpixel_radius=5, seed_count=1)
    cand = orc.seed_candidates(ref, roi, p)
    pts = _margin_points(roi, 5)
    tex = np.array([ref[y - 5:y + 6, x - 5:x + 6].astype(np.float64).std() for x, y in pts])
    j = int(np.flatnonzero(tex >= tex.max() - 1e-12)[0])
    np.testing.assert_array_equal(cand[0], pts[j])


def test_select_seed_recovers_known_displacement(orc, synth):
    W, H = 220, 190
    field = synth.affine(3.3, -2.2, 0.004, -0.002, 0.001, 0.003, W / 2, H / 2)
    ref = synth.render(W, H, W * H // 40, 101)
    dfm = synth.render(W, H, W * H // 40, 101, field)
    roi = np.ones((H, W), np.uint8)
    p = orc.seed_params(search=dict(subset_radius=10, search_radius=12), subpixel_radius=15, seed_count=8,
                        max_znssd=0.3, min_texture_std=0.02, subpixel_max_iterations=50,
                        subpixel_convergence_threshold=1e-6)
    ev, best = orc.select_seed(ref, dfm, roi, p)
    assert best >= 0 and ev.shape[0] == 8
    ok = ev["valid"] == 1
    assert ok.sum() >= 6
    u, v, _ = field.displacement(ev["x"][ok], ev["y"][ok])
    err = np.hypot(ev["displacement"]["p"][ok, 0] - u, ev["displacement"]["p"][ok, 1] - v)
    assert err.max() < 0.05, err
    b = ev[best]
    assert b["quality_passed"] == 1 and b["displacement"]["status"] == 0
    assert b["quality"] == b["displacement"]["correlation"]
    # the best has the lowest ZNSSD (1e-9 ties go to the larger displacement norm, :112-136)
    q = ev["quality"][ok & (ev["quality_passed"] == 1)]
    assert b["quality"] <= q.min() + 1e-9
    # invalid candidates keep Displacement2D's defaults
    assert np.all(ev["displacement"]["status"][~ok] == 2)


def test_seed_selection_reproduces_the_golden_seed(orc, stereo):
    """The reference's golden stereo-plate run seeded its propagation at
    (852, 624): that POI's golden u, v is the seed path's solution (IC-GN on
    the non-prefiltered spline, seed_selector.cpp:163-173), ~3e-3 px from the
    propagated value.  The oracle's seed selection with the golden
    configuration (coarse-grid search instead of the SIFT prior) must pick the
    same point and reproduce its golden u, v."""
    from conftest import read_gray
    rows = stereo["field_left_temporal"]
    ref, dfm = read_gray(stereo["img_00_L"]), read_gray(stereo["img_04_L"])
    p = orc.seed_params(search=dict(subset_radius=10, search_radius=30, pyramid_enabled=1, pyramid_scale=4,
                                    pyramid_refinement_radius=4),
                        subpixel_radius=41, subpixel_max_iterations=50, subpixel_convergence_threshold=1e-6,
                        seed_count=16, max_znssd=2.0, min_texture_std=0.02)
    ev, best = orc.select_seed(ref, dfm, stereo["roi"], p)
    assert best >= 0
    b = ev[best]
    assert (b["x"], b["y"]) == (852.0, 624.0)
    k = int(np.flatnonzero((rows[:, 0] == 852) & (rows[:, 1] == 624))[0])
    du = abs(b["displacement"]["p"][0] - rows[k, 2])
    dv = abs(b["displacement"]["p"][1] - rows[k, 3])
    assert max(du, dv) <= 1e-5, (du, dv)
    assert abs(b["displacement"]["correlation"] - rows[k, 4]) <= 1e-6
