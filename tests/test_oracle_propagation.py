"""Pin of the oracle's reliability-guided propagation
(reliability_propagation.cpp:130-394 restated in oracle/dic_oracle.c) on the
reference's own golden fields: a window of the stereo-plate case (images and
ROI as the reference holds them, cropped with the grid alignment kept),
seeded at one golden POI, must reproduce the golden valid set exactly.  The
golden run used other seeds, so every POI starts from another parent's
extrapolated guess; IC-GN stops when ||dp|| < 1e-6, which leaves the
converged point within ~1e-5 px of the minimum for well-conditioned subsets
and a few times that for slowly converging ones: u, v must agree within
1e-5 px for >= 97% of the POIs and within 1e-4 px for all."""
import numpy as np
import pytest

from conftest import gradient_init, read_gray

WINDOW = (560, 240, 900, 560)  # x0, y0, x1, y1; x0, y0 multiples of the stride 4


def crop_case(stereo, name):
    """Reference window WINDOW; deformed window shifted by the field's integer
    mean displacement (ox, oy), so in the crops' frames u' = u - ox, v' = v - oy
    stays small and every warped subset lies well inside the deformed crop."""
    rows = stereo["field_" + name]
    a, b = stereo["pair_" + name]
    x0, y0, x1, y1 = WINDOW
    ox, oy = int(round(np.median(rows[:, 2]))), int(round(np.median(rows[:, 3])))
    ref = read_gray(stereo["img_" + a])[y0:y1, x0:x1]
    dfm = read_gray(stereo["img_" + b])[y0 + oy:y1 + oy, x0 + ox:x1 + ox]
    roi = stereo["roi"][y0:y1, x0:x1].copy()
    g = gradient_init(rows)
    c = np.array([(x0 + x1) / 2, (y0 + y1) / 2])
    k = int(np.argmin(np.hypot(rows[:, 0] - c[0], rows[:, 1] - c[1])))
    seed = np.array([[rows[k, 0] - x0, rows[k, 1] - y0, rows[k, 2] - ox, rows[k, 3] - oy, *g[k], rows[k, 4]]])
    inside = (rows[:, 0] - 41 >= x0) & (rows[:, 1] - 41 >= y0) & (rows[:, 0] + 41 < x1) & (rows[:, 1] + 41 < y1)
    return ref, dfm, roi, seed, rows[inside] - [x0, y0, ox, oy, 0]


@pytest.mark.parametrize("name", ["left_temporal", "deformed_disparity"])
def test_oracle_propagation_reproduces_golden_window(orc, stereo, name):
    ref, dfm, roi, seed, golden = crop_case(stereo, name)
    pre = orc.Precomputed(dfm, degree=5)
    gx, gy = orc.reference_gradient_cache(ref)
    grid, n = orc.propagate(ref, pre, gx, gy, roi, seed, radius=41, max_iterations=50, threshold=1e-6, spacing=3,
                            max_znssd=2.0)
    flat = grid.ravel()
    valid = flat[flat["valid"] == 1]
    assert n == valid.size
    got = {(int(r["x"]), int(r["y"])): r for r in valid}
    want = {(int(r[0]), int(r[1])): r for r in golden}
    assert set(got) == set(want), (len(got), len(want))
    du = np.array([abs(got[k]["p"][0] - want[k][2]) for k in want])
    dv = np.array([abs(got[k]["p"][1] - want[k][3]) for k in want])
    bad = int(((du > 1e-5) | (dv > 1e-5)).sum())
    print(f"oracle propagation {name}: {len(got)} valid = golden, max |du| {du.max():.2e}, outside 1e-5: {bad}")
    assert bad <= 0.03 * len(want) and max(du.max(), dv.max()) <= 1e-4
    # rejected / unvisited grid points keep the defaults (reliability_propagation.cpp:167-171)
    rest = flat[flat["valid"] == 0]
    assert np.all(rest["status"] == 2) and np.all(rest["p"] == 0)
