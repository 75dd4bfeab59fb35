"""The oracle against the reference's own per-POI goldens
(tests/data/f0b/stereo_plate/fields/*.npz of Traditional-DIC, produced by its
pipeline with config/subset_2d.yaml: radius 41, quintic, IC-GN first order,
ZNSSD, 50 iterations, ||dp|| < 1e-6).  Every valid golden POI is re-solved
from init = golden (u, v) with affine terms from finite differences of the
golden field; a converged re-solve must land on the golden fixed point."""
import numpy as np
import pytest

from conftest import gradient_init, read_gray

FIELDS = ["reference_disparity", "left_temporal", "deformed_disparity"]


@pytest.mark.parametrize("name", FIELDS)
def test_golden_fields_oracle(orc, stereo, name):
    rows_total, valid = stereo["rows_" + name]
    assert rows_total == 97200 and valid == 7654  # 1440/4 x 1080/4 grid
    ref = read_gray(stereo["img_00_L"])
    a, b = stereo["pair_" + name]
    dfm = read_gray(stereo["img_" + b])
    rows = stereo["field_" + name]
    g = gradient_init(rows)
    sel = np.arange(0, rows.shape[0], 17)
    ini = np.zeros(sel.size, orc.INITIAL_DTYPE)
    ini["u"], ini["v"] = rows[sel, 2], rows[sel, 3]
    ini["du_dx"], ini["du_dy"], ini["dv_dx"], ini["dv_dy"] = g[sel].T
    pre = orc.Precomputed(dfm, degree=5)
    gx, gy = orc.reference_gradient_cache(ref)
    r = orc.icgn_solve(ref, pre, gx, gy, rows[sel, :2], ini, radius=41, max_iterations=50, threshold=1e-6)
    assert (r["status"] == 0).all()
    du = np.abs(r["p"][:, 0] - rows[sel, 2])
    dv = np.abs(r["p"][:, 1] - rows[sel, 3])
    dc = np.abs(r["correlation"] - rows[sel, 4])
    ok = (du <= 1e-5) & (dv <= 1e-5) & (dc <= 2e-6)
    # at most the propagation seed (stored from the seed path's non-prefiltered
    # interpolant, seed_selector.cpp:166-173) may differ
    assert (~ok).sum() <= 1, (du.max(), dv.max(), dc.max())
    assert np.percentile(du[ok], 99) < 1e-6
