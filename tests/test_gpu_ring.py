"""The whole 2D Subset-DIC workflow on the GPU against the reference's golden
ring run (case/mono_DIC/ring, config/subset_2d.yaml: radius 41, spacing 3,
first-order IC-GN, ZNSSD, 50 iterations, ||dp|| < 1e-6, max ZNSSD 2.0; the
golden summary is tests/data/f0/subset_ring/summary.json of the reference).

That summary comes from historical artifacts of a solver build its
provenance does not record (unlike the stereo-plate fields, regenerated by
the current code and matched here to 1e-5 px).  Its grid and valid set are
pinned exactly: 102400 points, 48928 valid.  Its field statistics are not
reproducible by the current algorithm (neither here nor with the other
interpolant variants: the historical field holds worse points, ZNSSD max
0.025 against 0.0066, |U| max 0.13 px larger), so they bound the result
loosely: means within 5e-4 px, spreads within 0.01 px, mean ZNSSD within
2e-4.  The SIFT prior of the golden run is replaced by the integer search."""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

RING_CONFIG = {
    "subset": {"radius": 41, "truncate_roi_subsets": False},
    "shape_function": {"order": 1},
    "optimization": {"method": "icgn", "max_iterations": 50, "convergence_threshold": 1e-6},
    "correlation": {"criterion": "znssd"},
    "interpolation": {"method": "bspline", "degree": 5},
    "initialization": {"method": "integer_search",
                       "integer_search": {"subset_radius": 10, "search_radius": 30, "sift_enabled": False,
                                          "pyramid_enabled": False, "pyramid_scale": 4,
                                          "pyramid_refinement_radius": 4},
                       "subpixel_refinement": {"enabled": True, "shape_function": "first_order",
                                               "optimizer": "icgn", "subset_radius": 41, "max_iterations": 50,
                                               "convergence_threshold": 1e-6}},
    "seed_selection": {"seed_count": 16, "max_znssd": 2.0, "min_displacement_norm": 0.0, "min_texture_std": 0.02,
                       "kmeans_iterations": 20, "kmeans_sample_limit": 20000},
    "reliability_propagation": {"spacing": 3, "max_znssd": 2.0},
}


def test_ring_workflow_reproduces_golden_stats(pkg, tmp_path):
    from PIL import Image
    d = np.load(os.path.join(GOLDEN, "subset_ring.npz"), allow_pickle=False)
    golden = json.loads(str(d["summary_json"]))["statistics"]
    paths = {}
    for key, name in (("img_reference", "001.bmp"), ("img_deformed", "002.bmp"), ("roi", "003.bmp")):
        a = d[key] * (255 if key == "roi" else 1)
        Image.fromarray(a.astype(np.uint8), "L").save(tmp_path / name)
        paths[key] = tmp_path / name
    print("ring workflow: 1280x1280, golden config without SIFT", flush=True)
    run = pkg.workflow.run_subset_2d(paths["img_reference"], [paths["img_deformed"]], paths["roi"], RING_CONFIG,
                                     tmp_path / "result" / "subset", run_id="ring")
    art = run.artifacts
    assert art["stats"] == [str(tmp_path / "result" / "subset" / "002_znssd_icgn_1st" / "stats.json")]
    got = json.loads(open(art["stats"][0]).read())
    print({k: (got[k], golden[k]) for k in golden}, flush=True)
    assert set(got) == set(golden)
    for k in ("total_points", "valid_points", "invalid_points"):
        assert got[k] == golden[k], k
    for k, tol in (("u_mean", 5e-4), ("v_mean", 5e-4), ("u_std", 0.01), ("v_std", 0.01), ("mag_mean", 0.01),
                   ("znssd_mean", 2e-4)):
        assert abs(got[k] - golden[k]) <= tol, (k, got[k], golden[k])
    assert got["znssd_max"] <= golden["znssd_max"] and got["mag_max"] <= golden["mag_max"]
    rows = np.genfromtxt(art["displacements"][0], delimiter=",", names=True)
    assert rows.shape[0] == golden["total_points"] and int(rows["valid"].sum()) == golden["valid_points"]
    m = pkg.workflow.derive_subset_metrics(art["stats"])
    assert m["valid_fraction"] == golden["valid_points"] / golden["total_points"]
    # the run contract beside the frame folder (run_contract.py:779-839)
    files = pkg.run_contract.validate_run_contract(tmp_path / "result" / "subset")
    assert files["status"]["execution_status"] == "SUCCESS" and files["status"]["quality_status"] == "QUALITY_OK"
    assert files["metrics"]["valid_points"] == golden["valid_points"] and files["metrics"]["run_id"] == "ring"
    assert files["result"]["primary_artifacts"] == ["displacements:002_znssd_icgn_1st__displacements.csv"]
    assert files["result"]["secondary_artifacts"] == ["stats:002_znssd_icgn_1st__stats.json"]
