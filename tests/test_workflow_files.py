"""The workflow's result files (workflows/subset_2d.py, run_contract.py) on the
CPU: displacements.csv header and rows, stats.json keys and values, the
method tag, and the run metrics -- checked against the reference's golden
summary of the ring case (tests/data/f0/subset_ring/summary.json) and its F4
contract (tests/data/f4/subset_ring/contract.json: metric keys total_points,
valid_points, valid_fraction)."""
import csv
import json
import os

import numpy as np

from conftest import GOLDEN


def _ring_summary():
    d = np.load(os.path.join(GOLDEN, "subset_ring.npz"), allow_pickle=False)
    return json.loads(str(d["summary_json"]))


def _result(n=7, seed=0):
    rng = np.random.default_rng(seed)
    r = {k: rng.standard_normal(n) for k in ("x", "y", "u", "v", "du_dx", "du_dy", "dv_dx", "dv_dy", "correlation")}
    r["valid"] = rng.random(n) > 0.3
    return r


def test_displacements_csv_header_and_rows(pkg, tmp_path):
    wf = pkg.workflow
    assert wf.DISPLACEMENT_HEADER == _ring_summary()["headers"]["displacements"]
    r = _result()
    wf.save_subset_csv(r, tmp_path / "d" / "displacements.csv")
    rows = list(csv.reader(open(tmp_path / "d" / "displacements.csv")))
    assert rows[0] == wf.DISPLACEMENT_HEADER and len(rows) == 8
    for i, row in enumerate(rows[1:]):
        assert [float(x) for x in row[:-1]] == [r[k][i] for k in wf.DISPLACEMENT_HEADER[:-1]]  # round-trip exact
        assert row[-1] == str(int(r["valid"][i]))
    # the reference's loader reads it back by header (tests/regression/test_subset_ring.py:20-22)
    t = np.genfromtxt(tmp_path / "d" / "displacements.csv", delimiter=",", names=True)
    assert list(t.dtype.names) == wf.DISPLACEMENT_HEADER


def test_stats_keys_match_the_golden_summary(pkg, tmp_path):
    wf = pkg.workflow
    r = _result(50, 1)
    st = wf.write_stats(r, tmp_path / "stats.json")
    assert set(st) == set(_ring_summary()["statistics"])
    assert json.loads((tmp_path / "stats.json").read_text()) == st
    v = r["valid"]
    assert st["valid_points"] + st["invalid_points"] == st["total_points"] == 50
    assert st["u_std"] == np.std(r["u"][v]) and st["mag_max"] == np.hypot(r["u"][v], r["v"][v]).max()
    assert "ssd_mean" in wf.subset_stats(r, "ssd")
    empty = wf.subset_stats({k: np.zeros(0) for k in r}, "znssd")
    assert empty["u_mean"] == 0.0 and empty["total_points"] == 0


def test_method_tag(pkg):
    tag = pkg.workflow.subset_method_tag
    assert tag(None) == "znssd_icgn_1st"  # the golden ring run's folder 002_znssd_icgn_1st
    assert tag({"correlation": {"criterion": "SSD"}, "optimization": {"method": "forward_gauss_newton"},
                "shape_function": {"order": 2}}) == "ssd_fgn_2nd"


def test_run_metrics_follow_the_f4_contract(pkg, tmp_path):
    wf = pkg.workflow
    paths = []
    for name, seed in (("002_znssd_icgn_1st", 2), ("003_znssd_icgn_1st", 3)):
        d = tmp_path / name
        d.mkdir()
        wf.write_stats(_result(40, seed), d / "stats.json")
        paths.append(d / "stats.json")
    m = wf.derive_subset_metrics(paths)
    for k in ("total_points", "valid_points", "valid_fraction"):  # contract.json metric_keys
        assert k in m
    assert m["total_points"] == 80 and set(m["frames"]) == {"002_znssd_icgn_1st", "003_znssd_icgn_1st"}
    assert m["valid_fraction"] == m["valid_points"] / 80
    s = _ring_summary()["statistics"]  # the ring's golden run: one frame
    assert s["valid_points"] / s["total_points"] == 48928 / 102400
