"""The run contract of the 2D Subset-DIC workflow (CPU): the cases of the
reference's tests/unit/test_run_contract.py, over run_contract here."""
import json
from pathlib import Path

import numpy as np
import pytest


@pytest.fixture(scope="module")
def rc(pkg):
    return pkg.run_contract


def _workspace(rc, tmp_path, kind="subset_2d", run_id="fixed-run"):
    ws = rc.RunWorkspace.create(tmp_path / "run", kind, repository_root=Path.cwd(), run_id=run_id,
                                started_at="2026-01-01T00:00:00Z")
    f = ws.root / "result" / "displacements.csv"
    f.parent.mkdir(parents=True, exist_ok=True)
    f.write_text("x,y,valid\n1,2,1\n", encoding="utf-8")
    return ws, rc.WorkflowRunResult(kind, ws.root, {"displacements": [f]})


def test_create_writes_running_status(rc, tmp_path):
    ws = rc.RunWorkspace.create(tmp_path / "r", "subset_2d", run_id="a.b-c_1")
    st = json.loads(ws.status_path.read_text())
    assert st["execution_status"] == "RUNNING" and st["schema_version"] == rc.SCHEMA_VERSION
    assert st["finished_at"] is None and st["completed_stages"] == [] and st["run_id"] == "a.b-c_1"
    assert rc.validate_run_contract is not None and not ws.manifest_path.exists()


def test_four_files_written_and_loaded(rc, tmp_path):
    ws, res = _workspace(rc, tmp_path)
    out = rc.finalize_run_contract(ws, res, finished_at="2026-01-01T00:00:01Z")
    assert set(out) == {"manifest", "status", "metrics", "result"}
    for name in rc.CONTRACT_FILES:
        assert (ws.root / name).is_file()
    files = rc.load_run(ws.root)
    assert files["status"]["execution_status"] == "SUCCESS" and files["status"]["duration_seconds"] == 1.0
    assert files["manifest"]["run_id"] == files["result"]["run_id"] == "fixed-run"
    d = files["manifest"]["artifacts"][0]
    assert d["path_scope"] == "workspace" and d["path"] == "result/displacements.csv"
    assert d["id"] == "displacements:result__displacements.csv" and d["kind"] == "displacement_field"
    assert d["stage"] == "subset_solve" and d["required"] and d["exists"] and d["format"] == "csv"
    assert d["size_bytes"] == 16 and len(d["sha256"]) == 64
    assert files["result"]["primary_artifacts"] == [d["id"]] and files["result"]["primary_result_type"] == "2d_displacement"
    assert files["status"]["completed_stages"] == ["subset_solve"]


def test_status_vocabulary_and_warnings(rc, tmp_path):
    assert rc.FINAL_STATUSES == {"SUCCESS", "SUCCESS_WITH_WARNINGS", "PARTIAL_SUCCESS", "VALIDATION_FAILED",
                                 "EXECUTION_FAILED"}
    ws, res = _workspace(rc, tmp_path)
    rc.finalize_run_contract(ws, res, warnings=[{"code": "TEST_WARNING", "message": "quality concern"}],
                             finished_at="2026-01-01T00:00:01Z")
    st = json.loads(ws.status_path.read_text())
    assert st["execution_status"] == "SUCCESS_WITH_WARNINGS" and st["quality_status"] == "QUALITY_WARNINGS"
    with pytest.raises(rc.RunContractError):
        rc.finalize_run_contract(ws, res, execution_status="RUNNING")


def test_partial_and_validation_statuses_kept(rc, tmp_path):
    for i, status in enumerate(("PARTIAL_SUCCESS", "VALIDATION_FAILED")):
        ws, res = _workspace(rc, tmp_path / str(i))
        rc.finalize_run_contract(ws, res, execution_status=status, finished_at="2026-01-01T00:00:01Z")
        assert json.loads(ws.status_path.read_text())["execution_status"] == status


def test_execution_failure(rc, tmp_path):
    ws = rc.RunWorkspace.create(tmp_path / "failed", "stereo_3d", run_id="failed-run")
    rc.write_execution_failure(ws, RuntimeError("solver failed"), stage="reconstruction")
    st = json.loads(ws.status_path.read_text())
    assert st["execution_status"] == "EXECUTION_FAILED" and st["failed_stage"] == "reconstruction"
    assert st["errors"][0]["exception_type"] == "RuntimeError" and st["errors"][0]["message"] == "solver failed"


def test_missing_required_artifact_fails(rc, tmp_path):
    ws = rc.RunWorkspace.create(tmp_path / "missing", "subset_2d", run_id="missing-run")
    res = rc.WorkflowRunResult("subset_2d", ws.root, {"displacements": [ws.root / "result" / "missing.csv"]})
    rc.finalize_run_contract(ws, res, finished_at="2026-01-01T00:00:01Z")
    st = json.loads(ws.status_path.read_text())
    assert st["execution_status"] == "EXECUTION_FAILED" and st["failed_stage"] == "contract_finalization"
    assert st["errors"][0]["code"] == "REQUIRED_ARTIFACT_MISSING"


def test_schema_and_run_id_fail_closed(rc, tmp_path):
    with pytest.raises(rc.RunContractError):
        rc.RunWorkspace.create(tmp_path / "bad", "subset_2d", run_id="bad run")
    ws, res = _workspace(rc, tmp_path)
    rc.finalize_run_contract(ws, res, finished_at="2026-01-01T00:00:01Z")
    st = json.loads(ws.status_path.read_text())
    st["schema_version"] = "unsupported"
    ws.status_path.write_text(json.dumps(st), encoding="utf-8")
    with pytest.raises(rc.RunContractError, match="unsupported run contract schema"):
        rc.validate_run_contract(ws.root)


def test_result_references_checked(rc, tmp_path):
    ws, res = _workspace(rc, tmp_path)
    rc.finalize_run_contract(ws, res, finished_at="2026-01-01T00:00:01Z")
    r = json.loads(ws.result_path.read_text())
    r["primary_artifacts"].append("missing-artifact")
    ws.result_path.write_text(json.dumps(r), encoding="utf-8")
    with pytest.raises(rc.RunContractError, match="unknown artifacts"):
        rc.validate_run_contract(ws.root)


def test_subset_metrics_and_quality(pkg, rc, tmp_path):
    """metrics.json from the frames' stats.json; zero valid points and non-finite
    metrics are quality warnings."""
    ws = rc.RunWorkspace.create(tmp_path / "run", "subset_2d", run_id="m")
    n = 30
    rng = np.random.default_rng(4)
    res = {k: rng.normal(size=n) for k in ("x", "y", "u", "v", "du_dx", "du_dy", "dv_dx", "dv_dy", "correlation")}
    arts = {"displacements": [], "stats": []}
    for name, valid in (("002_znssd_icgn_1st", np.ones(n, bool)), ("003_znssd_icgn_1st", np.arange(n) < 10)):
        d = ws.root / name
        d.mkdir()
        r = dict(res, valid=valid)
        pkg.workflow.save_subset_csv(r, d / "displacements.csv")
        pkg.workflow.write_stats(r, d / "stats.json")
        arts["displacements"].append(d / "displacements.csv")
        arts["stats"].append(d / "stats.json")
    out = rc.finalize_run_contract(ws, rc.WorkflowRunResult("subset_2d", ws.root, arts))
    m = out["metrics"]
    assert m["total_points"] == 60 and m["valid_points"] == 40 and m["valid_fraction"] == 40 / 60
    assert set(m["frames"]) == {"002_znssd_icgn_1st", "003_znssd_icgn_1st"} and m["metric_provenance"] == ["stats.json"]
    assert m["u_mean"] == m["frames"]["002_znssd_icgn_1st"]["u_mean"]  # the first frame's statistics
    assert out["status"]["execution_status"] == "SUCCESS" and out["result"]["secondary_artifacts"] == [
        "stats:002_znssd_icgn_1st__stats.json", "stats:003_znssd_icgn_1st__stats.json"]
    assert rc.quality_warnings({"valid_fraction": 0.0})[0]["code"] == "ZERO_VALID_POINTS"
    w = rc.quality_warnings({"a": {"b": [1.0, float("nan")]}})
    assert w == [{"code": "NONFINITE_METRIC", "message": "non-finite metric at a.b[1]", "stage": "quality"}]
