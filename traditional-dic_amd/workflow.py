"""The 2D Subset-DIC workflow's result files, over the GPU pipeline.

Reference interfaces mirrored (paths inside the Traditional-DIC tree):
  read_gray / read_mask         python/traditional_dic/workflows/subset_2d.py:21-31
  save_subset_csv               workflows/subset_2d.py:34-54 (displacements.csv)
  write_stats                   workflows/subset_2d.py:94-113 (stats.json)
  subset_method_tag             python/traditional_dic/config.py:77-92
  the per-frame loop            workflows/subset_2d.py:152-227 (result/<frame>_<tag>/)
  derive_metrics (subset_2d)    python/traditional_dic/run_contract.py:392-411
  the run contract files        run_contract.py (module run_contract here)

The figures, legends and the least-squares strain post-processing of the
reference workflow are not written here.
"""
from __future__ import annotations

import csv
import json
from pathlib import Path
from typing import Any

import numpy as np

DISPLACEMENT_HEADER = ["x", "y", "u", "v", "du_dx", "du_dy", "dv_dx", "dv_dy", "correlation", "valid"]


def read_gray(path) -> np.ndarray:
    """float32 image divided by its maximum (workflows/subset_2d.py:21-27)."""
    from PIL import Image
    arr = np.array(Image.open(path).convert("F"), dtype=np.float32, copy=True)
    m = float(np.max(arr))
    if m > 0.0:
        arr /= m
    return arr


def read_mask(path) -> np.ndarray:
    """0/1 uint8 mask of the pixels with L > 0 (workflows/subset_2d.py:30-31)."""
    from PIL import Image
    return (np.asarray(Image.open(path).convert("L")) > 0).astype(np.uint8)


def save_subset_csv(result: dict, path) -> None:
    """displacements.csv: DISPLACEMENT_HEADER, one row per grid point, valid as 0/1."""
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    cols = [np.asarray(result[k]) for k in DISPLACEMENT_HEADER[:-1]]
    valid = np.asarray(result["valid"])
    with path.open("w", newline="", encoding="utf-8") as f:
        w = csv.writer(f)
        w.writerow(DISPLACEMENT_HEADER)
        for i in range(len(valid)):
            w.writerow([c[i] for c in cols] + [int(valid[i])])


def subset_stats(result: dict, metric: str = "znssd") -> dict[str, Any]:
    """The statistics of stats.json over the valid points."""
    valid = np.asarray(result["valid"], dtype=bool)
    u = np.asarray(result["u"], dtype=np.float64)[valid]
    v = np.asarray(result["v"], dtype=np.float64)[valid]
    q = np.asarray(result["correlation"], dtype=np.float64)[valid]
    mag = np.hypot(u, v)
    n = int(len(result["x"]))
    return {
        "total_points": n,
        "valid_points": int(np.sum(valid)),
        "invalid_points": int(n - np.sum(valid)),
        "u_mean": float(u.mean()) if u.size else 0.0,
        "u_std": float(u.std()) if u.size else 0.0,
        "v_mean": float(v.mean()) if v.size else 0.0,
        "v_std": float(v.std()) if v.size else 0.0,
        "mag_mean": float(mag.mean()) if mag.size else 0.0,
        "mag_max": float(mag.max()) if mag.size else 0.0,
        f"{metric}_mean": float(q.mean()) if q.size else 0.0,
        f"{metric}_max": float(q.max()) if q.size else 0.0,
    }


def write_stats(result: dict, path, metric: str = "znssd") -> dict[str, Any]:
    stats = subset_stats(result, metric)
    Path(path).write_text(json.dumps(stats, indent=2), encoding="utf-8")
    return stats


def _tag(value, aliases: dict[str, str]) -> str:
    text = str(value).strip().lower()
    return aliases.get(text, text)


def subset_method_tag(config: dict | None) -> str:
    """<criterion>_<solver>_<order>, e.g. znssd_icgn_1st (config.py:77-92)."""
    cfg = config or {}
    criterion = _tag((cfg.get("correlation", {}) or {}).get("criterion", "znssd"), {})
    solver = _tag((cfg.get("optimization", {}) or {}).get("method", "icgn"),
                  {"fgn": "fgn", "forward_gauss_newton": "fgn"})
    order = _tag((cfg.get("shape_function", {}) or {}).get("order", "1"),
                 {"first_order": "1st", "first": "1st", "1": "1st",
                  "second_order": "2nd", "second": "2nd", "2": "2nd"})
    return f"{criterion}_{solver}_{order}"


def run_subset_2d(reference_path, deformed_paths, roi_path, config: dict, result_root, precision: str = "fp64",
                  device: int = 0, run_id: str | None = None, repository_root=None):
    """For each deformed frame: subset() on the GPU, then
    result_root/<frame>_<tag>/displacements.csv and stats.json
    (workflows/subset_2d.py:152-227 without figures and strain), and the run
    contract's manifest.json / status.json / metrics.json / result.json in
    result_root (run_contract.py:779-839): status RUNNING while the frames are
    solved, EXECUTION_FAILED (and the exception re-raised) if a frame fails.
    Returns a WorkflowRunResult with the artifacts and the contract paths."""
    from .seeding import subset
    from . import run_contract as rc
    root = Path(result_root)
    ws = rc.RunWorkspace.create(root, "subset_2d", repository_root=repository_root, run_id=run_id)
    try:
        reference = read_gray(reference_path)
        roi = read_mask(roi_path)
        artifacts: dict[str, list[str]] = {"displacements": [], "stats": []}
        tag = subset_method_tag(config)
        criterion = str((config.get("correlation", {}) or {}).get("criterion", "znssd")).strip().lower()
        for dp in deformed_paths:
            out = root / (Path(dp).stem + (f"_{tag}" if tag else ""))
            out.mkdir(parents=True, exist_ok=True)
            result = subset(reference, read_gray(dp), config=config, roi=roi, precision=precision, device=device)
            save_subset_csv(result, out / "displacements.csv")
            artifacts["displacements"].append(str(out / "displacements.csv"))
            write_stats(result, out / "stats.json", metric="ssd" if criterion == "ssd" else "znssd")
            artifacts["stats"].append(str(out / "stats.json"))
    except Exception as exc:
        rc.write_execution_failure(ws, exc)
        raise
    wr = rc.WorkflowRunResult("subset_2d", root, artifacts)
    try:
        contract = rc.finalize_run_contract(ws, wr, config=config)
        rc.validate_run_contract(ws.root)
    except Exception as exc:
        rc.write_execution_failure(ws, exc, stage="contract_finalization")
        raise
    wr.run_id, wr.contract = ws.run_id, contract
    wr.manifest_path, wr.status_path = ws.manifest_path, ws.status_path
    wr.metrics_path, wr.result_path = ws.metrics_path, ws.result_path
    return wr


def derive_subset_metrics(stats_paths) -> dict[str, Any]:
    """metrics.json fields of a subset_2d run (run_contract.py:392-411): the first
    frame's statistics, per-frame statistics under "frames", and the totals
    total_points / valid_points / valid_fraction over the frames."""
    stats = [(Path(p), json.loads(Path(p).read_text(encoding="utf-8"))) for p in sorted(map(str, stats_paths))]
    payload: dict[str, Any] = {"workflow_kind": "subset_2d", "metric_provenance": []}
    if stats:
        payload["frames"] = {p.parent.name: v for p, v in stats}
        payload.update(stats[0][1])
        total = sum(int(v.get("total_points", 0)) for _, v in stats)
        valid = sum(int(v.get("valid_points", 0)) for _, v in stats)
        payload["total_points"] = total
        payload["valid_points"] = valid
        payload["valid_fraction"] = float(valid / total) if total else 0.0
        payload["metric_provenance"].append("stats.json")
    return payload
