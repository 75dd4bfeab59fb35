"""Run metadata of the 2D Subset-DIC workflow: manifest.json, status.json,
metrics.json and result.json beside the frame folders.

Mirrors the reference's F4 run contract for the subset_2d workflow (paths
inside the Traditional-DIC tree): the four file names and the schema version
(python/traditional_dic/run_contract.py:25-37), RunWorkspace.create writing a
RUNNING status (:233-270), the status record (:289-317), the artifact
descriptors (:205-230, :320-364), the subset_2d metrics (:392-411), the
quality warnings (:473-491), finalize_run_contract's status rules and the
four payloads (:502-622), write_execution_failure (:625-651) and
validate_run_contract (:662-718).  Scientific arrays are never rewritten: the
contract only describes the files the workflow wrote.
"""
from __future__ import annotations

import hashlib
import json
import math
import os
import re
import tempfile
import uuid
from dataclasses import dataclass, field
from datetime import datetime, timezone
from pathlib import Path
from typing import Any, Iterable, Mapping, Sequence

SCHEMA_VERSION = "1.0"
FINAL_STATUSES = frozenset({"SUCCESS", "SUCCESS_WITH_WARNINGS", "PARTIAL_SUCCESS", "VALIDATION_FAILED",
                            "EXECUTION_FAILED"})
CONTRACT_FILES = ("manifest.json", "status.json", "metrics.json", "result.json")
WORKFLOW_KINDS = ("subset_2d", "mesh_2d", "stereo_3d", "multiview_3d")
PRIMARY_RESULT_TYPE = {"subset_2d": "2d_displacement", "mesh_2d": "2d_mesh_displacement",
                       "stereo_3d": "stereo_3d_displacement", "multiview_3d": "stitched_3d_surface"}
# What this build runs: the subset solver path only (the reference also offers
# mesh_2d and the 3-D workflows).
CAPABILITY_CONTRACT = {"subset_2d": True, "mesh_2d": False, "stereo_3d_solver": None, "multiview_3d_solver": None}

_RUN_ID = re.compile(r"[A-Za-z0-9][A-Za-z0-9._-]{0,127}")
_HASH_LIMIT = 8 << 20  # files above 8 MiB are described without a digest


class RunContractError(ValueError):
    """Malformed or inconsistent run metadata."""


@dataclass
class WorkflowRunResult:
    """Where a workflow wrote its outputs, and the contract files describing them."""
    workflow_kind: str
    output_root: Path
    artifacts: Mapping[str, Any] = field(default_factory=dict)
    warnings: tuple = ()
    run_id: str | None = None
    manifest_path: Path | None = None
    status_path: Path | None = None
    metrics_path: Path | None = None
    result_path: Path | None = None
    contract: Mapping[str, Any] | None = None


def utc_now() -> str:
    return datetime.now(timezone.utc).replace(microsecond=0).isoformat().replace("+00:00", "Z")


def _write_json(path: Path, payload: Mapping[str, Any]) -> None:
    """Atomic: written to a temporary file in the same folder, then renamed."""
    path.parent.mkdir(parents=True, exist_ok=True)
    data = (json.dumps(payload, indent=2, sort_keys=True, default=str) + "\n").encode("utf-8")
    fd, tmp = tempfile.mkstemp(prefix=f".{path.name}.", suffix=".tmp", dir=str(path.parent))
    try:
        with os.fdopen(fd, "wb") as f:
            f.write(data)
            f.flush()
            os.fsync(f.fileno())
        os.replace(tmp, path)
    except BaseException:
        if os.path.exists(tmp):
            os.unlink(tmp)
        raise


def _read_json(path: Path) -> dict[str, Any]:
    try:
        value = json.loads(path.read_text(encoding="utf-8"))
    except FileNotFoundError as exc:
        raise RunContractError(f"missing run contract file: {path}") from exc
    except (OSError, json.JSONDecodeError) as exc:
        raise RunContractError(f"malformed run contract file: {path}") from exc
    if not isinstance(value, dict):
        raise RunContractError(f"run contract file must contain an object: {path}")
    if value.get("schema_version") != SCHEMA_VERSION:
        raise RunContractError(f"unsupported run contract schema in {path}: {value.get('schema_version')!r}")
    return value


def check_run_id(run_id: str) -> str:
    if not isinstance(run_id, str) or not _RUN_ID.fullmatch(run_id):
        raise RunContractError("run_id must match [A-Za-z0-9][A-Za-z0-9._-]{0,127}")
    return run_id


@dataclass
class RunWorkspace:
    root: Path
    run_id: str
    workflow_kind: str
    repository_root: Path
    started_at: str

    @classmethod
    def create(cls, root, workflow_kind: str, *, repository_root=None, run_id: str | None = None,
               started_at: str | None = None) -> "RunWorkspace":
        """Makes the folder and writes a RUNNING status.json."""
        path = Path(root).resolve()
        rid = check_run_id(run_id or f"run-{uuid.uuid4().hex}")
        path.mkdir(parents=True, exist_ok=True)
        ws = cls(path, rid, str(workflow_kind), Path(repository_root or path).resolve(), started_at or utc_now())
        ws.write_status(status_record(ws, "RUNNING"))
        return ws

    manifest_path = property(lambda self: self.root / "manifest.json")
    status_path = property(lambda self: self.root / "status.json")
    metrics_path = property(lambda self: self.root / "metrics.json")
    result_path = property(lambda self: self.root / "result.json")

    def write_status(self, payload: Mapping[str, Any]) -> None:
        _write_json(self.status_path, payload)


def status_record(ws: RunWorkspace, execution_status: str, *, finished_at: str | None = None,
                  duration_seconds: float | None = None, completed_stages: Sequence[str] = (),
                  failed_stage: str | None = None, errors: Sequence[Mapping[str, Any]] = (),
                  warnings: Sequence[Mapping[str, Any]] = (), quality_status: str = "UNKNOWN") -> dict[str, Any]:
    return {
        "schema_version": SCHEMA_VERSION,
        "run_id": ws.run_id,
        "workflow_kind": ws.workflow_kind,
        "execution_status": execution_status,
        "quality_status": quality_status,
        "started_at": ws.started_at,
        "finished_at": finished_at,
        "duration_seconds": duration_seconds,
        "completed_stages": list(dict.fromkeys(map(str, completed_stages))),
        "failed_stage": failed_stage,
        "errors": [dict(e) for e in errors],
        "warnings": [dict(w) for w in warnings],
        "updated_at": finished_at or utc_now(),
    }


# ---- artifact descriptors ----

def _paths(value: Any) -> Iterable[Path]:
    if isinstance(value, (str, Path)):
        yield Path(value)
    elif isinstance(value, Mapping):
        for v in value.values():
            yield from _paths(v)
    elif isinstance(value, Sequence) and not isinstance(value, (bytes, bytearray)):
        for v in value:
            yield from _paths(v)


def _slug(text: str) -> str:
    return re.sub(r"[^A-Za-z0-9_.-]+", "_", text).strip("_") or "artifact"


def _where(path: Path, ws: RunWorkspace) -> tuple[str, str]:
    """Portable path: relative to the workspace, else to the repository, else absolute."""
    for base, scope in ((ws.root, "workspace"), (ws.repository_root, "repository")):
        try:
            return path.relative_to(base).as_posix(), scope
        except ValueError:
            pass
    return str(path), "absolute"


def artifact_kind(key: str, path: Path) -> str:
    """Kind of a file of the 2D Subset-DIC workflow (the reference's rules for these keys)."""
    k, name, suffix = key.lower(), path.name.lower(), path.suffix.lower()
    if k == "figures" or suffix in (".png", ".jpg", ".jpeg", ".svg"):
        return "figure"
    if "mask" in k or "mask" in name:
        return "pair_mask"
    if k in ("fields", "pairwise_2d") or "displacement" in k or "displacements" in name:
        return "displacement_field"
    if "strain" in k or "strain" in name:
        return "strain_field"
    if "summary" in k or "summary" in name or suffix == ".json":
        return "summary"
    return "workflow_artifact"


def artifact_stage(key: str, path: Path) -> str:
    k = key.lower()
    if k in ("displacements", "stats"):
        return "subset_solve"
    if k == "strains":
        return "strain"
    if k == "figures":
        return "visualization"
    if "mask" in k or "mask" in path.name.lower():
        return "masks"
    return "write_outputs"


def _digest(path: Path) -> str | None:
    if not path.is_file() or path.stat().st_size > _HASH_LIMIT:
        return None
    return hashlib.sha256(path.read_bytes()).hexdigest()


def describe_artifacts(ws: RunWorkspace, artifacts: Mapping[str, Any]) -> list[dict[str, Any]]:
    """One descriptor per file (folders are expanded; the contract files themselves are skipped)."""
    out, seen = [], set()
    for key in sorted(artifacts):
        for p in _paths(artifacts[key]):
            p = (p if p.is_absolute() else ws.root / p).resolve()
            files = sorted(c for c in p.rglob("*") if c.is_file()) if p.is_dir() else [p]
            for f in files:
                f = f.resolve()
                if f in seen or f.name in CONTRACT_FILES:
                    continue
                seen.add(f)
                rel, scope = _where(f, ws)
                kind = artifact_kind(key, f)
                exists = f.is_file()
                d = {"id": f"{_slug(key)}:{_slug(rel.replace('/', '__'))}", "kind": kind, "path": rel,
                     "path_scope": scope, "format": f.suffix.lower().lstrip(".") or "file",
                     "stage": artifact_stage(key, f), "required": key.lower() != "figures" and kind != "figure",
                     "exists": exists, "description": f"{key} artifact"}
                if exists:
                    d["size_bytes"] = f.stat().st_size
                    h = _digest(f)
                    if h is not None:
                        d["sha256"] = h
                out.append(d)
    return sorted(out, key=lambda d: d["id"])


# ---- metrics and quality ----

def subset_metrics(ws: RunWorkspace) -> dict[str, Any]:
    """metrics.json of a subset_2d run: the first frame's stats.json, every frame's
    under "frames", and the point totals over the frames."""
    m: dict[str, Any] = {"schema_version": SCHEMA_VERSION, "run_id": ws.run_id, "workflow_kind": ws.workflow_kind,
                         "metric_provenance": []}
    stats = []
    for p in sorted(ws.root.rglob("stats.json")):
        try:
            v = json.loads(p.read_text(encoding="utf-8"))
        except (OSError, json.JSONDecodeError):
            continue
        if isinstance(v, dict):
            stats.append((p, v))
    if stats:
        m["frames"] = {p.parent.name: v for p, v in stats}
        m.update(stats[0][1])
        total = sum(int(v.get("total_points", 0)) for _, v in stats)
        valid = sum(int(v.get("valid_points", 0)) for _, v in stats)
        m["total_points"], m["valid_points"] = total, valid
        m["valid_fraction"] = float(valid / total) if total else 0.0
        m["metric_provenance"].append("stats.json")
    return m


def quality_warnings(metrics: Mapping[str, Any]) -> list[dict[str, Any]]:
    out = []
    vf = metrics.get("valid_fraction")
    if vf is not None and float(vf) == 0.0:
        out.append({"code": "ZERO_VALID_POINTS", "message": "no valid scientific points were reported",
                    "stage": "quality"})

    def visit(v: Any, where: str) -> None:
        if isinstance(v, float) and not math.isfinite(v):
            out.append({"code": "NONFINITE_METRIC", "message": f"non-finite metric at {where}", "stage": "quality"})
        elif isinstance(v, Mapping):
            for k, c in v.items():
                visit(c, f"{where}.{k}" if where else str(k))
        elif isinstance(v, Sequence) and not isinstance(v, (str, bytes, bytearray)):
            for i, c in enumerate(v):
                visit(c, f"{where}[{i}]")

    visit(metrics, "")
    return out


# ---- the four files ----

def finalize_run_contract(ws: RunWorkspace, workflow_result: WorkflowRunResult, *,
                          warnings: Sequence[Mapping[str, Any]] = (), execution_status: str = "SUCCESS",
                          failed_stage: str | None = None, errors: Sequence[Mapping[str, Any]] = (),
                          finished_at: str | None = None, config: Mapping[str, Any] | None = None
                          ) -> dict[str, Any]:
    """Writes metrics.json, result.json, manifest.json and the final status.json.
    A missing required artifact turns the run into EXECUTION_FAILED; quality or
    workflow warnings turn SUCCESS into SUCCESS_WITH_WARNINGS."""
    if execution_status not in FINAL_STATUSES:
        raise RunContractError(f"invalid final execution status: {execution_status!r}")
    desc = describe_artifacts(ws, workflow_result.artifacts)
    errors = list(errors)
    missing = [d["id"] for d in desc if d["required"] and not d["exists"]]
    if missing:
        execution_status = "EXECUTION_FAILED"
        errors.append({"code": "REQUIRED_ARTIFACT_MISSING", "message": "required workflow artifacts are missing",
                       "stage": "contract_finalization", "details": {"artifact_ids": missing}})
        failed_stage = failed_stage or "contract_finalization"
    metrics = subset_metrics(ws) if ws.workflow_kind == "subset_2d" else {
        "schema_version": SCHEMA_VERSION, "run_id": ws.run_id, "workflow_kind": ws.workflow_kind,
        "metric_provenance": []}
    warns = [dict(w) for w in warnings] + quality_warnings(metrics)
    quality = "QUALITY_WARNINGS" if warns else "QUALITY_OK"
    if execution_status == "SUCCESS" and warns:
        execution_status = "SUCCESS_WITH_WARNINGS"
    if execution_status == "SUCCESS" and workflow_result.warnings:
        warns += [{"code": "WORKFLOW_WARNING", "message": str(w), "stage": "workflow"}
                  for w in workflow_result.warnings]
        execution_status, quality = "SUCCESS_WITH_WARNINGS", "QUALITY_WARNINGS"
    finished = finished_at or utc_now()
    try:
        t0 = datetime.fromisoformat(ws.started_at.replace("Z", "+00:00"))
        t1 = datetime.fromisoformat(finished.replace("Z", "+00:00"))
        duration = max(0.0, (t1 - t0).total_seconds())
    except ValueError:
        duration = None
    status = status_record(ws, execution_status, finished_at=finished, duration_seconds=duration,
                           completed_stages=sorted({d["stage"] for d in desc if d["exists"]}),
                           failed_stage=failed_stage, errors=errors, warnings=warns, quality_status=quality)
    metrics.update(quality_status=quality, quality_warnings=warns, status_file="status.json")
    figures = [d["id"] for d in desc if d["kind"] == "figure"]
    primary = [d["id"] for d in desc if d["required"] and d["kind"] not in ("figure", "summary", "pair_mask")]
    result = {
        "schema_version": SCHEMA_VERSION, "run_id": ws.run_id, "workflow_kind": ws.workflow_kind,
        "primary_result_type": PRIMARY_RESULT_TYPE.get(ws.workflow_kind, "workflow_output"),
        "primary_artifacts": primary,
        "secondary_artifacts": [d["id"] for d in desc if d["id"] not in primary and d["id"] not in figures],
        "figures": figures, "metrics_file": "metrics.json", "status_file": "status.json",
        "manifest_file": "manifest.json",
    }
    manifest = {
        "schema_version": SCHEMA_VERSION, "run_id": ws.run_id, "workflow_kind": ws.workflow_kind,
        "created_at": ws.started_at, "repository_root": str(ws.repository_root),
        "capability_contract": dict(CAPABILITY_CONTRACT),
        "workspace": {"root": ".", "artifacts": "artifacts", "figures": "figures", "logs": "logs"},
        "artifacts": desc, "figures": figures, "logs": [],
        "workflow_result": {"workflow_kind": workflow_result.workflow_kind, "output_root": "."},
    }
    if config is not None:
        manifest["config"] = dict(config)
    _write_json(ws.metrics_path, metrics)
    _write_json(ws.result_path, result)
    _write_json(ws.manifest_path, manifest)
    ws.write_status(status)
    return {"manifest": manifest, "status": status, "metrics": metrics, "result": result}


def write_execution_failure(ws: RunWorkspace, exc: BaseException, *, stage: str = "workflow_execution",
                            finished_at: str | None = None) -> None:
    """EXECUTION_FAILED status naming the stage and the exception; never masks the exception itself."""
    finished = finished_at or utc_now()
    err = {"code": "WORKFLOW_EXECUTION_FAILED", "message": str(exc) or type(exc).__name__, "stage": stage,
           "exception_type": type(exc).__name__}
    try:
        ws.write_status(status_record(ws, "EXECUTION_FAILED", finished_at=finished, failed_stage=stage,
                                      errors=[err]))
    except Exception:  # the scientific error is the one to report
        pass


def validate_run_contract(root) -> dict[str, Any]:
    """Reads the four files and checks identity, vocabulary, descriptors and references."""
    root = Path(root).resolve()
    files = {name[:-5]: _read_json(root / name) for name in CONTRACT_FILES}
    if len({str(v.get("run_id")) for v in files.values()}) != 1 or \
            len({str(v.get("workflow_kind")) for v in files.values()}) != 1:
        raise RunContractError("run contract files disagree on run_id or workflow_kind")
    man, st, res = files["manifest"], files["status"], files["result"]
    check_run_id(str(man.get("run_id", "")))
    if man.get("workflow_kind") not in WORKFLOW_KINDS:
        raise RunContractError(f"unsupported workflow_kind in run contract: {man.get('workflow_kind')!r}")
    required = {
        "manifest": {"schema_version", "run_id", "workflow_kind", "capability_contract", "artifacts"},
        "status": {"schema_version", "run_id", "workflow_kind", "execution_status", "errors", "warnings"},
        "metrics": {"schema_version", "run_id", "workflow_kind"},
        "result": {"schema_version", "run_id", "workflow_kind", "primary_artifacts", "secondary_artifacts",
                   "figures"},
    }
    for name, keys in required.items():
        gone = sorted(keys - set(files[name]))
        if gone:
            raise RunContractError(f"{name}.json is missing required fields: {gone}")
    if st.get("execution_status") not in FINAL_STATUSES | {"RUNNING"}:
        raise RunContractError(f"invalid execution_status: {st.get('execution_status')!r}")
    arts = man.get("artifacts")
    if not isinstance(arts, list):
        raise RunContractError("manifest.artifacts must be a list")
    by_id = {str(a.get("id")): a for a in arts if isinstance(a, Mapping)}
    if len(by_id) != len(arts):
        raise RunContractError("manifest artifact IDs must be unique")
    fields = {"id", "kind", "path", "path_scope", "format", "stage", "required", "exists"}
    for aid, a in by_id.items():
        gone = sorted(fields - set(a))
        if gone:
            raise RunContractError(f"artifact {aid!r} is missing fields: {gone}")
        if a.get("path_scope") not in ("workspace", "repository", "absolute"):
            raise RunContractError(f"artifact {aid!r} has invalid path_scope")
    if (res.get("manifest_file"), res.get("status_file"), res.get("metrics_file")) != \
            ("manifest.json", "status.json", "metrics.json"):
        raise RunContractError("result metadata file references are inconsistent")
    refs = list(res.get("primary_artifacts", [])) + list(res.get("secondary_artifacts", [])) + \
        list(res.get("figures", []))
    unknown = sorted(set(refs) - set(by_id))
    if unknown:
        raise RunContractError(f"result references unknown artifacts: {unknown}")
    if st.get("execution_status") != "RUNNING":
        repo = Path(str(man.get("repository_root", root)))
        repo = repo.resolve() if repo.is_absolute() else (root / repo).resolve()
        base = {"workspace": root, "repository": repo}
        absent = [a["id"] for a in arts if a.get("required") and a.get("exists") and not (
            (base[a["path_scope"]] / a["path"]) if a["path_scope"] != "absolute" else Path(a["path"])).is_file()]
        if absent:
            raise RunContractError(f"required artifacts are not present: {absent}")
    return files


def load_run(root) -> dict[str, Any]:
    """The four files of a finished run, validated."""
    return validate_run_contract(root)
