#!/usr/bin/env python3
"""Write the reference's own retained data as compact test fixtures.

Reads data files (meshes, site lists, field outputs, walk traces, summaries,
request argument lists) that the reference repository keeps under
benchmarks/compiled_bunny/output/ and benchmarks/submission_study/ and stores
them as small .npz / .json fixtures beside this script.  Nothing here
imports or runs reference code: the files are parsed as text.

Run once in a container where the reference tree is mounted:
    python tests/golden/make_fixtures.py [/root/reference]
"""

from __future__ import annotations

import csv
import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent


def read_ascii_ply(path: Path) -> tuple[np.ndarray, np.ndarray]:
    """ASCII triangle PLY, read the way load_ascii_ply reads it
    (benchmarks/compiled_bunny/src/wih_twos.cpp:437-493): the first three
    numbers of each vertex line and `3 a b c` face lines."""
    lines = path.read_text().splitlines()
    nv = nf = 0
    start = None
    for i, line in enumerate(lines):
        words = line.split()
        if words[:2] == ["element", "vertex"]:
            nv = int(words[2])
        elif words[:2] == ["element", "face"]:
            nf = int(words[2])
        elif words[:1] == ["end_header"]:
            start = i + 1
            break
    assert start is not None and nv > 0 and nf > 0
    vertices = np.array([[float(w) for w in lines[start + i].split()[:3]] for i in range(nv)])
    faces = np.array([[int(w) for w in lines[start + nv + i].split()[1:4]] for i in range(nf)],
                     dtype=np.uint32)
    return vertices, faces


def read_field(path: Path) -> dict[str, np.ndarray]:
    with path.open() as handle:
        rows = list(csv.DictReader(handle))
    return {
        "u": np.array([float(r["u"]) for r in rows]),
        "score": np.array([float(r["score"]) for r in rows]),
        "score_se": np.array([float(r["score_se"]) for r in rows]),
        "steps": np.array([int(r["steps"]) for r in rows], dtype=np.int64),
        "limit_hits": np.array([int(r["limit_hits"]) for r in rows], dtype=np.int64),
        "xyz": np.array([[float(r["x"]), float(r["y"]), float(r["z"])] for r in rows]),
    }


def main() -> None:
    root = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    out = root / "benchmarks" / "compiled_bunny" / "output"
    # Config.seed of the bunny study (benchmarks/stanford_bunny/run_bunny_benchmark.py:33).
    base_seed = 20260912

    for faces in (1000, 69674):
        v, f = read_ascii_ply(out / f"bunny-{faces:05d}-faces.ply")
        np.savez_compressed(HERE / f"bunny_{faces}.npz", vertices=v, faces=f)

    points = np.loadtxt(out / "evaluation-points.csv", delimiter=",")
    np.save(HERE / "bunny_sites.npy", points)

    trace = np.loadtxt(out / "walk-traces.csv", delimiter=",", skiprows=1)
    np.save(HERE / "bunny_walk_traces.npy", trace)

    fields = {}
    # field-reference.csv: 2048 walks, seed Config.seed + 720, repeat 0
    # (benchmarks/compiled_bunny/run_compiled_benchmark.py:670-675).
    fields["reference"] = (out / "field-reference.csv", 2048, base_seed + 720, 69674)
    # field-twos-NNNN.csv: seed Config.seed + 1000 + walks (:686-692).
    for walks in (4, 8, 16, 32, 64, 128, 256, 384, 512):
        fields[f"twos_{walks}"] = (out / f"field-twos-{walks:04d}.csv", walks,
                                   base_seed + 1000 + walks, 69674)
    # field-bunny-NNNNN-faces.csv: 32 walks, seed Config.seed + 1800 (:760-766).
    fields["faces_1000"] = (out / "field-bunny-01000-faces.csv", 32, base_seed + 1800, 1000)
    fields["faces_69674"] = (out / "field-bunny-69674-faces.csv", 32, base_seed + 1800, 69674)

    index = {}
    for name, (path, walks, seed, nfaces) in fields.items():
        data = read_field(path)
        assert np.array_equal(data["xyz"], points), name
        del data["xyz"]
        np.savez_compressed(HERE / f"field_{name}.npz", **data)
        index[name] = {"walks": walks, "seed": seed, "faces": nfaces, "repeat": 0,
                       "epsilon": 5e-4, "max_steps": 512, "source_y_slope": 0.65,
                       "source": str(path.relative_to(root))}

    summary = json.loads((out / "reference-summary.json").read_text())
    meta = {
        "fields": index,
        # walk-traces.csv: 12 traced sites, site = i * 384 / 12, seed
        # Config.seed + 720 + 900000 + site (wih_twos.cpp:1725-1744).
        "traces": {"sites": 12, "seed": base_seed + 720, "key_offset": 900000, "faces": 69674,
                   "epsilon": 5e-4, "max_steps": 512},
        "reference_summary": {k: summary[k] for k in ("mean_steps", "paths_per_second",
                                                      "distance_queries_per_second",
                                                      "runtime_seconds_median", "threads")},
        # Recorded Embree-vs-libigl start-point distance agreement
        # (benchmarks/compiled_bunny/output/summary.json).
        "embree_libigl_max_error": json.loads((out / "summary.json").read_text())[
            "maximum_embree_libigl_distance_error"],
    }
    (HERE / "bunny_meta.json").write_text(json.dumps(meta, indent=1) + "\n")


def study_arguments(request_json: Path) -> dict[str, str]:
    """Flag -> value of a retained wih_study command line (request.json)."""
    argv = json.loads(request_json.read_text())
    return {argv[i]: argv[i + 1] for i in range(1, len(argv) - 1) if argv[i].startswith("--")}


def study_fixtures(root: Path) -> None:
    """Paired and hot-only requests of the submission study on the branch mesh
    (benchmarks/submission_study/**/raw/*/: request.json, points.txt,
    holes.txt, field.csv).  Source configuration: the pilot and the
    '-v1-remote-sources' validation used the archived first version
    (pilot/source-config-v1/branch.txt, pilot/README.txt); the later runs the
    current assets/branch.txt, which adds a sensor self-heating bump."""
    study = root / "benchmarks" / "submission_study"
    v, f = read_ascii_ply(study / "assets" / "branch.ply")
    np.savez_compressed(HERE / "study_branch_mesh.npz", vertices=v, faces=f)
    def walk_requests(pattern: str) -> list[Path]:
        keep = []
        for rq in sorted(study.glob(pattern)):
            a = study_arguments(rq)
            if a["--mode"] == "twos" and a.get("--diff", "0") == "0" and a.get("--fd", "0") == "0":
                keep.append(rq)
        return keep

    picks = walk_requests("pilot/channels/branch/twos-1611/raw/round*-*/request.json")
    picks += walk_requests("pilot/pipeline/branch/twos-1611/raw/round*-rescore/request.json")
    for pattern in ("pilot/pipeline/branch/evaluation/1611/raw/*/request.json",
                    "validation-local-cavity/raw/*/request.json", "validation/raw/*/request.json",
                    "validation-v1-remote-sources/raw/*/request.json"):
        picks += walk_requests(pattern)[:2]
    arrays, requests = {}, []
    for k, rq in enumerate(picks):
        d = rq.parent
        a = study_arguments(rq)
        assert a["--mesh"].endswith("/branch.ply") and a["--config"].endswith("/branch.txt")
        rel = str(d.relative_to(study))
        points = np.loadtxt(d / "points.txt", ndmin=2).reshape(-1, 3)
        holes_text = (d / "holes.txt").read_text() if (d / "holes.txt").exists() else ""
        holes = np.array([float(x) for x in holes_text.split()]).reshape(-1, 3)
        with (d / "field.csv").open() as handle:
            rows = list(csv.DictReader(handle))
        assert [int(r["site"]) for r in rows] == list(range(len(points)))
        arrays[f"{k}_points"] = points
        arrays[f"{k}_holes"] = holes
        arrays[f"{k}_values"] = np.array([[float(r[c]) for c in ("u", "p", "score", "same_index")] for r in rows])
        arrays[f"{k}_counts"] = np.array([[int(r["steps"]), int(r["caps"])] for r in rows], dtype=np.int64)
        requests.append({"source": rel, "seed": int(a["--seed"]), "walks": int(a["--walks"]),
                         "epsilon": float(a["--epsilon"]), "max_steps": int(a["--max-steps"]),
                         "cavity_radius": float(a["--cavity-radius"]), "hot": a["--hot-only"] != "0",
                         "config": "v1" if rel.startswith("pilot/") or "-v1-" in rel else "current"})
    np.savez_compressed(HERE / "study_branch_requests.npz", **arrays)
    meta = {"mesh": "benchmarks/submission_study/assets/branch.ply",
            "configs": {"v1": (study / "pilot" / "source-config-v1" / "branch.txt").read_text(),
                        "current": (study / "assets" / "branch.txt").read_text()},
            "requests": requests}
    (HERE / "study_branch_meta.json").write_text(json.dumps(meta, indent=1) + "\n")


def diff_fixtures(root: Path) -> None:
    """Differential WoS requests (wih_study --diff 1) on the branch mesh: all
    twelve of validation-local-cavity (one cavity, 128 sites, offsets 0.002,
    0.004, 0.008) and the first two of validation (six cavities, 256 sites).
    Per request: sites, cavity centres, the field's per-cavity columns
    dc{k}_x,y,z, dc{k}_se_x,y,z, dR{k}, dR{k}_se, and kernel.json's request
    totals (steps = outer + nested steps, distance queries, parity rays, caps)."""
    study = root / "benchmarks" / "submission_study"
    picks = sorted(study.glob("validation-local-cavity/raw/gradient-*/request.json"))
    picks += sorted(study.glob("validation/raw/gradient-*/request.json"))[:2]
    arrays, requests = {}, []
    for k, rq in enumerate(picks):
        d = rq.parent
        a = study_arguments(rq)
        assert a["--diff"] == "1" and a["--mesh"].endswith("/branch.ply")
        assert a["--config"].endswith("assets/branch.txt") and a["--repeats"] == "1"
        points = np.loadtxt(d / "points.txt", ndmin=2).reshape(-1, 3)
        holes = np.loadtxt(d / "holes.txt", ndmin=2).reshape(-1, 3)
        with (d / "field.csv").open() as handle:
            rows = list(csv.DictReader(handle))
        assert [int(r["site"]) for r in rows] == list(range(len(points)))
        cols = [[f"dc{c}_x", f"dc{c}_y", f"dc{c}_z", f"dc{c}_se_x", f"dc{c}_se_y", f"dc{c}_se_z", f"dR{c}",
                 f"dR{c}_se"] for c in range(len(holes))]
        arrays[f"{k}_points"] = points
        arrays[f"{k}_holes"] = holes
        arrays[f"{k}_values"] = np.array([[[float(r[x]) for x in cs] for cs in cols] for r in rows])
        kernel = json.loads((d / "kernel.json").read_text())
        requests.append({"source": str(d.relative_to(study)), "seed": int(a["--seed"]), "walks": int(a["--walks"]),
                         "epsilon": float(a["--epsilon"]), "max_steps": int(a["--max-steps"]),
                         "cavity_radius": float(a["--cavity-radius"]), "offset": float(a["--diff-offset"]),
                         "config": "current",
                         "totals": {x: int(kernel[x]) for x in ("steps", "distance_queries", "parity_rays", "caps")}})
    np.savez_compressed(HERE / "study_branch_diff.npz", **arrays)
    (HERE / "study_branch_diff.json").write_text(json.dumps({"requests": requests}, indent=1) + "\n")


def cert_fixtures(root: Path) -> None:
    """Certified minimum-benefit validations (wih_twos --mode validate) of the
    pilot in benchmarks/minimum_benefit/pilot-floor-policy/{bunny,housing}/
    attempt-*/: the finalist (validation-points.csv), the cavities accepted
    before it (validation-holes.csv), the first 2,048 of its 65,536 samples
    (validation.csv: value, steps, capped, clipped) and the per-stage summary
    (validation.json).  Seeds follow the runner (CONFIG seed 2026091601 +
    100 step + 2, runner_snapshot_after_arithmetic_cleanup.py:213 with
    configuration.json), as do the caps: radius cap 0.1, value cap 0.06 (bunny,
    the 69,674-face watertight bunny = bunny_69674.npz, affine source slope
    0.65) or 0.18 (housing, the housing mesh, its chips), eps 1e-6, 128 steps."""
    base = root / "benchmarks" / "minimum_benefit" / "pilot-floor-policy"
    config = json.loads((base / "configuration.json").read_text())
    arrays, attempts = {}, []
    k = 0
    for kind in ("bunny", "housing"):
        for d in sorted((base / kind).glob("attempt-*")):
            step = int(d.name.split("-")[1])
            decision = json.loads((d / "decision.json").read_text())
            holes_path = d / "validation-holes.csv"
            holes = (np.loadtxt(holes_path, delimiter=",", ndmin=2).reshape(-1, 3) if holes_path.stat().st_size
                     else np.zeros((0, 3)))
            with (d / "validation.csv").open() as handle:
                rows = list(csv.DictReader(handle))
            assert len(rows) == 65536 and all(int(r["site"]) == 0 for r in rows)
            head = rows[:2048]
            summary = json.loads((d / "validation.json").read_text())
            arrays[f"{k}_point"] = np.loadtxt(d / "validation-points.csv", delimiter=",", ndmin=2).reshape(-1, 3)
            arrays[f"{k}_holes"] = holes
            arrays[f"{k}_value"] = np.array([float(r["value"]) for r in head])
            arrays[f"{k}_steps"] = np.array([int(r["steps"]) for r in head], dtype=np.int32)
            arrays[f"{k}_capped"] = np.array([int(r["capped"]) for r in head], dtype=np.int8)
            arrays[f"{k}_clipped"] = np.array([int(r["clipped"]) for r in head], dtype=np.int8)
            attempts.append({"source": str(d.relative_to(root)), "kind": kind, "step": step,
                             "seed": config["seed"] + 100 * step + 2, "radius": decision["radius"],
                             "value_cap": config["value_caps"][kind], "radius_cap": config["radius_cap"],
                             "epsilon": config["epsilon"], "max_steps": config["max_steps"], "walks": 65536,
                             "slope": 0.65, "stages": summary["stages"],
                             "containment_failures": summary["containment_failures"]})
            k += 1
    np.savez_compressed(HERE / "cert_requests.npz", **arrays)
    (HERE / "cert_meta.json").write_text(json.dumps({"attempts": attempts}, indent=1) + "\n")


def housing_fixtures(root: Path) -> None:
    """Scout and validation fields of the engineering-housing design loop
    (benchmarks/engineering_housing/output/field-rep-1-*.csv with their
    points/cavities files): wih_twos --mode twos --source-mode housing,
    eps 1e-3, max_steps 512, cavity radius 0.095, repeat 0.  Seeds follow
    run_housing_benchmark.py:826-869: replicate 1 of "TWoS + validation" /
    "TWoS equal-time" uses SEED + 100000, scout step k adds 1000 + k,
    validation step k adds 2000 + k (SEED = 20260912, :28); walk counts are
    that script's defaults (:1116-1120)."""
    housing = root / "benchmarks" / "engineering_housing"
    v, f = read_ascii_ply(housing / "assets" / "stylized-finned-housing.ply")
    np.savez_compressed(HERE / "housing_mesh.npz", vertices=v, faces=f)
    base = 20260912 + 100000
    picks = [("rep-1-twos-scout-01", base + 1001, 12), ("rep-1-twos-scout-03", base + 1003, 12),
             ("rep-1-twos-validate-02", base + 2002, 192), ("rep-1-twos-validate-05", base + 2005, 192),
             ("rep-1-twos-equal-scout-02", base + 1002, 96), ("rep-1-twos-equal-validate-04", base + 2004, 768)]
    arrays, requests = {}, []
    for k, (tag, seed, walks) in enumerate(picks):
        data = read_field(housing / "output" / f"field-{tag}.csv")
        cav_path = housing / "output" / f"cavities-{tag}.csv"
        cav = np.loadtxt(cav_path, delimiter=",", ndmin=2) if cav_path.exists() else np.zeros((0, 3))
        summary = json.loads((housing / "output" / f"summary-{tag}.json").read_text())
        assert summary["walks_per_site"] == walks and summary["cavities"] == len(cav)
        arrays[f"{k}_points"] = data["xyz"]
        arrays[f"{k}_cavities"] = cav
        arrays[f"{k}_values"] = np.c_[data["u"], data["score"], data["score_se"]]
        arrays[f"{k}_counts"] = np.c_[data["steps"], data["limit_hits"]]
        requests.append({"tag": tag, "seed": seed, "walks": walks, "epsilon": summary["epsilon"],
                         "cavity_radius": 0.095, "max_steps": 512})
    np.savez_compressed(HERE / "housing_requests.npz", **arrays)
    (HERE / "housing_meta.json").write_text(json.dumps({"requests": requests}, indent=1) + "\n")


def design_fixture(root: Path) -> None:
    """The pilot greedy insertion design on the branch mesh
    (benchmarks/submission_study/pilot/pipeline/branch/twos-1611/: design.json
    and its raw/round{k}-scout, -rescore requests): the candidate pool (round 0's
    scout sites), the chosen centres, and per round the scout and rescore
    seeds, walk counts, feasible-candidate counts, the shortlist as indices
    into that round's feasible candidates, and both rounds' score columns."""
    run = root / "benchmarks" / "submission_study" / "pilot" / "pipeline" / "branch" / "twos-1611"
    design = json.loads((run / "design.json").read_text())
    pool = np.loadtxt(run / "raw" / "round0-scout" / "points.txt", ndmin=2)
    rounds = []
    for k in range(len(design["centers"])):
        entry = {}
        for kind in ("scout", "rescore"):
            d = run / "raw" / f"round{k}-{kind}"
            a = study_arguments(d / "request.json")
            pts = np.loadtxt(d / "points.txt", ndmin=2)
            with (d / "field.csv").open() as handle:
                rows = list(csv.DictReader(handle))
            entry[kind] = {"seed": int(a["--seed"]), "walks": int(a["--walks"]), "sites": len(pts),
                           "score": [float(r["score"]) for r in rows]}
            if kind == "scout":
                scout_pts = pts
            else:
                entry["shortlist"] = [int(np.nonzero(np.all(scout_pts == q, axis=1))[0][0]) for q in pts]
        entry["radius"] = float(a["--cavity-radius"])
        entry["epsilon"] = float(a["--epsilon"])
        entry["max_steps"] = int(a["--max-steps"])
        rounds.append(entry)
    out = {"source": str(run.relative_to(root)), "seed": design["seed"], "radius": design["radius"],
           "protocol": design["protocol"], "centers": design["centers"], "pool": pool.tolist(), "rounds": rounds,
           "config": (root / "benchmarks" / "submission_study" / "pilot" / "source-config-v1" /
                      "branch.txt").read_text()}
    (HERE / "study_branch_design.json").write_text(json.dumps(out) + "\n")


if __name__ == "__main__":
    main()
    study_fixtures(Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference"))
    housing_fixtures(Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference"))
    design_fixture(Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference"))
    diff_fixtures(Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference"))
    cert_fixtures(Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference"))
