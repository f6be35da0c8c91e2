"""Loader of the submission-study fixtures (tests/golden/study_branch_*).

Each request is one retained wih_study call on the branch mesh
(benchmarks/submission_study/**/raw/*/): sites, cavity centres, seed, walk
count, eps, max_steps, and the reference's output columns
u, p, score, same_index (binary64) and steps, caps (integers).
"""

from __future__ import annotations

import json
from dataclasses import dataclass
from pathlib import Path

import numpy as np


@dataclass
class StudyRequest:
    source: str
    points: np.ndarray
    holes: np.ndarray
    values: np.ndarray   # [n, 4]: u, p, score, same_index
    counts: np.ndarray   # [n, 2]: steps, caps
    seed: int
    walks: int
    epsilon: float
    max_steps: int
    cavity_radius: float
    hot: bool
    config: str          # source configuration text (sensor line, heat bumps)


def load(golden: Path):
    meta = json.loads((golden / "study_branch_meta.json").read_text())
    mesh = np.load(golden / "study_branch_mesh.npz")
    arrays = np.load(golden / "study_branch_requests.npz")
    requests = []
    for k, r in enumerate(meta["requests"]):
        requests.append(StudyRequest(
            source=r["source"], points=arrays[f"{k}_points"], holes=arrays[f"{k}_holes"],
            values=arrays[f"{k}_values"], counts=arrays[f"{k}_counts"], seed=r["seed"], walks=r["walks"],
            epsilon=r["epsilon"], max_steps=r["max_steps"], cavity_radius=r["cavity_radius"], hot=r["hot"],
            config=meta["configs"][r["config"]]))
    return mesh["vertices"], mesh["faces"], requests


def prefix(req: StudyRequest, n: int) -> StudyRequest:
    """The request restricted to its first n sites (a site's streams depend on
    its index, so a prefix keeps every remaining site's values)."""
    return StudyRequest(**{**req.__dict__, "points": req.points[:n], "values": req.values[:n],
                           "counts": req.counts[:n]})


def compare(req: StudyRequest, u, p, score, same, steps, caps, rtol: float) -> np.ndarray:
    """Integer columns equal; value columns equal within rtol (relative to the
    larger of |reference| and 1e-3 times the column's largest |value|).
    Returns, per site, whether all four values are bitwise equal.

    Why not every site is bitwise: the reference's point-query callback keeps a
    triangle whose binary64 distance is below the binary32 rounding of the best
    distance so far (benchmarks/compiled_bunny/src/wih_twos.cpp:416-430), so
    among triangles within binary32 rounding of each other (typically the two
    faces of an edge the point is closest to) it returns the last one Embree
    visits, not the minimum; the one-ulp difference in the radius then grows
    along the path (a WoS step can at most double a position perturbation)."""
    np.testing.assert_array_equal(steps, req.counts[:, 0])
    np.testing.assert_array_equal(caps, req.counts[:, 1])
    ours = np.c_[u, p, score, same]
    ref = req.values
    for c in range(4):
        scale = np.maximum(np.abs(ref[:, c]), 1e-3 * np.max(np.abs(ref[:, c])))
        err = np.abs(ours[:, c] - ref[:, c])
        assert np.all(err <= rtol * scale), (req.source, c, np.max(err / np.maximum(scale, 1e-300)))
    return np.all(ours == ref, axis=1)


@dataclass
class HousingRequest:
    tag: str
    points: np.ndarray
    cavities: np.ndarray
    values: np.ndarray   # [n, 3]: u, score, score_se
    counts: np.ndarray   # [n, 2]: steps, limit_hits
    seed: int
    walks: int
    epsilon: float
    max_steps: int
    cavity_radius: float


def load_housing(golden: Path):
    """Scout/validation fields of the engineering-housing design loop
    (wih_twos --source-mode housing; tests/golden/housing_*)."""
    meta = json.loads((golden / "housing_meta.json").read_text())
    mesh = np.load(golden / "housing_mesh.npz")
    arrays = np.load(golden / "housing_requests.npz")
    requests = [HousingRequest(tag=r["tag"], points=arrays[f"{k}_points"], cavities=arrays[f"{k}_cavities"],
                               values=arrays[f"{k}_values"], counts=arrays[f"{k}_counts"], seed=r["seed"],
                               walks=r["walks"], epsilon=r["epsilon"], max_steps=r["max_steps"],
                               cavity_radius=r["cavity_radius"])
                for k, r in enumerate(meta["requests"])]
    return mesh["vertices"], mesh["faces"], requests


def compare_housing(req: HousingRequest, u, score, score_se, steps, caps, rtol: float) -> np.ndarray:
    """Step and cap counts equal; u, score, score_se within rtol; returns the
    per-site bitwise agreement of the three values."""
    np.testing.assert_array_equal(steps, req.counts[:, 0])
    np.testing.assert_array_equal(caps, req.counts[:, 1])
    ours = np.c_[u, score, score_se]
    for c in range(3):
        ref = req.values[:, c]
        scale = np.maximum(np.abs(ref), 1e-3 * np.max(np.abs(ref)))
        err = np.abs(ours[:, c] - ref)
        assert np.all(err <= rtol * scale), (req.tag, c, np.max(err / scale))
    return np.all(ours == req.values, axis=1)


@dataclass
class DiffRequest:
    source: str
    points: np.ndarray
    holes: np.ndarray
    values: np.ndarray   # [n, n_cavities, 8]: dc x y z, dc se x y z, dR, dR se
    totals: dict         # request totals: steps, distance_queries, parity_rays, caps
    seed: int
    walks: int
    epsilon: float
    max_steps: int
    cavity_radius: float
    offset: float
    config: str


def load_diff(golden: Path):
    """Differential WoS requests (wih_study --diff 1) on the branch mesh
    (tests/golden/study_branch_diff.*): the mesh and the requests."""
    meta = json.loads((golden / "study_branch_meta.json").read_text())
    mesh = np.load(golden / "study_branch_mesh.npz")
    arrays = np.load(golden / "study_branch_diff.npz")
    reqs = json.loads((golden / "study_branch_diff.json").read_text())["requests"]
    out = [DiffRequest(source=r["source"], points=arrays[f"{k}_points"], holes=arrays[f"{k}_holes"],
                       values=arrays[f"{k}_values"], totals=r["totals"], seed=r["seed"], walks=r["walks"],
                       epsilon=r["epsilon"], max_steps=r["max_steps"], cavity_radius=r["cavity_radius"],
                       offset=r["offset"], config=meta["configs"][r["config"]])
           for k, r in enumerate(reqs)]
    return mesh["vertices"], mesh["faces"], out


def compare_diff(req: DiffRequest, values, rtol: float) -> np.ndarray:
    """Every derivative column within rtol, relative to the larger of
    |reference| and 1e-3 times that column's largest |value| over the request;
    returns per site whether all of its values are bitwise equal (near-tie
    closest triangles move a few paths by ulps, see compare)."""
    ref = req.values
    assert values.shape == ref.shape, (values.shape, ref.shape)
    scale = np.maximum(np.abs(ref), 1e-3 * np.max(np.abs(ref), axis=0, keepdims=True))
    err = np.abs(values - ref)
    bad = err > rtol * scale
    assert not np.any(bad), (req.source, np.argwhere(bad)[:4], np.max(err / np.maximum(scale, 1e-300)))
    return np.all(values == ref, axis=(1, 2))
