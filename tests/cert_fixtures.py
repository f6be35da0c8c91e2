"""Loader of the certified minimum-benefit fixtures (tests/golden/cert_*):
the pilot's --mode validate calls on the watertight bunny and the housing."""

from __future__ import annotations

import json
from dataclasses import dataclass
from pathlib import Path

import numpy as np


@dataclass
class CertAttempt:
    source: str
    kind: str             # "bunny" (affine source) or "housing" (chips)
    point: np.ndarray     # [1, 3] the finalist
    holes: np.ndarray     # [k, 3] cavities accepted before it
    radius: float
    seed: int
    walks: int
    epsilon: float
    max_steps: int
    radius_cap: float
    value_cap: float
    slope: float
    value: np.ndarray     # the first 2,048 samples
    steps: np.ndarray
    capped: np.ndarray
    clipped: np.ndarray
    stages: list          # per stage n: mean, mean_steps, clipped, path_caps, clearances
    containment_failures: int


def load(golden: Path):
    meta = json.loads((golden / "cert_meta.json").read_text())
    arrays = np.load(golden / "cert_requests.npz")
    out = []
    for k, a in enumerate(meta["attempts"]):
        out.append(CertAttempt(source=a["source"], kind=a["kind"], point=arrays[f"{k}_point"],
                               holes=arrays[f"{k}_holes"], radius=a["radius"], seed=a["seed"], walks=a["walks"],
                               epsilon=a["epsilon"], max_steps=a["max_steps"], radius_cap=a["radius_cap"],
                               value_cap=a["value_cap"], slope=a["slope"], value=arrays[f"{k}_value"],
                               steps=arrays[f"{k}_steps"], capped=arrays[f"{k}_capped"],
                               clipped=arrays[f"{k}_clipped"], stages=a["stages"],
                               containment_failures=a["containment_failures"]))
    return out


def mesh(golden: Path, kind: str):
    z = np.load(golden / ("bunny_69674.npz" if kind == "bunny" else "housing_mesh.npz"))
    return z["vertices"], z["faces"]
