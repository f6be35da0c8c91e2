"""Writers of the reference's input formats (ASCII PLY, points CSV) for the
drop-in executable tests, and a reader of their CSV outputs."""

from __future__ import annotations

import csv
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
BIN = ROOT / "walking-into-holes_amd" / "bin"


def write_ply(path: Path, vertices, faces) -> None:
    v = np.asarray(vertices, dtype=np.float64)
    f = np.asarray(faces, dtype=np.int64)
    lines = ["ply", "format ascii 1.0", f"element vertex {len(v)}", "property float x", "property float y",
             "property float z", f"element face {len(f)}", "property list uchar int vertex_indices", "end_header"]
    lines += [f"{a:.17g} {b:.17g} {c:.17g}" for a, b, c in v]
    lines += [f"3 {a} {b} {c}" for a, b, c in f]
    path.write_text("\n".join(lines) + "\n")


def write_points(path: Path, points, sep: str = ",") -> None:
    p = np.asarray(points, dtype=np.float64).reshape(-1, 3)
    path.write_text("".join(sep.join(f"{x:.17g}" for x in row) + "\n" for row in p))


def read_csv(path: Path) -> dict[str, np.ndarray]:
    with path.open() as handle:
        rows = list(csv.DictReader(handle))
    return {k: np.array([float(r[k]) for r in rows]) for k in rows[0]}
