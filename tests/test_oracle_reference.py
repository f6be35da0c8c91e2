"""The CPU oracle against the reference's own retained outputs.

Fixtures (tests/golden/, written by make_fixtures.py from the reference's
benchmarks/compiled_bunny/output/ files): the 69,674- and 1,000-face Stanford
Bunny meshes, the 384 evaluation sites, walk-traces.csv (12 traced paths) and the
field CSVs (u, score, score_se, steps per site).  The oracle must reproduce them
bit for bit: mt19937_64 streams, splitmix64 keys, random_unit_vector, the walk,
closest-point distance (binary32-rounded, as the retained outputs are), and the
cross-index square + jackknife combine.
"""

import json

import numpy as np
import pytest

import oracle as O


@pytest.fixture(scope="module")
def bunny(golden):
    meta = json.loads((golden / "bunny_meta.json").read_text())
    sites = np.load(golden / "bunny_sites.npy")
    meshes = {}
    for faces in (1000, 69674):
        m = np.load(golden / f"bunny_{faces}.npz")
        meshes[faces] = O.Mesh(m["vertices"], m["faces"])
    return meta, sites, meshes


def test_walk_traces_bitwise(bunny, golden):
    meta, sites, meshes = bunny
    traces = np.load(golden / "bunny_walk_traces.npy")
    geo = O.Geometry(O.MESH, mesh=meshes[69674], mesh_radius_f32=True)
    side = O.reference_affine_side()
    for path in range(12):
        rows = traces[traces[:, 0] == path]
        site = path * 384 // 12
        key = meta["traces"]["seed"] + meta["traces"]["key_offset"] + site
        got = O.trace(geo, sites[site], side, rng=O.RNG_MT_KEY, key=key, eps=5e-4, max_steps=512)
        assert len(got) == len(rows)
        np.testing.assert_array_equal(got, rows[:, 2:6])


def test_walk_traces_need_binary32_radius(bunny, golden):
    """The retained traces are not reproduced with the binary64 distance: the
    float query radius is part of the recorded behaviour."""
    meta, sites, meshes = bunny
    traces = np.load(golden / "bunny_walk_traces.npy")
    geo = O.Geometry(O.MESH, mesh=meshes[69674], mesh_radius_f32=False)
    rows = traces[traces[:, 0] == 0]
    key = meta["traces"]["seed"] + meta["traces"]["key_offset"]
    got = O.trace(geo, sites[0], O.reference_affine_side(), rng=O.RNG_MT_KEY, key=key)
    assert np.abs(got[:, 3] - rows[:, 5]).max() > 0


@pytest.mark.parametrize("name,count", [("twos_4", 384), ("twos_8", 384), ("twos_16", 384),
                                        ("twos_32", 384), ("twos_64", 384), ("faces_1000", 384),
                                        ("faces_69674", 384), ("twos_128", 96), ("reference", 24)])
def test_field_bitwise(bunny, golden, name, count):
    meta, sites, meshes = bunny
    f = meta["fields"][name]
    ref = np.load(golden / f"field_{name}.npz")
    geo = O.Geometry(O.MESH, mesh=meshes[f["faces"]], mesh_radius_f32=True)
    # Sites keep their global index in the walk key, so a prefix is replayed.
    r = O.estimate(geo, sites, [O.reference_affine_side(f["source_y_slope"])], combine=O.SQUARE,
                   rng=O.RNG_MT_SITE, seed=f["seed"], repeat=f["repeat"], walks=f["walks"],
                   eps=f["epsilon"], max_steps=f["max_steps"]) if count == 384 else None
    if r is None:
        r = _prefix(geo, sites, f, count)
    np.testing.assert_array_equal(r["steps"][:count], ref["steps"][:count])
    np.testing.assert_array_equal(r["caps"][:count], ref["limit_hits"][:count])
    np.testing.assert_array_equal(r["u"][:count], ref["u"][:count])
    np.testing.assert_array_equal(r["score"][:count], ref["score"][:count])
    if f["walks"] >= 4:
        np.testing.assert_array_equal(r["score_se"][:count], ref["score_se"][:count])


def _prefix(geo, sites, f, count):
    """Estimate only the first `count` sites (site index = position)."""
    out = O.estimate(geo, sites[:count], [O.reference_affine_side(f["source_y_slope"])],
                     combine=O.SQUARE, rng=O.RNG_MT_SITE, seed=f["seed"], repeat=f["repeat"],
                     walks=f["walks"], eps=f["epsilon"], max_steps=f["max_steps"])
    return out


def test_start_distances_match_libigl_record(bunny, golden):
    """Step-0 radii of the traces agree with libigl's distances to the
    recorded 3.04e-7 (run_compiled_benchmark.py:810-822); the oracle's exact
    binary64 distance agrees with the traced binary32 radius to float rounding."""
    meta, sites, meshes = bunny
    traces = np.load(golden / "bunny_walk_traces.npy")
    for path in range(12):
        site = path * 384 // 12
        d, *_ = meshes[69674].closest(sites[site])
        r0 = traces[(traces[:, 0] == path) & (traces[:, 1] == 0)][0, 5]
        assert abs(d - r0) <= 1e-7 * max(d, 1e-3)
        assert abs(d - r0) <= meta["embree_libigl_max_error"] + 1e-7
