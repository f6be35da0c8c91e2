"""Oracle against the engineering-housing design loop's retained fields
(wih_twos --mode twos --source-mode housing: evaluate_site,
benchmarks/compiled_bunny/src/wih_twos.cpp:674-717, with the housing chip
source of :553-566 sampled at a Beta(2,2) Green point, :648-653).

These outputs were written with the binary32 query radius (as the
compiled-bunny fields were), so the replay uses the binary32-rounded mesh
distance: step and cap counts match exactly, and u, score and score_se
bitwise on nearly every site.
"""

import numpy as np
import pytest

import oracle as O
import study_fixtures as SF


@pytest.fixture(scope="module")
def housing(golden):
    v, f, requests = SF.load_housing(golden)
    return O.Mesh(v, f), requests


def test_housing_fields(housing):
    mesh, requests = housing
    off, chips = O.housing_chip_terms()
    exact = []
    for req in requests:
        g = O.Geometry(container=O.MESH, mesh=mesh, cavities=req.cavities if len(req.cavities) else None,
                       cavity_radius=req.cavity_radius, mesh_radius_f32=True)
        r = O.estimate(g, req.points, [O.side(terms=chips, offset=off)], combine=O.SQUARE, rng=O.RNG_MT_SITE,
                       seed=req.seed, walks=req.walks, eps=req.epsilon, max_steps=req.max_steps, threads=8)
        exact.append(SF.compare_housing(req, r["u"], r["score"], r["score_se"], r["steps"], r["caps"], rtol=1e-6))
    exact = np.concatenate(exact)
    assert exact.mean() > 0.97, exact.mean()


def test_housing_chip_source_matters(housing):
    """A different chip weight changes u by far more than the tolerance."""
    mesh, requests = housing
    req = requests[0]
    off, chips = O.housing_chip_terms()
    chips = [list(t) for t in chips]
    chips[0][7] = 5.3
    g = O.Geometry(container=O.MESH, mesh=mesh, cavity_radius=req.cavity_radius, mesh_radius_f32=True)
    r = O.estimate(g, req.points[:64], [O.side(terms=chips, offset=off)], combine=O.SQUARE, rng=O.RNG_MT_SITE,
                   seed=req.seed, walks=req.walks, eps=req.epsilon, max_steps=req.max_steps, threads=8)
    np.testing.assert_array_equal(r["steps"], req.counts[:64, 0])
    assert np.max(np.abs(r["u"] - req.values[:64, 0]) / np.abs(req.values[:64, 0])) > 1e-4
