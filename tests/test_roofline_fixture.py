"""The frozen roofline unit of the triangle-mesh workload
(tests/golden/roofline_c3.json) is what the oracle's BVH counts today: the
fixture is recomputed from scratch and must match exactly, and its figures
are sane against the tree's shape (a query visits at least one root-to-leaf
path and tests at least one leaf's triangles)."""

import json
import math

import pytest

from golden import make_roofline_c3 as M  # noqa: E402


@pytest.fixture(scope="module")
def stored(golden):
    return json.loads((golden / "roofline_c3.json").read_text())


def test_fixture_recomputes_exactly(stored):
    assert M.measure() == stored


def test_fixture_is_sane(stored):
    depth = math.ceil(math.log2(stored["triangles"] / 4))  # median split, <= 4 triangles per leaf
    for key in ("first_step", "walk"):
        e = stored[key]
        assert e["nodes_per_query"] >= depth + 1 and e["tris_per_query"] >= 1
        assert e["bytes_per_query"] == pytest.approx(32 * e["nodes_per_query"] + 36 * e["tris_per_query"])
    # A path's later queries sit near the surface and cost less than its
    # first one from the site.
    assert stored["walk"]["bytes_per_query"] < stored["first_step"]["bytes_per_query"]
    assert stored["frozen_bytes_per_query"] == stored["walk"]["bytes_per_query"]
