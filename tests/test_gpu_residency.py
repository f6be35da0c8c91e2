"""Resident meshes keyed by content (twos_mesh_create): the same vertex and
face arrays, bit for bit, return the resident mesh instead of a second
build; each create is matched by one release, and the last release frees
the mesh.  A mesh differing in one vertex bit is a different mesh."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

twos = pytest.importorskip("twos")


def test_identical_meshes_share_one_resident_copy(golden):
    m = np.load(golden / "bunny_1000.npz")
    v, f = m["vertices"], m["faces"]
    with twos.Context() as ctx:
        a = ctx.mesh(v, f)
        b = ctx.mesh(v.copy(), f.copy())
        assert b.id == a.id and b.build_ms == 0.0
        w = v.copy()
        w[0, 0] = np.nextafter(w[0, 0], np.inf)
        c = ctx.mesh(w, f)
        assert c.id != a.id
        ctx.release_mesh(c)
        ctx.release_mesh(b)
        # The first handle still holds the mesh.
        g = ctx.geometry(twos.MESH, mesh=a)
        sites = np.load(golden / "bunny_sites.npy")
        site = sites[ctx.inside(a, sites)][:1]
        assert len(site) == 1
        r = ctx.estimate(g, site, [twos.reference_affine_side()], combine=twos.SQUARE, walks=64, eps=1e-3)
        assert r.steps[0] > 0
        ctx.release_geometry(g)
        ctx.release_mesh(a)
        with pytest.raises(twos.NotFoundError):
            ctx.release_mesh(a)
