"""The product's host ring construction (xf_mesh_rings, used by scenes with
vertex-ring energies) against the reference's OneRing arrays in the golden
fixture, and against the oracle's builder on a mirrored block and an
irregular Delaunay mesh. Runs on the CPU (no device call)."""
import os

import numpy as np
import pytest

import pyoracle as po
import xpbdfem as xf

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_mesh_rings_match_golden():
    g = np.load(os.path.join(GOLDEN, "ref_rings.npz"))
    r = xf.mesh_rings(g["tets"])
    for k, gk in (("node_off", "ring_node_off"), ("nodes", "ring_nodes"), ("tet_off", "ring_tet_off"),
                  ("tets", "ring_tets"), ("idx", "ring_idx")):
        assert np.array_equal(r[k], g[gk]), k


@pytest.mark.parametrize("mesh", ["mirrored", "delaunay"])
def test_mesh_rings_match_oracle(mesh):
    if mesh == "mirrored":
        _, tets = xf.tet_block(7, 3, 4, (0.1, 0.1, 0.1), (0, 0, 0), mirrored=True)
    else:
        spatial = pytest.importorskip("scipy.spatial")
        pts = np.random.default_rng(1).uniform(size=(400, 3))
        tets = spatial.Delaunay(pts).simplices.astype(np.uint32).reshape(-1)
    r = xf.mesh_rings(tets)
    o = po.Rings(tets)
    assert r["largest"] == o.largest and r["node_off"].size == o.n_rings + 1
    for k in ("node_off", "nodes", "tet_off", "tets", "idx"):
        assert np.array_equal(r[k], getattr(o, k)), k


def test_ring_limits_rejected():
    """A vertex in 129 tets (a fan around node 0 with shared rim nodes: 133
    nodes, under the 256-node limit) exceeds OneRingConstructor's 128 tets."""
    fan = np.array([[0, 1 + i, 2 + i, 132] for i in range(129)], np.uint32).reshape(-1)
    with pytest.raises(xf.XfError):
        xf.mesh_rings(fan)
    with pytest.raises(ValueError):
        po.Rings(fan)
    ok = np.array([[0, 1 + i, 2 + i, 132] for i in range(128)], np.uint32).reshape(-1)
    assert xf.mesh_rings(ok)["tet_off"][1] == 128
