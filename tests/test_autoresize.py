"""GeoLinear3d::Init's autoResize (Geo.cpp:712-729): xpbdfem.autoresize
and the C++ adapter's rescale give the reference's FP64 positions bit for
bit (golden fixture; live on other meshes where the compiled reference is
built). The GPU trajectories from those positions are checked in
test_gpu_adapter.py."""
import os

import numpy as np
import pytest

import pyoracle as po
import xpbdfem as xf

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_autoresize_matches_golden():
    g = np.load(os.path.join(GOLDEN, "ref_autoresize.npz"))
    assert np.array_equal(xf.autoresize(g["xyz"]), g["X0"])


def test_autoresize_closed_form():
    """Centred on the bounding box, largest extent 0.085 (the box is taken in
    float, so both hold to float rounding of the input coordinates)."""
    xyz, _ = xf.tet_block(7, 3, 2, (0.3, 0.1, 0.2), (4.0, -1.0, 2.0))
    p = xf.autoresize(xyz).reshape(-1, 3)
    ext = p.max(0) - p.min(0)
    assert ext.max() == pytest.approx(0.085, rel=1e-6)
    assert np.allclose(0.5 * (p.max(0) + p.min(0)), 0.0, atol=1e-7)
    assert ext[1] / ext[0] == pytest.approx(0.3 / 2.1, rel=1e-5)


@pytest.mark.skipif(not po.have_ref(), reason="compiled reference not built here")
@pytest.mark.parametrize("mesh", ["mirrored", "armadillo"])
def test_live_autoresize_matches_reference(mesh):
    if mesh == "armadillo":
        xyz, stream = po.ref_armadillo_t4()
        tets = stream.reshape(-1, 5)[:, 1:].reshape(-1).astype(np.uint32)
    else:
        xyz, tets = po.tet_block(5, 3, 4, (0.13, 0.1, 0.07), (1.0, 2.0, -3.0), mirrored=True)
    rb = po.RefBody(xyz, tets, 1000.0, autoresize=True)
    X0, _ = rb.state()
    assert np.array_equal(xf.autoresize(xyz), X0)
