"""Tet mesh ingestion: tetgen .node/.ele files and the reference's element
index stream (GeoLinear3d::Init input, Geo.cpp:697), and the substep on an
irregular mesh (the reference's Armadillo tet mesh, MeshGen.cpp:23-80, live
from the compiled reference where it is built)."""
import numpy as np
import pytest

import pyoracle as po
import xpbdfem as xf

scipy_spatial = pytest.importorskip("scipy.spatial")


def jittered_delaunay(nx=6, ny=4, nz=4, h=0.1, jitter=0.2, seed=0):
    rng = np.random.default_rng(seed)
    g = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"), -1).reshape(-1, 3) * h
    p = g + rng.uniform(-jitter * h, jitter * h, g.shape)
    p[:, 0] -= p[:, 0].min()
    d = scipy_spatial.Delaunay(p)
    return p.astype(np.float32), d.simplices.astype(np.uint32)


def signed_volumes(xyz, tets):
    x = xyz.reshape(-1, 3).astype(np.float64)
    t = tets.reshape(-1, 4)
    P = x[t[:, :3]] - x[t[:, 3]][:, None, :]
    return np.linalg.det(P) / 6.0


@pytest.mark.parametrize("base", [0, 1])
def test_tetgen_round_trip(tmp_path, base):
    xyz, tets = jittered_delaunay()
    node, ele = xf.write_tetgen(tmp_path / "m", xyz, tets, base=base)
    # tetgen extras: a comment line, attributes and boundary markers
    lines = open(node).read().splitlines()
    lines[0] = lines[0].split()[0] + " 3 1 1   # count dim attrs markers"
    lines = [lines[0], "# a comment line"] + [l + " 0.5 1" for l in lines[1:]]
    open(node, "w").write("\n".join(lines) + "\n")
    x2, t2 = xf.read_tetgen(node, ele, orient=False)
    assert np.array_equal(x2, xyz.reshape(-1)) and np.array_equal(t2, tets.reshape(-1))


def test_tetgen_orientation(tmp_path):
    xyz, tets = jittered_delaunay(seed=3)
    assert (signed_volumes(xyz, tets) < 0).any()  # Delaunay output is mixed
    node, ele = xf.write_tetgen(tmp_path / "m", xyz, tets)
    x2, t2 = xf.read_tetgen(node, ele, orient=True)
    assert (signed_volumes(x2, t2) > 0).all()
    assert sorted(map(tuple, np.sort(t2.reshape(-1, 4), 1))) == sorted(map(tuple, np.sort(tets, 1)))
    el = po.init_elements(x2, t2, 1000.0)
    assert (el["vol"] > 0).all() and (el["w"] > 0).all()


def test_tetgen_errors(tmp_path):
    xyz, tets = jittered_delaunay()
    node, ele = xf.write_tetgen(tmp_path / "m", xyz, tets)
    with pytest.raises(xf.XfError, match="cannot open"):
        xf.read_tetgen(tmp_path / "missing.node", ele)
    bad = tmp_path / "bad.ele"
    bad.write_text("1 10 0\n0 0 1 2 3 4 5 6 7 8 9\n")
    with pytest.raises(xf.XfError, match="4 nodes"):
        xf.read_tetgen(node, bad)
    bad.write_text(f"1 4 0\n0 0 1 2 {len(xyz)}\n")
    with pytest.raises(xf.XfError, match="out of range"):
        xf.read_tetgen(node, bad)


def test_stream_conversion():
    xyz, tets = xf.tet_block(3, 2, 2, (0.1, 0.1, 0.1), (0, 0, 0))
    t = tets.reshape(-1, 4)
    stream = np.concatenate([np.full((len(t), 1), 4, np.uint32), t], 1).reshape(-1)
    assert np.array_equal(xf.tets_from_stream(stream), tets)
    with pytest.raises(xf.XfError):
        xf.tets_from_stream(np.array([8, 0, 1, 2, 3, 4, 5, 6, 7], np.uint32))  # a hex


needs_ref = pytest.mark.skipif(not po.have_ref(), reason="reference build (oracle/_ref) not present")


@needs_ref
def test_armadillo_matches_reference(tmp_path):
    """The reference's own irregular tet mesh: stream ingestion, element
    constants and 20 substeps over a greedy colour order (62 colours), oracle
    == compiled reference; the tetgen writer/reader round-trips it."""
    xyz, stream = po.ref_armadillo_t4()
    tets = xf.tets_from_stream(stream)
    n, t = xyz.size // 3, tets.size // 4
    pinned = (xyz[0::3] - xyz[0::3].min() < 0.02 * np.ptp(xyz[0::3])).astype(np.uint8)
    order, _, nc = xf.mesh_colour(n, tets)
    print(f"\narmadillo: {n} nodes, {t} tets, {nc} colours")
    rb = po.RefBody(xyz, tets, 1000.0, pinned, order)
    rt = np.zeros(4 * t, np.uint32)
    po.ref().xr_get_tets(rb.h, rt)
    assert np.array_equal(rt, tets)
    ob = po.OracleBody(xyz, tets, 1000.0, pinned, compliance=1e-6, variant="ref64")
    re = rb.elements()
    for k in ("qi", "qq", "qr", "vol"):
        assert np.array_equal(getattr(ob, k), re[k]), k
    assert np.array_equal(ob.area, rb.area())
    dt = float(np.float32(1.0) / np.float32(1200.0))
    rb.substeps(rb.settings(1e-6, energy=7, simultaneous=True), np.float32(dt), 20)
    ob.substep(po.make_params(dt, energy=7, simultaneous=True), order, count=20)
    X, V = rb.state()
    assert np.array_equal(X, ob.X) and np.array_equal(V, ob.V)
    node, ele = xf.write_tetgen(tmp_path / "armadillo", xyz, tets, base=1)
    x2, t2 = xf.read_tetgen(node, ele, orient=False)
    assert np.array_equal(x2, xyz) and np.array_equal(t2, tets)


@pytest.mark.parametrize("count", ["-1", "nan", "1e30", "2.5", "4294967297"])
def test_tetgen_malformed_counts(tmp_path, count):
    """Header counts that are negative, non-finite, fractional or beyond
    2^31 are rejected with a message (no undefined integer cast)."""
    xyz, tets = jittered_delaunay()
    node, ele = xf.write_tetgen(tmp_path / "m", xyz, tets)
    bad = tmp_path / "bad.node"
    bad.write_text(f"{count} 3 0 0\n0 0 0 0\n")
    with pytest.raises(xf.XfError, match="node count"):
        xf.read_tetgen(bad, ele)
    bad = tmp_path / "bad.ele"
    bad.write_text(f"{count} 4 0\n0 0 1 2 3\n")
    with pytest.raises(xf.XfError, match="tet count"):
        xf.read_tetgen(node, bad)
    bad.write_text("1 4 0\n0 0 1 2 nan\n")
    with pytest.raises(xf.XfError, match="out of range"):
        xf.read_tetgen(node, bad)
