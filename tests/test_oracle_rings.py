"""Vertex-ring ("continuous") energies in the oracle: ContinuousPixar (8),
ContinuousMixed (9) and ContinuousSkin (10) of GeoLinear3d (Geo.cpp:13-225,
Fem.cpp:941-973).

Pins:
* ring construction == the reference's OneRing arrays (golden fixture, and
  live on block and Armadillo meshes where the compiled reference is built);
* xo_substep_rings_ref64 == the reference's Geo3d::Substep, bit for bit
  (golden fixture over every Rayleigh type + PBD; live with a caller order);
* closed forms the mathematics fixes, in FP64: at rest the ring's weighted
  invariants are I1 = 3 and J = 1; under x -> s x they are 3 s^2 and s^3;
  a rotation leaves both unchanged; the gradients match central differences;
  each gradient field sums to zero (translation) and has zero torque
  (rotation), so a free body keeps zero linear momentum.
"""
import os

import numpy as np
import pytest

import pyoracle as po

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def gold():
    return np.load(os.path.join(GOLDEN, "ref_rings.npz"))


def test_rings_match_golden(gold):
    R = po.Rings(gold["tets"])
    assert R.n_rings == gold["ring_node_off"].size - 1
    assert np.array_equal(R.node_off, gold["ring_node_off"])
    assert np.array_equal(R.nodes, gold["ring_nodes"])
    assert np.array_equal(R.tet_off, gold["ring_tet_off"])
    assert np.array_equal(R.tets, gold["ring_tets"])
    assert np.array_equal(R.idx, gold["ring_idx"])


def test_minstd_ring_order_matches_golden(gold):
    n_rings = gold["ring_node_off"].size - 1
    assert np.array_equal(po.minstd_ring_order(gold["tets"].size // 4, n_rings), gold["ring_order"])


@pytest.mark.parametrize("case", range(10))
def test_ring_substeps_match_golden(gold, case):
    energy, nu, comp, rtype, damp, pbd, pbda = gold["cases"][case]
    ob = po.OracleBody(gold["xyz"], gold["tets"], 1000.0, gold["pinned"], compliance=float(comp), nu=float(nu),
                       variant="ref64")
    ob.tick = int(gold["tick0"][0])
    p = po.make_params(float(gold["dt"][0]), energy=int(energy), compliance=float(comp), poisson=float(nu),
                       damping=float(damp), rayleigh_type=int(rtype), pbd=float(pbd), pbd_amortized=float(pbda))
    ob.substep(p, gold["order"], count=20, ring_order=gold["ring_order"])
    assert np.array_equal(ob.X, gold[f"X{case}"])
    assert np.array_equal(ob.V, gold[f"V{case}"])
    assert np.abs(ob.X - gold["xyz"]).max() > 1e-4  # the case moves


@pytest.mark.skipif(not po.have_ref(), reason="compiled reference not built here")
@pytest.mark.parametrize("mesh", ["block", "mirrored", "armadillo"])
def test_live_rings_match_reference(mesh):
    if mesh == "armadillo":
        xyz, stream = po.ref_armadillo_t4()
        tets = stream.reshape(-1, 5)[:, 1:].reshape(-1).astype(np.uint32)
    else:
        xyz, tets = po.tet_block(5, 4, 3, (0.1, 0.12, 0.1), (0, 0, 0), mirrored=mesh == "mirrored")
    R = po.Rings(tets)
    rb = po.RefBody(xyz, tets, 1000.0)
    assert R.n_rings == rb.ring_order().size
    for v in range(R.n_rings):
        a, b = R.ring(v), rb.ring(v)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), v
        assert np.array_equal(a[2], b[2].astype(np.uint8)), v


@pytest.mark.skipif(not po.have_ref(), reason="compiled reference not built here")
@pytest.mark.parametrize("energy", [8, 9, 10])
@pytest.mark.parametrize("rtype,damp", [(0, 0.0), (1, 0.03), (3, 0.03)])
def test_live_ring_substeps_caller_order(energy, rtype, damp):
    """A caller-set ring order (reversed MINSTD) and a mirrored mesh."""
    xyz, tets = po.tet_block(5, 3, 2, (0.1, 0.1, 0.1), (0, 0, 0), mirrored=True)
    pinned = (xyz[0::3] == 0.0).astype(np.uint8)
    comp = 1e-6 if energy == 8 else 2e-5
    rb = po.RefBody(xyz, tets, 1000.0, pinned)
    vorder = rb.ring_order()[::-1].copy()
    rb.set_ring_order(vorder)
    st = rb.settings(comp, nu=0.45, energy=energy, damping=damp, rayleigh_type=rtype, pbd_damping=1.0,
                     volume_pbd=0.004, amortized_volume_pbd=0.016, tick_id=2)
    rb.substeps(st, np.float32(1.0 / 1200.0), 15)
    Xr, Vr = rb.state()
    ob = po.OracleBody(xyz, tets, 1000.0, pinned, compliance=comp, nu=0.45, variant="ref64")
    ob.tick = 2
    p = po.make_params(float(np.float32(1.0 / 1200.0)), energy=energy, compliance=comp, poisson=0.45,
                       damping=damp, rayleigh_type=rtype, pbd=0.004, pbd_amortized=0.016)
    ob.substep(p, rb.order(), count=15, ring_order=vorder)
    assert np.array_equal(ob.X, Xr) and np.array_equal(ob.V, Vr)


# ------------------------------------------------------------ closed forms

@pytest.fixture(scope="module")
def block():
    xyz, tets = po.tet_block(4, 3, 3, (0.1, 0.1, 0.1), (-0.2, -0.15, -0.15))
    el = po.init_elements(xyz, tets, 1000.0)
    return xyz, tets, po.Rings(tets), el


def ring_ug(X, R, v, el):
    U = np.zeros(2)
    vol = np.zeros(1)
    nn = int(R.node_off[v + 1] - R.node_off[v])
    g = np.zeros(6 * nn)
    po.lib().xo_ring_ug_f64(np.ascontiguousarray(X, np.float64), po.C.byref(R.view), v, el["qq"], el["qr"],
                            el["vol"], U, vol, g)
    return U, vol[0], g[:3 * nn].reshape(nn, 3), g[3 * nn:].reshape(nn, 3)


def rotation(seed):
    q, _ = np.linalg.qr(np.random.default_rng(seed).normal(size=(3, 3)))
    return q * np.sign(np.linalg.det(q))


def test_ring_invariants_at_rest_and_under_scaling(block):
    xyz, tets, R, el = block
    X0 = xyz.astype(np.float64)
    centre = 6 + 5 * 2 + 20 * 2  # an interior vertex (x=1, y=2, z=2 of 5x4x4 nodes)
    for v in (0, centre, R.n_rings - 1):
        U, vol, _, _ = ring_ug(X0, R, v, el)
        tv = R.tets[R.tet_off[v]:R.tet_off[v + 1]]
        assert vol == pytest.approx(0.25 * el["vol"][tv].astype(np.float64).sum(), rel=1e-6)
        assert U[0] == pytest.approx(3.0, abs=1e-5) and U[1] == pytest.approx(1.0, abs=1e-6)
        for s in (0.8, 1.25):
            U, _, _, _ = ring_ug(s * X0, R, v, el)
            assert U[0] == pytest.approx(3.0 * s * s, rel=1e-5)
            assert U[1] == pytest.approx(s ** 3, rel=1e-5)


def test_ring_energies_rotation_invariant(block):
    xyz, tets, R, el = block
    rng = np.random.default_rng(3)
    X = xyz.astype(np.float64) + rng.normal(scale=0.01, size=xyz.size)
    Q = rotation(5)
    Xr = (X.reshape(-1, 3) @ Q.T).reshape(-1)
    for v in range(0, R.n_rings, 7):
        U, _, g0, g1 = ring_ug(X, R, v, el)
        Ur, _, g0r, g1r = ring_ug(Xr, R, v, el)
        assert Ur == pytest.approx(U, rel=1e-10)
        assert np.allclose(g0r, g0 @ Q.T, atol=1e-9 * np.abs(g0).max())
        assert np.allclose(g1r, g1 @ Q.T, atol=1e-9 * np.abs(g1).max())


def test_ring_gradients_finite_difference(block):
    xyz, tets, R, el = block
    rng = np.random.default_rng(11)
    X = xyz.astype(np.float64) + rng.normal(scale=0.015, size=xyz.size)
    h = 1e-6
    for v in (0, 37, R.n_rings // 2):
        _, _, g0, g1 = ring_ug(X, R, v, el)
        nodes = R.nodes[R.node_off[v]:R.node_off[v + 1]]
        fd0, fd1 = np.zeros_like(g0), np.zeros_like(g1)
        for n, node in enumerate(nodes):
            for r in range(3):
                Xp, Xm = X.copy(), X.copy()
                Xp[3 * node + r] += h
                Xm[3 * node + r] -= h
                Up = ring_ug(Xp, R, v, el)[0]
                Um = ring_ug(Xm, R, v, el)[0]
                fd0[n, r] = (Up[0] - Um[0]) / (2 * h)
                fd1[n, r] = (Up[1] - Um[1]) / (2 * h)
        assert np.allclose(g0, fd0, rtol=1e-6, atol=1e-6 * np.abs(g0).max()), v
        assert np.allclose(g1, fd1, rtol=1e-6, atol=1e-6 * np.abs(g1).max()), v


def test_ring_gradients_sum_and_torque_vanish(block):
    xyz, tets, R, el = block
    rng = np.random.default_rng(2)
    X = xyz.astype(np.float64) + rng.normal(scale=0.02, size=xyz.size)
    P = X.reshape(-1, 3)
    for v in range(0, R.n_rings, 5):
        _, _, g0, g1 = ring_ug(X, R, v, el)
        x = P[R.nodes[R.node_off[v]:R.node_off[v + 1]]]
        for g in (g0, g1):
            scale = np.abs(g).max() * max(np.abs(x).max(), 1.0)
            assert np.abs(g.sum(0)).max() <= 1e-12 * scale * len(g)
            assert np.abs(np.cross(x, g).sum(0)).max() <= 1e-11 * scale * len(g)


@pytest.mark.parametrize("energy", [8, 9, 10])
def test_free_body_keeps_zero_momentum(block, energy):
    """No pins, no gravity, a deformed start at rest: after a substep of the
    FP64 oracle, sum m V = 0 (each ring moves its nodes by w lambda g with
    sum g = 0). Angular momentum of V = (X - O) / dt is not exact for a
    position-based sweep (second-order term sum m (X - O) x dX), so rotation
    invariance is pinned on the gradients above instead."""
    xyz, tets, R, el = block
    rng = np.random.default_rng(energy)
    X0 = (xyz.astype(np.float64) + rng.normal(scale=0.01, size=xyz.size)).astype(np.float32)
    ob = po.OracleBody(X0, tets, 1000.0, None, compliance=1e-6 if energy == 8 else 1e-4, nu=0.4, variant="f64")
    p = po.make_params(1e-3, gravity=(0.0, 0.0), energy=energy, compliance=1e-6 if energy == 8 else 1e-4,
                       poisson=0.4)
    ob.substep(p, np.arange(ob.t, dtype=np.uint32), count=1, ring_order=po.minstd_ring_order(ob.t, R.n_rings))
    m = 1.0 / ob.w.astype(np.float64)
    V = ob.V.reshape(-1, 3)
    x = ob.X.reshape(-1, 3)
    p_lin = (m[:, None] * V).sum(0)
    ref = (m[:, None] * np.abs(V)).sum()
    # floor: FP64 rounding of positions |x| <= 0.3 turned into velocities by 1 / dt
    floor = 64 * np.finfo(np.float64).eps * np.abs(x).max() / 1e-3 * m.sum()
    assert ref > 1e4 * floor
    assert np.abs(p_lin).max() <= floor
