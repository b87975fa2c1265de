"""The rest of TSolveElement's energy switch for linear tets (Fem.cpp:870-889)
in the oracle, against the compiled reference: Pixar (0), PixarReduced (1),
PixarSel (2) -- SolveElementPixar with one quadrature point --, CubeNeo (11)
and CubeSkin (12), which act on quads and hexes only and leave tets alone,
and energies 13..31, which take the switch's default (MixedSelective).
Live, where the compiled reference is built; every Rayleigh type and PBD
for the Pixar energies. The golden fixture ref_dispatch.npz keeps the same
checks where the reference is absent."""
import os

import numpy as np
import pytest

import pyoracle as po

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
needs_ref = pytest.mark.skipif(not po.have_ref(), reason="compiled reference not built here")
DT = float(np.float32(1) / np.float32(1200))


def block():
    xyz, tets = po.tet_block(6, 3, 3, (0.1, 0.1, 0.1), (0.0, -0.15, -0.15))
    return xyz, tets, (xyz[0::3] == 0.0).astype(np.uint8), po.minstd_order(tets.size // 4)


def oracle_run(xyz, tets, pinned, order, energy, sim, nu, comp, rtype=0, damp=0.0, pbd=0.0, pbda=0.0, n=40):
    ob = po.OracleBody(xyz, tets, 1000.0, pinned, compliance=comp, nu=nu, variant="ref64")
    ob.tick = 5
    p = po.make_params(DT, energy=energy, simultaneous=sim, compliance=comp, poisson=nu, damping=damp,
                       rayleigh_type=rtype, pbd=pbd, pbd_amortized=pbda)
    ob.substep(p, order, count=n)
    return ob


@needs_ref
@pytest.mark.parametrize("energy", [0, 1, 2, 11, 12, 13, 31])
@pytest.mark.parametrize("sim", [True, False])
@pytest.mark.parametrize("nu", [0.45, 0.3])
def test_live_energy_dispatch(energy, sim, nu):
    xyz, tets, pinned, order = block()
    comp = 1e-6 if energy <= 2 else 2e-5
    rb = po.RefBody(xyz, tets, 1000.0, pinned, order)
    rb.substeps(rb.settings(comp, nu=nu, energy=energy, simultaneous=sim, tick_id=5), DT, 40)
    X, V = rb.state()
    ob = oracle_run(xyz, tets, pinned, order, energy, sim, nu, comp)
    assert np.array_equal(X, ob.X) and np.array_equal(V, ob.V)
    assert np.abs(X - xyz).max() > 1e-4


@needs_ref
@pytest.mark.parametrize("energy", [0, 2])
@pytest.mark.parametrize("rtype,damp,pbd,pbda", [(0, 0.02, 0.0, 0.0), (1, 0.02, 0.0, 0.0), (2, 0.02, 0.003, 0.0),
                                                 (3, 0.02, 0.003, 0.012)])
def test_live_pixar_damping(energy, rtype, damp, pbd, pbda):
    xyz, tets, pinned, order = block()
    rb = po.RefBody(xyz, tets, 1000.0, pinned, order)
    rb.substeps(rb.settings(1e-6, nu=0.45, energy=energy, damping=damp, rayleigh_type=rtype,
                            pbd_damping=1.0 if pbd else 0.0, volume_pbd=pbd, amortized_volume_pbd=pbda, tick_id=5),
                DT, 40)
    X, V = rb.state()
    ob = oracle_run(xyz, tets, pinned, order, energy, True, 0.45, 1e-6, rtype, damp, pbd, pbda)
    assert np.array_equal(X, ob.X) and np.array_equal(V, ob.V)


def test_cube_energies_leave_tets_free():
    """CubeNeo / CubeSkin do not constrain tets: the body falls freely (pins
    aside), exactly as with no elastic term."""
    xyz, tets, pinned, order = block()
    ob = oracle_run(xyz, tets, pinned, order, 11, True, 0.45, 2e-5, n=10)
    free = pinned == 0
    y0 = xyz[1::3].astype(np.float64)
    dy = ob.X[1::3] - y0
    assert np.allclose(dy[free], dy[free][0], rtol=0, atol=1e-12)  # rigid drop of the free nodes
    assert np.all(dy[~free] == 0.0)


def test_pixar_full_equals_reduced_for_linear_tets():
    xyz, tets, pinned, order = block()
    a = oracle_run(xyz, tets, pinned, order, 0, True, 0.45, 1e-6, n=10)
    b = oracle_run(xyz, tets, pinned, order, 1, False, 0.45, 1e-6, n=10)
    assert np.array_equal(a.X, b.X)


def test_dispatch_matches_golden():
    g = np.load(os.path.join(GOLDEN, "ref_dispatch.npz"))
    for k, (energy, sim, nu, comp, rtype, damp, pbd, pbda) in enumerate(g["cases"]):
        ob = oracle_run(g["xyz"], g["tets"], g["pinned"], g["order"], int(energy), bool(sim), float(nu), float(comp),
                        int(rtype), float(damp), float(pbd), float(pbda), n=40)
        assert np.array_equal(ob.X, g[f"X{k}"]) and np.array_equal(ob.V, g[f"V{k}"]), k
