"""The oracle against the reference's own outputs.

Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py
from the compiled, unmodified reference) pin every oracle function that the
reference has: mesh generation, the MINSTD sweep order, element constants,
SolveElement on random deformed tets, and multi-frame Geo3d::Substep
trajectories. The live tests repeat this against oracle/_ref/libxfref.so
where it was built. Bar: bit-exact.
"""
import os

import numpy as np
import pytest

import pyoracle as po

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def beam():
    return dict(np.load(os.path.join(GOLDEN, "ref_beam12.npz")))


@pytest.fixture(scope="module")
def kats():
    return dict(np.load(os.path.join(GOLDEN, "ref_tet_kats.npz")))


def test_mesh_generation_matches_reference(beam):
    W, H, D, h = 12, 4, 4, np.float32(0.1)
    origin = (h * np.float32(-0.5 * W), h * np.float32(-0.5 * H), h * np.float32(-0.5 * D))
    xyz, tets = po.tet_block(W, H, D, (h, h, h), origin)
    assert np.array_equal(xyz, beam["gen_xyz"])
    assert np.array_equal(tets, beam["gen_tets"])


def test_minstd_order_matches_reference(beam):
    assert np.array_equal(po.minstd_order(beam["tets"].size // 4), beam["minstd_order"])


def test_element_constants_match_reference(beam):
    el = po.init_elements(beam["xyz"], beam["tets"], 1000.0)
    for k in ("qi", "qq", "qr", "vol", "w"):
        assert np.array_equal(el[k], beam["el_" + k]), k


def test_natural_colour_order_is_stable(beam):
    n = beam["xyz"].size // 3
    t = beam["tets"].size // 4
    nc, col = po.colour_greedy(n, beam["tets"], np.arange(t, dtype=np.uint32))
    order = np.concatenate([np.nonzero(col == c)[0] for c in range(nc)]).astype(np.uint32)
    assert np.array_equal(order, beam["sweep_order"])


@pytest.mark.parametrize("energy,sim", [(4, True), (7, True), (4, False)])
def test_substep_trajectory_matches_reference(beam, energy, sim):
    tag = f"e{energy}{'s' if sim else 'g'}"
    ob = po.OracleBody(beam["xyz"], beam["tets"], 1000.0, beam["pinned"], compliance=1e-5, nu=0.5, variant="ref64")
    p = po.make_params(float(beam["dt"][0]), energy=energy, simultaneous=sim)
    done = 0
    for f in (1, 10, 120):
        ob.substep(p, beam["sweep_order"], count=20 * f - done)
        done = 20 * f
        assert np.array_equal(ob.X, beam[f"{tag}_X{f}"]), f"frame {f}"
        assert np.array_equal(ob.V, beam[f"{tag}_V{f}"]), f"frame {f}"


def test_solve_element_kats(kats):
    comp = float(kats["compliance"][0])
    dt = float(kats["dt"][0])
    for k in range(kats["X"].shape[0]):
        c0, c1, a = po.material(kats["vol"][k], compliance=comp, nu=float(kats["nu"][k]))
        p = po.make_params(dt, energy=int(kats["energy"][k]), simultaneous=bool(kats["sim"][k]))
        X = kats["X"][k].copy()
        po.lib().xo_solve_tet_ref64(p, X, kats["w"][k], kats["qi"][k], kats["qq"][k], kats["qr"][k], float(c0[0]),
                                    float(c1[0]), float(a[0]))
        assert np.array_equal(X, kats["Xout"][k]), f"kat {k}"


# ------------------------------------------------------------ live reference

def test_volume_and_damping_passes_match_golden():
    """ref_passes.npz (compiled reference): volume passes, every Rayleigh
    damping type and PBD damping over 20 substeps from tickId 5."""
    g = np.load(os.path.join(GOLDEN, "ref_passes.npz"))
    dt = float(g["dt"][0])
    comp = float(g["compliance"][0])
    for k, (energy, sim, vp, rtype, damp, pbd, pbda) in enumerate(g["cases"]):
        ob = po.OracleBody(g["xyz"], g["tets"], 1000.0, g["pinned"], compliance=comp, variant="ref64")
        ob.tick = int(g["tick0"][0])
        p = po.make_params(dt, energy=int(energy), simultaneous=bool(sim), volume_passes=int(vp), compliance=comp,
                           damping=float(np.float32(damp)), rayleigh_type=int(rtype), pbd=float(np.float32(pbd)),
                           pbd_amortized=float(np.float32(pbda)))
        ob.substep(p, g["order"], count=20)
        assert np.array_equal(ob.X, g[f"X{k}"]) and np.array_equal(ob.V, g[f"V{k}"]), k


needs_ref = pytest.mark.skipif(not po.have_ref(), reason="reference build (oracle/_ref) not present")


@needs_ref
@pytest.mark.parametrize("dims,mirrored", [((5, 3, 2), False), ((4, 4, 3), True)])
def test_live_mesh_and_elements(dims, mirrored):
    W, H, D = dims
    n = (W + 1) * (H + 1) * (D + 1)
    rx = np.zeros(3 * n, np.float32)
    rt = np.zeros(24 * W * H * D, np.uint32)
    po.ref().xr_gen_tet_block(W, H, D, 0.07, 0.05, 0.11, int(mirrored), rx.ctypes.data, rt.ctypes.data)
    s = np.float32([0.07, 0.05, 0.11])
    origin = (s[0] * np.float32(-0.5 * W), s[1] * np.float32(-0.5 * H), s[2] * np.float32(-0.5 * D))
    xyz, tets = po.tet_block(W, H, D, s, origin, mirrored)
    assert np.array_equal(xyz, rx) and np.array_equal(tets, rt)
    rb = po.RefBody(xyz, tets, 730.0)
    ref_el = rb.elements()
    el = po.init_elements(xyz, tets, 730.0)
    for k in ("qi", "qq", "qr", "vol", "w"):
        assert np.array_equal(el[k], ref_el[k]), k


@needs_ref
@pytest.mark.parametrize("energy", [3, 4, 5, 6, 7])
@pytest.mark.parametrize("sim", [True, False])
@pytest.mark.parametrize("nu", [0.5, 0.45])
def test_live_substeps_all_energies(energy, sim, nu):
    xyz, tets = po.tet_block(6, 3, 3, (0.1, 0.1, 0.1), (0.0, -0.15, -0.15))
    pinned = (xyz[0::3] == 0.0).astype(np.uint8)
    order = po.minstd_order(tets.size // 4)  # any serial order: here the reference's own
    rb = po.RefBody(xyz, tets, 1000.0, pinned, order)
    st = rb.settings(2e-5, nu=nu, energy=energy, simultaneous=sim)
    ob = po.OracleBody(xyz, tets, 1000.0, pinned, compliance=2e-5, nu=nu, variant="ref64")
    p = po.make_params(float(np.float32(1) / np.float32(1200)), energy=energy, simultaneous=sim)
    rb.substeps(st, p.dt, 40)
    ob.substep(p, order, count=40)
    X, V = rb.state()
    assert np.array_equal(X, ob.X) and np.array_equal(V, ob.V)


@needs_ref
@pytest.mark.parametrize("energy", [3, 4, 7])
@pytest.mark.parametrize("sim", [True, False])
@pytest.mark.parametrize("passes", [1, 3])
def test_live_volume_passes(energy, sim, passes):
    """Settings::volumePasses: SolveElementVolume sweeps after the main sweep
    (Geo.cpp:779-782), oracle == compiled reference."""
    xyz, tets = po.tet_block(6, 3, 3, (0.1, 0.1, 0.1), (0.0, -0.15, -0.15))
    pinned = (xyz[0::3] == 0.0).astype(np.uint8)
    order = po.minstd_order(tets.size // 4)
    rb = po.RefBody(xyz, tets, 1000.0, pinned, order)
    st = rb.settings(2e-5, energy=energy, simultaneous=sim, volume_passes=passes)
    ob = po.OracleBody(xyz, tets, 1000.0, pinned, compliance=2e-5, variant="ref64")
    p = po.make_params(float(np.float32(1) / np.float32(1200)), energy=energy, simultaneous=sim,
                       volume_passes=passes, compliance=2e-5)
    rb.substeps(st, p.dt, 40)
    ob.substep(p, order, count=40)
    X, V = rb.state()
    assert np.array_equal(X, ob.X) and np.array_equal(V, ob.V)
    plain = po.OracleBody(xyz, tets, 1000.0, pinned, compliance=2e-5, variant="ref64")
    plain.substep(po.make_params(p.dt, energy=energy, simultaneous=sim), order, count=40)
    assert not np.array_equal(plain.X, ob.X)  # the passes do something


@needs_ref
def test_live_ldlt_usage_example():
    A = np.array([1, 2, 3, 4, 5, 6], np.float32)
    b = np.array([7, 8, 9], np.float32)
    xr = np.zeros(3, np.float32)
    xo = np.zeros(3, np.float32)
    po.ref().xr_ldlt(3, A, b, xr)
    po.lib().xo_ldlt_solve(3, A, b, xo)
    assert np.array_equal(xr, xo)
