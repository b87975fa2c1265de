"""Vertex-ring ("continuous") energies on the GPU: ContinuousPixar (8),
ContinuousMixed (9) and ContinuousSkin (10) through the C ABI against the
oracle's serial ring sweep (xo_substep_rings_ref64, itself bit-exact with
the compiled reference) over the scene's ring order. Bar: bit-exact
positions and velocities."""
import numpy as np
import pytest

import pyoracle as po
import workloads
import xpbdfem as xf

pytestmark = pytest.mark.gpu

DAMPING = {  # rayleigh type, damping, pbd factor, amortized pbd factor
    "none": (0, 0.0, 0.0, 0.0),
    "paper": (0, 0.02, 0.0, 0.0),
    "limit": (1, 0.02, 0.0, 0.0),
    "post": (2, 0.02, 0.003, 0.0),
    "amortized": (3, 0.02, 0.003, 0.012),
}


def ring_beam(energy, damping="none", nu=0.45, W=12, H=4, D=4, mirrored=False, jitter=0.0, seed=0):
    xyz, tets = xf.tet_block(W, H, D, (0.1, 0.1, 0.1), (0.0, -0.2, -0.2), mirrored=mirrored)
    if jitter:
        xyz = (xyz + np.random.default_rng(seed).uniform(-jitter, jitter, xyz.size)).astype(np.float32)
    rtype, damp, pbd, pbda = DAMPING[damping]
    return dict(
        name=f"ring_beam{W}x{H}x{D}_e{energy}", xyz=xyz, tets=tets, density=1000.0,
        pinned=(np.asarray(xyz[0::3]) < 0.05).astype(np.uint8), velocity=None, tet_mu=None,
        settings=dict(energy=energy, compliance=1e-6 if energy == 8 else 2e-5, poissons_ratio=nu,
                      gravity=(0.0, -9.81), collide=xf.COLLIDE_GROUND, ground_y=-0.35, friction=0.5,
                      damping=damp, rayleigh_type=rtype, pbd_damping=pbd, pbd_damping_amortized=pbda),
        dt=np.float32(1.0) / np.float32(1200.0), substeps=20)


def oracle(wl, order, ring_order, substeps, tick=0):
    ob = po.body_from_workload(wl, "ref64")
    ob.tick = tick
    ob.substep(po.params_from_settings(wl["dt"], wl["settings"]), order, count=substeps, ring_order=ring_order)
    return ob


def run(ctx, wl, chunks=2, tick=0, **opts):
    sc = workloads.make_scene(ctx, wl, **opts)
    if tick:
        sc.set_tick(tick)
    for _ in range(chunks):
        sc.step(wl["substeps"], float(wl["dt"]))
    x, v = sc.get_state(0)
    order, _ = sc.order(0)
    rorder, rcol = sc.ring_order(0)
    info = sc.info()
    t = sc.tick()
    sc.close()
    return x, v, order, rorder, rcol, info, t


@pytest.mark.parametrize("energy", [8, 9, 10])
@pytest.mark.parametrize("damping", list(DAMPING))
def test_ring_energies_bit_exact(ctx, energy, damping):
    wl = ring_beam(energy, damping)
    S = wl["substeps"]
    x, v, order, rorder, rcol, info, t = run(ctx, wl, tick=3)
    assert t == 3 + 2 * S
    ob = oracle(wl, order, rorder, 2 * S, tick=3)
    moved = np.abs(x - wl["xyz"]).max()
    print(f"\nenergy {energy} {damping}: {info['n_rings']} rings, {info['n_ring_colours']} ring colours, "
          f"moved {moved:.3e}, max|dx| vs oracle {np.abs(x - ob.X).max():.3e}")
    assert moved > 1e-4
    assert np.array_equal(x, ob.X) and np.array_equal(v, ob.V)


def test_ring_colouring_is_valid(ctx):
    wl = ring_beam(9, jitter=0.0)
    _, _, _, rorder, rcol, info, _ = run(ctx, wl, chunks=1)
    R = po.Rings(wl["tets"])
    assert info["n_rings"] == R.n_rings and sorted(rorder.tolist()) == list(range(R.n_rings))
    assert np.all(np.diff(rcol[rorder].astype(np.int64)) >= 0)  # colour-major
    for c in range(info["n_ring_colours"]):
        members = np.nonzero(rcol == c)[0]
        nodes = np.concatenate([R.ring(v)[0] for v in members])
        assert np.unique(nodes).size == nodes.size, c  # rings of a colour share no node


@pytest.mark.parametrize("energy", [8, 10])
@pytest.mark.parametrize("block,slabs", [(64, False), (256, False), (64, True), (128, True)])
@pytest.mark.parametrize("lanes", [16, 32])
def test_ring_launch_shapes_bit_exact(ctx, energy, block, slabs, lanes):
    """Rings of 24 tets: with 16 lanes per ring a lane takes 1 or 2 tets."""
    wl = ring_beam(energy, "limit")
    x, v, order, rorder, _, _, _ = run(ctx, wl, block_size=block, xcd_slabs=slabs, ring_lanes=lanes)
    ob = oracle(wl, order, rorder, 2 * wl["substeps"])
    assert np.array_equal(x, ob.X) and np.array_equal(v, ob.V)


def test_ring_minstd_visit_and_irregular_mesh(ctx):
    """The reference's vOrder as colouring visit order, on a mirrored block
    with jittered nodes (ring sizes vary)."""
    wl = ring_beam(10, "amortized", W=8, H=3, D=3, mirrored=True, jitter=0.02, seed=4)
    x, v, order, rorder, _, _, _ = run(ctx, wl, colour_order=xf.COLOUR_MINSTD)
    ob = oracle(wl, order, rorder, 2 * wl["substeps"])
    assert np.array_equal(x, ob.X) and np.array_equal(v, ob.V)


def test_ring_graph_equals_eager(ctx):
    wl = ring_beam(9, "post")
    xg, vg, _, _, _, _, _ = run(ctx, wl, use_graph=True)
    xe, ve, _, _, _, _, _ = run(ctx, wl, use_graph=False)
    assert np.array_equal(xg, xe) and np.array_equal(vg, ve)


def test_ring_batched_bodies_bit_exact(ctx):
    """Three bodies with energies 8, 9, 10 and different damping in one scene:
    each equals the oracle's sweep of its own ring order."""
    bodies = [ring_beam(8, "paper", W=6), ring_beam(9, "amortized", W=7, nu=0.5), ring_beam(10, "post", W=5)]
    batch = dict(name="rings3", bodies=bodies, dt=bodies[0]["dt"], substeps=20)
    sc = workloads.make_batch_scene(ctx, batch)
    sc.step(20, float(batch["dt"]))
    sc.step(20, float(batch["dt"]))
    for i, wl in enumerate(bodies):
        x, v = sc.get_state(i)
        ob = oracle(wl, sc.order(i)[0], sc.ring_order(i)[0], 40)
        assert np.array_equal(x, ob.X) and np.array_equal(v, ob.V), wl["name"]
    sc.close()


@pytest.mark.parametrize("damping", ["paper", "post"])
def test_ring_energies_sixteen_lanes(ctx, damping):
    """Damped forms with 16 lanes per ring (the automatic choice here is 32)."""
    for energy in (8, 9, 10):
        wl = ring_beam(energy, damping)
        x, v, order, rorder, _, _, _ = run(ctx, wl, ring_lanes=16)
        ob = oracle(wl, order, rorder, 2 * wl["substeps"])
        assert np.array_equal(x, ob.X) and np.array_equal(v, ob.V), energy


def test_ring_mode_rejections(ctx):
    wl = ring_beam(9)
    dt = float(wl["dt"])
    sc = xf.Scene(ctx)
    sc.add_body(wl["xyz"], wl["tets"], wl["density"], pinned=wl["pinned"])
    with pytest.raises(xf.XfError):
        sc.finalize(ring_lanes=8)
    sc.close()
    sc = workloads.make_scene(ctx, wl, launch_mode=xf.LAUNCH_PERSISTENT)
    with pytest.raises(xf.XfError):
        sc.step(1, dt)
    sc.close()
    sc = workloads.make_scene(ctx, dict(wl, tet_mu=np.full(wl["tets"].size // 4, 1e5, np.float32)))
    with pytest.raises(xf.XfError):
        sc.step(1, dt)
    sc.close()
    mixed = dict(name="mixed", bodies=[wl, dict(ring_beam(9), settings=dict(wl["settings"], energy=4))],
                 dt=wl["dt"], substeps=20)
    sc = workloads.make_batch_scene(ctx, mixed)
    with pytest.raises(xf.XfError):
        sc.step(1, dt)
    sc.close()
    sc = workloads.make_scene(ctx, wl)
    S = wl["substeps"]
    with pytest.raises(xf.XfError):
        sc.adjoint(S, dt, np.array([S], np.uint32), np.zeros((1, wl["xyz"].size)))
    sc.close()


def test_ring_autotune_keeps_results(ctx):
    """The autotuner picks a per-colour shape in ring mode, restores the state,
    and the tuned scene still equals the oracle."""
    wl = ring_beam(10, "amortized")
    sc = workloads.make_scene(ctx, wl)
    mode, size, ms = sc.autotune(wl["substeps"], float(wl["dt"]), reps=1)
    assert mode == xf.LAUNCH_PER_COLOUR and size in (64, 128, 256) and ms > 0.0
    assert sc.tick() == 0
    for _ in range(2):
        sc.step(wl["substeps"], float(wl["dt"]))
    x, v = sc.get_state(0)
    ob = oracle(wl, sc.order(0)[0], sc.ring_order(0)[0], 2 * wl["substeps"])
    sc.close()
    assert np.array_equal(x, ob.X) and np.array_equal(v, ob.V)
