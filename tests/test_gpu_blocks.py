"""Sweep-block launches (XF_LAUNCH_BLOCKS) on the GPU: each workgroup sweeps
one block through its phases in LDS, blocks of one launch colour in
parallel. Bar: bit-exact with the oracle's serial sweep in the order the
scene reports (colour-major over the (launch colour, phase) colours), and
with per-colour launches over the same colouring."""
import numpy as np
import pytest

import pyoracle as po
import workloads
import xpbdfem as xf

pytestmark = pytest.mark.gpu


def oracle(wl, order, substeps, tick=0):
    ob = po.body_from_workload(wl, "ref64")
    ob.tick = tick
    ob.substep(po.params_from_settings(wl["dt"], wl["settings"]), order, count=substeps)
    return ob


def run(ctx, wl, frames=2, tick=0, **opts):
    opts.setdefault("sweep_blocks", True)
    opts.setdefault("launch_mode", xf.LAUNCH_BLOCKS)
    sc = workloads.make_scene(ctx, wl, **opts)
    sc.set_tick(tick)
    S = wl["substeps"]
    for _ in range(frames):
        sc.step(S, float(wl["dt"]))
    x, v = sc.get_state(0)
    order, colour = sc.order(0)
    info = sc.info()
    t, k = sc.time(), sc.tick()
    sc.close()
    return x, v, order, colour, info, t, k


def check_colouring(wl, order, colour):
    t = wl["tets"].reshape(-1, 4)
    co = colour[order]
    assert np.all(np.diff(co.astype(np.int64)) >= 0)
    pairs = np.stack([np.repeat(colour.astype(np.int64), 4), t.ravel().astype(np.int64)], 1)
    assert len(np.unique(pairs, axis=0)) == len(pairs)


@pytest.mark.parametrize("energy,sim", [(4, True), (7, True), (4, False), (3, True), (5, True), (6, False),
                                        (7, False)])
def test_blocks_bit_exact(ctx, energy, sim):
    wl = workloads.beam_neo_12x4x4(energy=energy, simultaneous=sim)
    x, v, order, colour, info, _, _ = run(ctx, wl)
    assert info["n_blocks"] == 96 and info["block_launches"] == 4 and info["n_colours"] == 24
    check_colouring(wl, order, colour)
    ob = oracle(wl, order, 2 * wl["substeps"])
    print(f"\nenergy {energy} sim {sim}: max|dx| vs oracle {np.abs(x - ob.X).max():.3e}")
    assert np.array_equal(x, ob.X) and np.array_equal(v, ob.V)


def test_blocks_workgroup_sizes(ctx):
    """64, 128 and 256 threads per block (lanes loop over a phase's tets when
    it has more tets than lanes): the same results."""
    wl = workloads.beam_neo_12x4x4(energy=7)
    runs = [run(ctx, wl, block_size=b) for b in (64, 128, 256)]
    ob = oracle(wl, runs[0][2], 2 * wl["substeps"])
    for r in runs:
        assert np.array_equal(r[2], runs[0][2])
        assert np.array_equal(r[0], ob.X) and np.array_equal(r[1], ob.V)


@pytest.mark.parametrize("lanes,block", [(128, 64), (256, 256), (256, 128)])
def test_blocks_slot_lanes(ctx, lanes, block):
    """Blocks merged into slots of up to 128 / 256 tets per phase (the lanes
    of smaller workgroups take several tets in turn) on a 24x8x8 Yeoh beam:
    the same order, fewer workgroups, the oracle's result."""
    wl = workloads.beam_yeoh(24, 8, 8, structured=False)
    base = run(ctx, wl, frames=1)
    r = run(ctx, wl, frames=1, block_lanes=lanes, block_size=block)
    assert base[4]["block_lanes"] == 64
    assert r[4]["block_lanes"] == lanes and r[4]["n_block_slots"] < base[4]["n_block_slots"]
    assert np.array_equal(r[2], base[2])
    ob = oracle(wl, r[2], wl["substeps"])
    assert np.array_equal(r[0], ob.X) and np.array_equal(r[1], ob.V)
    assert np.array_equal(base[0], ob.X) and np.array_equal(base[1], ob.V)


def test_blocks_more_tets_than_lanes(ctx):
    """A mirrored 4x4x150 block: no direction pair beats (x, y), whose
    columns of 900 tets have 12+ phases of up to 75 tets; on 64-thread
    workgroups lanes run two tets in those phases."""
    for cells in [(4, 4, 150)]:
        xyz, tets = xf.tet_block(*cells, (0.05, 0.05, 0.05), (0.0, 0.0, 0.0), mirrored=True)
        wl = workloads.beam_neo_12x4x4(energy=4)
        wl = dict(wl, xyz=xyz, tets=tets, pinned=(xyz.reshape(-1, 3)[:, 0] == 0).astype(np.uint8),
                  velocity=None, tet_colour=None)
        sb = xf.mesh_sweep_blocks(xyz, tets)
        big = np.bincount(sb["block"].astype(np.int64) * 64 + sb["phase"]).max()
        x, v, order, _, _, _, _ = run(ctx, wl, frames=1, block_size=64)
        ob = oracle(wl, order, wl["substeps"])
        print(f"\ncells {cells}: largest phase {big} tets")
        assert np.array_equal(x, ob.X) and np.array_equal(v, ob.V)
        assert big > 64


@pytest.mark.parametrize("use_graph", [True, False])
def test_blocks_equal_per_colour_launches(ctx, use_graph):
    """The same sweep-block colouring run with one launch per colour."""
    wl = workloads.beam_neo_12x4x4(energy=7)
    a = run(ctx, wl, use_graph=use_graph)
    b = run(ctx, wl, launch_mode=xf.LAUNCH_PER_COLOUR, use_graph=use_graph)
    assert np.array_equal(a[2], b[2])
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert a[5] == b[5] and a[6] == b[6]  # clock advanced once per substep


@pytest.mark.parametrize("energy,sim", [(7, True), (4, False)])
def test_block_colouring_persistent_launch(ctx, energy, sim):
    """The persistent launch over the sweep-block colouring (24 colours of
    (launch colour, phase)) is a serial sweep in the reported order too."""
    wl = workloads.beam_neo_12x4x4(energy=energy, simultaneous=sim)
    x, v, order, colour, info, _, _ = run(ctx, wl, launch_mode=xf.LAUNCH_PERSISTENT)
    assert info["n_colours"] == 24
    ob = oracle(wl, order, 2 * wl["substeps"])
    assert np.array_equal(x, ob.X) and np.array_equal(v, ob.V)


def test_blocks_big_beam_bit_exact(ctx):
    """The 96x24x24 Yeoh beam (331,776 tets; 4,608 blocks along x and y - z,
    6 phases each, 4 launches per substep: 24 phases) with ground and moving
    sphere."""
    wl = workloads.beam_yeoh_96x24x24()
    x, v, order, colour, info, _, _ = run(ctx, wl, frames=1)
    assert info["n_blocks"] == 96 * 48 and info["n_colours"] == 24
    assert info["block_lanes"] == 64  # larger slots would leave fewer than 4 per compute unit
    ob = oracle(wl, order, wl["substeps"])
    print(f"\nbig beam: max|dx| vs oracle {np.abs(x - ob.X).max():.3e}")
    assert np.array_equal(x, ob.X) and np.array_equal(v, ob.V)


def test_blocks_large_beam_equals_per_colour(ctx):
    """The 288x72x72 Yeoh beam (8,957,952 tets), the throughput-bound shape
    the bench's large leg runs: 256-lane slots on 256-thread workgroups, more
    than two per compute unit (32-byte LDS nodes, the node phase in its own
    launch). One frame of block launches equals one frame of per-colour
    launches over the same colouring, bit for bit, and stays finite (the
    oracle would take minutes at this size; the 96x24x24 test pins the
    colouring's serial sweep against it)."""
    wl = workloads.beam_yeoh(288, 72, 72)
    a = run(ctx, wl, frames=1, block_size=256, block_lanes=256)
    b = run(ctx, wl, frames=1, block_size=256, block_lanes=256, launch_mode=xf.LAUNCH_PER_COLOUR)
    info = a[4]
    assert info["n_tets"] == 288 * 72 * 72 * 6 and info["n_colours"] == 24
    assert info["block_node_stride"] == 32 and info["fused_finish"] == 0
    assert info["n_block_slots"] > 4 * 2 * 256
    assert np.array_equal(a[2], b[2])
    assert np.all(np.isfinite(a[0])) and np.all(np.isfinite(a[1]))
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_blocks_batch256_equals_per_colour(ctx):
    """The 256-box batch the bench's batched leg runs (3,790,950 tets, MixedSel
    and YeohSkinFast, per-tet material slots): one frame of block launches
    (128-lane slots, 256-thread workgroups) equals per-colour launches over
    the same colouring for every body, bit for bit."""
    b = workloads.boxes_batch()
    dt, S = float(b["dt"]), b["substeps"]
    states = []
    for mode in (xf.LAUNCH_BLOCKS, xf.LAUNCH_PER_COLOUR):
        sc = workloads.make_batch_scene(ctx, b, sweep_blocks=True, launch_mode=mode, block_size=256,
                                        block_lanes=128)
        sc.step(S, dt)
        states.append([sc.get_state(i) for i in range(len(b["bodies"]))] + [sc.order(0)[0]])
        sc.close()
    assert np.array_equal(states[0][-1], states[1][-1])
    for i in range(len(b["bodies"])):
        (xa, va), (xb, vb) = states[0][i], states[1][i]
        assert np.all(np.isfinite(xa))
        assert np.array_equal(xa, xb) and np.array_equal(va, vb), i


@pytest.mark.parametrize("energy", [7, 4])
def test_blocks_fused_node_phase(ctx, energy):
    """The end of each substep (pins, ground and moving-sphere contact with
    friction, velocity, next prediction) done by the last block launch for
    every node, its own block nodes from LDS and the rest from memory: the
    same state as the separate node launch and as the oracle, over two step
    calls (whose last substeps store the velocity instead)."""
    wl = workloads.beam_neo_12x4x4(energy=energy)
    wl["settings"] = dict(wl["settings"], ground_y=-0.2004, collide=xf.COLLIDE_GROUND | xf.COLLIDE_SPHERE,
                          sphere_c0=(0.3, -0.45, 0.0), sphere_v=(0.0, 6.0, 0.0), sphere_r=0.3)
    fused = run(ctx, wl, frames=3)
    sep = run(ctx, wl, frames=3, separate_finish=True)
    assert fused[4]["fused_finish"] == 1 and sep[4]["fused_finish"] == 0
    assert np.array_equal(fused[2], sep[2])
    ob = oracle(wl, fused[2], 3 * wl["substeps"])
    assert np.array_equal(fused[0], ob.X) and np.array_equal(fused[1], ob.V)
    assert np.array_equal(sep[0], ob.X) and np.array_equal(sep[1], ob.V)
    assert fused[5] == sep[5] and fused[6] == sep[6] == 3 * wl["substeps"]
    # contact and the sphere were active
    assert np.any(np.abs(ob.X.reshape(-1, 3)[:, 1] - np.float32(-0.2004)) < 1e-9)


def test_blocks_fused_node_phase_conditions(ctx):
    """Volume passes or damping keep the node phase separate; so does a mesh
    whose blocks fall in fewer than two launch colours (the clock would be
    advanced and read in one launch)."""
    wl = workloads.beam_neo_12x4x4(energy=7)
    for extra in (dict(volume_passes=1), dict(damping=0.01, rayleigh_type=0), dict(pbd_damping=0.002)):
        w = dict(wl, settings=dict(wl["settings"], **extra))
        sc = workloads.make_scene(ctx, w, sweep_blocks=True, launch_mode=xf.LAUNCH_BLOCKS)
        assert sc.info()["fused_finish"] == 0, extra
        sc.close()
    xyz, tets = xf.tet_block(1, 1, 1, (0.1, 0.1, 0.1), (0.0, -0.05, -0.05))
    corner = tets[:4].astype(np.int64)
    one_tet = dict(xyz=xyz.reshape(-1, 3)[corner].ravel().copy(), tets=np.arange(4, dtype=np.uint32))
    for name, mesh, launches, fused in (("cell", dict(xyz=xyz, tets=tets), 2, 1), ("tet", one_tet, 1, 0)):
        w = dict(wl, name=name, pinned=np.zeros(mesh["xyz"].size // 3, np.uint8), **mesh)
        x, v, order, _, info, _, _ = run(ctx, w, frames=2)
        assert info["block_launches"] == launches and info["fused_finish"] == fused, name
        ob = oracle(w, order, 2 * wl["substeps"])
        assert np.array_equal(x, ob.X) and np.array_equal(v, ob.V), name


def test_blocks_minstd_visit_and_volume_passes(ctx):
    wl = workloads.beam_neo_12x4x4(energy=4)
    wl["settings"] = dict(wl["settings"], volume_passes=2)
    x, v, order, colour, _, _, _ = run(ctx, wl, colour_order=xf.COLOUR_MINSTD)
    check_colouring(wl, order, colour)
    ob = oracle(wl, order, 2 * wl["substeps"])
    assert np.array_equal(x, ob.X) and np.array_equal(v, ob.V)


@pytest.mark.parametrize("rtype", [0, 1, 2, 3])
def test_blocks_damping_bit_exact(ctx, rtype):
    """Rayleigh Paper / Limit damp inside the projection (O read from global
    memory); Post / PostAmortized and PBD run per colour after the velocity
    update."""
    wl = workloads.beam_neo_12x4x4(energy=7)
    wl["settings"] = dict(wl["settings"], damping=0.02, rayleigh_type=rtype, pbd_damping=0.003,
                          pbd_damping_amortized=0.012)
    x, v, order, _, _, _, k = run(ctx, wl, tick=5)
    assert k == 5 + 2 * wl["substeps"]
    ob = oracle(wl, order, 2 * wl["substeps"], tick=5)
    assert np.array_equal(x, ob.X) and np.array_equal(v, ob.V)


@pytest.mark.parametrize("energy", [0, 2, 12, 31])
def test_blocks_energy_dispatch_bit_exact(ctx, energy):
    """Pixar (0), PixarSel (2), CubeSkin (12: no constraint on tets, the clock
    still advances) and the default MixedSel (31)."""
    wl = workloads.beam_neo_12x4x4(energy=energy)
    wl["settings"] = dict(wl["settings"], compliance=1e-6 if energy <= 2 else 2e-5, poissons_ratio=0.45)
    x, v, order, _, _, t, k = run(ctx, wl, tick=3)
    assert k == 3 + 2 * wl["substeps"]
    assert np.isfinite(t) and t > 0.0
    ob = oracle(wl, order, 2 * wl["substeps"], tick=3)
    assert np.array_equal(x, ob.X) and np.array_equal(v, ob.V)


def test_blocks_per_tet_mu(ctx):
    """Per-tet mu (element data with comp1 / a / body per tet)."""
    wl = workloads.beam_fit_16x4x4()
    x, v, order, _, _, _, _ = run(ctx, wl, frames=1)
    ob = oracle(wl, order, wl["substeps"])
    assert np.array_equal(x, ob.X) and np.array_equal(v, ob.V)


def test_blocks_batch_bit_exact(ctx):
    """Many bodies, one launch per launch colour: each batched body equals the
    same body alone and the oracle's sweep in its reported order."""
    b = workloads.boxes_batch(n_bodies=6, seed=5, cells=(3, 6), frames=2)
    dt, S = float(b["dt"]), b["substeps"]
    sc = workloads.make_batch_scene(ctx, b, sweep_blocks=True, launch_mode=xf.LAUNCH_BLOCKS)
    # the multi-body form of the fused node phase (body parameters per node)
    assert sc.info()["fused_finish"] == 1
    sc.step(S, dt)
    sc.step(S, dt)
    sep = workloads.make_batch_scene(ctx, b, sweep_blocks=True, launch_mode=xf.LAUNCH_BLOCKS, separate_finish=True)
    assert sep.info()["fused_finish"] == 0
    sep.step(S, dt)
    sep.step(S, dt)
    for i in range(len(b["bodies"])):
        xa, va = sc.get_state(i)
        xb, vb = sep.get_state(i)
        assert np.array_equal(xa, xb) and np.array_equal(va, vb)
    sep.close()
    for i, wl in enumerate(b["bodies"]):
        order, colour = sc.order(i)
        check_colouring(wl, order, colour)
        ob = po.body_from_workload(wl, "ref64")
        ob.substep(po.params_from_settings(b["dt"], wl["settings"]), order, count=2 * S)
        x, v = sc.get_state(i)
        assert np.array_equal(x, ob.X) and np.array_equal(v, ob.V), wl["name"]
        single = workloads.make_batch_scene(ctx, b, bodies=[i], sweep_blocks=True, launch_mode=xf.LAUNCH_BLOCKS)
        single.step(S, dt)
        single.step(S, dt)
        xs, vs = single.get_state(0)
        single.close()
        assert np.array_equal(xs, x) and np.array_equal(vs, v), wl["name"]
    sc.close()


def test_blocks_autotune_keeps_results(ctx):
    wl = workloads.beam_neo_12x4x4(energy=7)
    sc = workloads.make_scene(ctx, wl, sweep_blocks=True)
    x0, v0 = sc.get_state(0)
    mode, size, ms = sc.autotune(wl["substeps"], float(wl["dt"]), reps=2)
    assert mode in (xf.LAUNCH_PER_COLOUR, xf.LAUNCH_PERSISTENT, xf.LAUNCH_BLOCKS) and ms > 0
    x1, v1 = sc.get_state(0)
    assert np.array_equal(x0, x1) and np.array_equal(v0, v1) and sc.tick() == 0
    sc.step(wl["substeps"], float(wl["dt"]))
    x, v = sc.get_state(0)
    order, _ = sc.order(0)
    sc.close()
    ob = oracle(wl, order, wl["substeps"])
    assert np.array_equal(x, ob.X) and np.array_equal(v, ob.V)


def test_blocks_autotune_slot_caps(ctx):
    """The autotuner repacks the sweep-block slots for each slot cap it
    tries: the beam steps bit-exactly at every cap set explicitly, and after
    tuning at the cap it chose (reported by info and in the chosen options)."""
    wl = workloads.beam_yeoh(24, 8, 8)
    S, dt = wl["substeps"], float(wl["dt"])
    ref = None
    slots = {}
    for cap in (32, 64, 128, 256):
        sc = workloads.make_scene(ctx, wl, sweep_blocks=True, launch_mode=xf.LAUNCH_BLOCKS, block_lanes=cap)
        sc.step(S, dt)
        x, v = sc.get_state(0)
        slots[cap] = sc.info()["n_block_slots"]
        assert sc.info()["block_lanes"] == cap
        sc.close()
        if ref is None:
            ref = (x, v)
        assert np.array_equal(x, ref[0]) and np.array_equal(v, ref[1]), cap
    assert slots[32] > slots[256], slots
    sc = workloads.make_scene(ctx, wl, sweep_blocks=True)
    mode, size, ms = sc.autotune(S, dt, reps=1)
    info = sc.info()
    assert info["block_lanes"] in slots and info["n_block_slots"] == slots[info["block_lanes"]]
    sc.step(S, dt)
    x, v = sc.get_state(0)
    sc.close()
    assert np.array_equal(x, ref[0]) and np.array_equal(v, ref[1])


def separate_tets(k):
    """k unit tets 0.3 m apart along x, no pins (positions, tets, density)."""
    one, t4 = xf.tet_block(1, 1, 1, (0.1, 0.1, 0.1), (0.0, 0.0, 0.0))
    p = one.reshape(-1, 3)[t4[:4]]
    xyz = np.concatenate([p + np.float32([0.3 * j, 0.0, 0.0]) for j in range(k)]).ravel().astype(np.float32)
    return xyz, np.arange(4 * k, dtype=np.uint32), 1000.0


def test_blocks_beyond_64k_lds(ctx):
    """A slot of 2,800 nodes (700 separate tets in one block: 89.6 KB of LDS,
    more than 64 KiB) runs on a device whose workgroups may use more, bit
    for bit as the oracle's serial sweep."""
    xyz, tets, rho = separate_tets(700)
    wl = dict(name="tets700", xyz=xyz, tets=tets, density=rho, pinned=None, velocity=None, tet_mu=None,
              settings=dict(energy=xf.ENERGY_MIXED_SEL, simultaneous=1, compliance=1e-5, poissons_ratio=0.5,
                            gravity=(0.0, -9.81)),
              dt=np.float32(1.0) / np.float32(1200.0), substeps=20)
    sc = workloads.make_scene(ctx, wl, sweep_blocks=True, launch_mode=xf.LAUNCH_BLOCKS)
    info = sc.info()
    assert info["block_lds_limit"] >= 65536
    if info["block_lds_limit"] < 2800 * 32:
        sc.close()
        pytest.skip("the device's workgroups hold at most 64 KiB of LDS")
    assert info["n_block_nodes"] == 2800
    sc.step(wl["substeps"], float(wl["dt"]))
    x, v = sc.get_state(0)
    order, _ = sc.order(0)
    sc.close()
    ob = oracle(wl, order, wl["substeps"])
    assert np.array_equal(x, ob.X) and np.array_equal(v, ob.V)


def test_blocks_rejections(ctx):
    wl = workloads.beam_neo_12x4x4()
    cases = [dict(launch_mode=xf.LAUNCH_BLOCKS), dict(sweep_blocks=True, xcd_slabs=True),
             dict(sweep_blocks=2), dict(launch_mode=3), dict(sweep_blocks=True, block_lanes=300)]
    for opts in cases:
        sc = xf.Scene(ctx)
        sc.add_body(wl["xyz"], wl["tets"], wl["density"], pinned=wl["pinned"])
        with pytest.raises(xf.XfError):
            sc.finalize(**opts)
        sc.close()
    sc = xf.Scene(ctx)
    sc.add_body(wl["xyz"], wl["tets"], wl["density"], pinned=wl["pinned"], tet_colour=xf.tet_block_colours(12, 4, 4))
    with pytest.raises(xf.XfError):
        sc.finalize(sweep_blocks=True)
    sc.close()
    # separate tets: every component's levels restart at 0, so all fall in
    # one block; more nodes than a workgroup's LDS holds (32 B each)
    probe = xf.Scene(ctx)
    probe.add_body(*separate_tets(1))
    probe.finalize()
    limit = probe.info()["block_lds_limit"]
    probe.close()
    sc = xf.Scene(ctx)
    sc.add_body(*separate_tets(limit // 32 // 4 + 1))
    with pytest.raises(xf.XfError, match="LDS"):
        sc.finalize(sweep_blocks=True)
    sc.close()
    # vertex-ring energies run per colour only
    sc = workloads.make_scene(ctx, wl, sweep_blocks=True, launch_mode=xf.LAUNCH_BLOCKS)
    sc.set_settings(0, **dict(wl["settings"], energy=xf.ENERGY_CONTINUOUS_SKIN))
    with pytest.raises(xf.XfError):
        sc.step(1, float(wl["dt"]))
    sc.close()


@pytest.mark.parametrize("block_size", [128, 256])
@pytest.mark.parametrize("sphere_dy", [0.0, 0.05])
def test_bench_launch_shape_bit_exact(ctx, sphere_dy, block_size):
    """The headline run's exact launch shape on the 96x24x24 Yeoh beam with
    ground and moving sphere: sweep blocks, 128- or 256-thread workgroups (the
    shapes the autotuner picks there, within 1 % of each other), 64-lane slots,
    48 bytes per node in LDS, the frame captured in a hipGraph and replayed,
    the node phase fused into the last block launch of each substep. Two
    frames (capture, then a replay of the same graph) against the oracle's
    serial sweep, bit for bit. sphere_dy = 0 is the workload as benchmarked
    (the sphere reaches the sagging underside later); with the sphere raised
    5 cm into the beam, contacts with kinetic friction start at once."""
    wl = workloads.beam_yeoh_96x24x24()
    s = dict(wl["settings"])
    c0 = np.array(s["sphere_c0"], np.float64) + (0.0, sphere_dy, 0.0)
    s["sphere_c0"] = tuple(c0)
    wl["settings"] = s
    S = wl["substeps"]
    sc = workloads.make_scene(ctx, wl, use_graph=True, block_size=block_size, sweep_blocks=True,
                              launch_mode=xf.LAUNCH_BLOCKS)
    info = sc.info()
    assert info["fused_finish"] == 1 and info["n_colours"] == 24 and info["block_launches"] == 4
    sc.step(S, float(wl["dt"]))
    x1, v1 = sc.get_state(0)
    sc.step(S, float(wl["dt"]))
    x2, v2 = sc.get_state(0)
    order = sc.order(0)[0]
    assert sc.info()["block_node_stride"] == 48  # 264 workgroups per launch: latency-bound layout
    sc.close()
    ob = oracle(wl, order, S)
    print(f"\nbench shape, frame 1: max|dx| vs oracle {np.abs(x1 - ob.X).max():.3e}")
    assert np.array_equal(x1, ob.X) and np.array_equal(v1, ob.V)
    ob.substep(po.params_from_settings(wl["dt"], wl["settings"]), order, count=S)
    assert np.array_equal(x2, ob.X) and np.array_equal(v2, ob.V)
    if sphere_dy > 0.0:
        # nodes held on the surface of the moving sphere
        c = np.float32(c0) + np.float32(s["sphere_v"]) * np.float32(2 * S * wl["dt"])
        d = np.linalg.norm(x2.reshape(-1, 3) - c, axis=1)
        assert np.sum(np.abs(d - s["sphere_r"]) < 1e-4) >= 3


@pytest.mark.parametrize("case", ["wide", "damped", "quad", "serial", "many_slots"])
def test_blocks_node_stride(ctx, case):
    """48 bytes per slot node in LDS where every block launch has at most two
    workgroups per compute unit and the projection is simultaneous MixedSel /
    YeohSkinFast without damping with one lane per tet; 32 otherwise. Both
    layouts bit-exact against the oracle."""
    wl = workloads.beam_neo_12x4x4(energy=7, simultaneous=case != "serial")
    opts = dict(block_size=64)
    if case == "damped":
        wl["settings"] = dict(wl["settings"], damping=0.01)
    if case == "quad":
        opts["tet_lanes"] = 4
    if case == "many_slots":
        # 300 x 8 x 4 cells of 1 cm, one block per workgroup (block_lanes 1):
        # about 900 workgroups per launch colour, more than two per compute unit
        xyz, tets = xf.tet_block(300, 8, 4, (0.01, 0.01, 0.01), (0.0, 0.0, 0.0))
        wl = dict(wl, xyz=xyz, tets=tets, pinned=(xyz.reshape(-1, 3)[:, 0] == 0).astype(np.uint8),
                  velocity=None, tet_colour=None)
        opts["block_lanes"] = 1
    x, v, order, _, info, _, _ = run(ctx, wl, frames=1, **opts)
    want = 48 if case == "wide" else 32
    assert info["block_node_stride"] == want, info
    if case == "many_slots":
        assert info["n_block_slots"] > 4 * 2 * 256
    ob = oracle(wl, order, wl["substeps"])
    assert np.array_equal(x, ob.X) and np.array_equal(v, ob.V)
