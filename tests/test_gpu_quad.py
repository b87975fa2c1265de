"""Sweep-block launches with four lanes per tet (xf_scene_opts.tet_lanes = 4):
each tet's per-node work is split over the four lanes of a quad, the
per-tet scalars computed by all four alike. Bar: bit-exact with the
oracle's serial sweep in the reported order and with one lane per tet, for
MixedSel and YeohSkinFast, one and many bodies, phases with more tets than
a workgroup has quads, and the fused node phase."""
import numpy as np
import pytest

import pyoracle as po
import workloads
import xpbdfem as xf

pytestmark = pytest.mark.gpu


def run(ctx, wl, frames=2, **opts):
    opts.setdefault("sweep_blocks", True)
    opts.setdefault("launch_mode", xf.LAUNCH_BLOCKS)
    sc = workloads.make_scene(ctx, wl, **opts)
    for _ in range(frames):
        sc.step(wl["substeps"], float(wl["dt"]))
    x, v = sc.get_state(0)
    order, _ = sc.order(0)
    info = sc.info()
    sc.close()
    return x, v, order, info


def oracle(wl, order, substeps):
    ob = po.body_from_workload(wl, "ref64")
    ob.substep(po.params_from_settings(wl["dt"], wl["settings"]), order, count=substeps)
    return ob


@pytest.mark.parametrize("energy", [4, 7])
@pytest.mark.parametrize("block_size", [64, 128, 256])
def test_quad_bit_exact(ctx, energy, block_size):
    """12x4x4 beam: at 64 threads a workgroup has 16 quads for phases of up
    to 24 tets (quads take several tets in turn); ground contact; the last
    block launch of each substep ends it (fused node phase)."""
    wl = workloads.beam_neo_12x4x4(energy=energy)
    x, v, order, info = run(ctx, wl, tet_lanes=4, block_size=block_size)
    assert info["tet_lanes"] == 4 and info["fused_finish"] == 1
    ob = oracle(wl, order, 2 * wl["substeps"])
    assert np.array_equal(x, ob.X) and np.array_equal(v, ob.V)
    x1, v1, order1, info1 = run(ctx, wl, tet_lanes=1, block_size=block_size)
    assert info1["tet_lanes"] == 1 and np.array_equal(order1, order)
    assert np.array_equal(x1, x) and np.array_equal(v1, v)


def test_quad_separate_finish(ctx):
    wl = workloads.beam_neo_12x4x4(energy=7)
    x, v, order, info = run(ctx, wl, tet_lanes=4, block_size=256, separate_finish=True)
    assert info["fused_finish"] == 0
    ob = oracle(wl, order, 2 * wl["substeps"])
    assert np.array_equal(x, ob.X) and np.array_equal(v, ob.V)


def test_quad_big_beam_bench_shape(ctx):
    """The 96x24x24 Yeoh beam with ground and moving sphere (raised 5 cm into
    the beam so contacts start at once), 256-thread workgroups, graph
    replay: one frame bit-exact with the oracle."""
    wl = workloads.beam_yeoh_96x24x24()
    s = dict(wl["settings"])
    s["sphere_c0"] = (s["sphere_c0"][0], s["sphere_c0"][1] + 0.05, s["sphere_c0"][2])
    wl["settings"] = s
    x, v, order, info = run(ctx, wl, frames=1, tet_lanes=4, block_size=256, use_graph=True)
    assert info["fused_finish"] == 1 and info["n_colours"] == 24
    ob = oracle(wl, order, wl["substeps"])
    assert np.array_equal(x, ob.X) and np.array_equal(v, ob.V)


def test_quad_batch_mixed_energies(ctx):
    """Many bodies with MixedSel and YeohSkinFast mixed (energy chosen per
    body, per-tet material slots): every body equals the oracle."""
    b = workloads.boxes_batch(n_bodies=6, seed=5, cells=(3, 6), frames=2)
    dt, S = float(b["dt"]), b["substeps"]
    assert len({wl["settings"]["energy"] for wl in b["bodies"]}) == 2
    sc = workloads.make_batch_scene(ctx, b, sweep_blocks=True, launch_mode=xf.LAUNCH_BLOCKS, tet_lanes=4,
                                    block_size=128)
    assert sc.info()["fused_finish"] == 1
    sc.step(S, dt)
    sc.step(S, dt)
    for i, wl in enumerate(b["bodies"]):
        order, _ = sc.order(i)
        ob = po.body_from_workload(wl, "ref64")
        ob.substep(po.params_from_settings(b["dt"], wl["settings"]), order, count=2 * S)
        x, v = sc.get_state(i)
        assert np.array_equal(x, ob.X) and np.array_equal(v, ob.V), wl["name"]
    sc.close()


def test_quad_more_tets_than_quads(ctx):
    """A mirrored 4x4x150 block: phases of up to 75 tets on 64-thread
    workgroups (16 quads), so each quad takes up to 5 tets of a phase."""
    xyz, tets = xf.tet_block(4, 4, 150, (0.05, 0.05, 0.05), (0.0, 0.0, 0.0), mirrored=True)
    wl = workloads.beam_neo_12x4x4(energy=4)
    wl = dict(wl, xyz=xyz, tets=tets, pinned=(xyz.reshape(-1, 3)[:, 0] == 0).astype(np.uint8), velocity=None,
              tet_colour=None)
    x, v, order, _ = run(ctx, wl, frames=1, tet_lanes=4, block_size=64)
    ob = oracle(wl, order, wl["substeps"])
    assert np.array_equal(x, ob.X) and np.array_equal(v, ob.V)


@pytest.mark.parametrize("settings", [dict(energy=5), dict(simultaneous=0), dict(damping=0.01)])
def test_quad_rejects_other_projections(ctx, settings):
    wl = workloads.beam_neo_12x4x4()
    sc = workloads.make_scene(ctx, wl, sweep_blocks=True, launch_mode=xf.LAUNCH_BLOCKS, tet_lanes=4)
    sc.set_settings(0, **dict(wl["settings"], **settings))
    with pytest.raises(xf.XfError, match="tet_lanes"):
        sc.step(1, float(wl["dt"]))
    sc.close()
