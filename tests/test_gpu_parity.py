"""GPU parity: the HIP path (through the C ABI) against the oracle.

Bar: bit-exact. With the same colour order, GPU positions and velocities
equal those of
  * the oracle's reference-precision sweep (xo_substep_ref64: FP64 node
    state, FP32 element math), and
  * the compiled reference itself (oracle/_ref/libxfref.so driving the
    reference's GeoLinear3d::Substep), where that build is present,
for every energy (3-7) and both solve modes. The error against the FP64
oracle is reported; it is bounded by FP64_REPORT_BOUND only to catch gross
failures, since the substep amplifies rounding (the reference's own result
differs from FP64 by ~4e-4 of the bbox diagonal after one frame).
"""
import numpy as np
import pytest

import pyoracle as po
import workloads
import xpbdfem as xf

pytestmark = pytest.mark.gpu

FP64_REPORT_BOUND = 5e-3


def bbox_diag(xyz):
    p = xyz.reshape(-1, 3)
    return float(np.linalg.norm(p.max(0) - p.min(0)))


def rel_err(x, ref, xyz0):
    return float(np.abs(x.reshape(-1, 3) - ref.reshape(-1, 3)).max() / bbox_diag(xyz0))


def run_gpu(ctx, wl, substeps, use_graph=True, colour_order=xf.COLOUR_NATURAL, chunks=1,
            launch_mode=xf.LAUNCH_PER_COLOUR, persistent_blocks=0, **opts):
    sc = workloads.make_scene(ctx, wl, use_graph=use_graph, colour_order=colour_order, launch_mode=launch_mode,
                              persistent_blocks=persistent_blocks, **opts)
    for _ in range(chunks):
        sc.step(substeps, float(wl["dt"]))
    x, v = sc.get_state(0)
    order, _ = sc.order(0)
    sc.close()
    return x, v, order


def run_oracle(wl, order, substeps, variant="ref64"):
    ob = po.body_from_workload(wl, variant=variant)
    ob.substep(po.params_from_settings(wl["dt"], wl["settings"]), order, count=substeps)
    return ob


@pytest.mark.parametrize("energy", [3, 4, 5, 6, 7])
@pytest.mark.parametrize("simultaneous", [True, False])
def test_one_frame_bit_exact(ctx, energy, simultaneous):
    wl = workloads.beam_neo_12x4x4(energy=energy, simultaneous=simultaneous)
    x, v, order = run_gpu(ctx, wl, wl["substeps"])
    ob = run_oracle(wl, order, wl["substeps"])
    print(f"\nenergy {energy} sim {simultaneous}: max|dx| vs oracle {np.abs(x - ob.X).max():.3e}")
    assert np.array_equal(x, ob.X)
    assert np.array_equal(v, ob.V)


@pytest.mark.skipif(not po.have_ref(), reason="reference build (oracle/_ref) not present")
@pytest.mark.parametrize("energy", [4, 7])
def test_frames_bit_exact_vs_compiled_reference(ctx, energy):
    """No collisions (the reference has none): 5 frames, GPU == reference."""
    wl = workloads.beam_neo_12x4x4(energy=energy, collide=False)
    x, v, order = run_gpu(ctx, wl, wl["substeps"], chunks=5)
    rb = po.RefBody(wl["xyz"], wl["tets"], wl["density"], wl["pinned"], order)
    s = wl["settings"]
    rb.substeps(rb.settings(s["compliance"], nu=s["poissons_ratio"], energy=energy), wl["dt"], 5 * wl["substeps"])
    xr, vr = rb.state()
    assert np.array_equal(x, xr)
    assert np.array_equal(v, vr)


def test_120_frames_bit_exact_and_fp64_error(ctx):
    wl = workloads.beam_neo_12x4x4()
    sc = workloads.make_scene(ctx, wl)
    order, _ = sc.order(0)
    ob = po.body_from_workload(wl, "ref64")
    f64 = po.body_from_workload(wl, "f64")
    p = po.params_from_settings(wl["dt"], wl["settings"])
    for frame in range(1, wl["frames"] + 1):
        sc.step(wl["substeps"], float(wl["dt"]))
        ob.substep(p, order, count=wl["substeps"])
        f64.substep(p, order, count=wl["substeps"])
        if frame in (1, 2, 5, 10, 30, 60, 120):
            x, v = sc.get_state(0)
            assert np.array_equal(x, ob.X), f"frame {frame}"
            assert np.array_equal(v, ob.V), f"frame {frame}"
            e = rel_err(x, f64.X, wl["xyz"])
            print(f"\nframe {frame}: rel pos err vs FP64 oracle {e:.3e}; min y {x[1::3].min():.4f}")
            if frame == 1:
                assert e <= FP64_REPORT_BOUND
    assert x[1::3].min() >= wl["settings"]["ground_y"] - 1e-9
    sc.close()


def test_graph_equals_eager(ctx):
    wl = workloads.beam_neo_12x4x4(energy=7)
    xg, vg, _ = run_gpu(ctx, wl, 2 * wl["substeps"], use_graph=True)
    xe, ve, _ = run_gpu(ctx, wl, 2 * wl["substeps"], use_graph=False)
    assert np.array_equal(xg, xe) and np.array_equal(vg, ve)


def test_graph_replays_match_eager(ctx):
    """Several replays of one captured frame, with host reads between them."""
    wl = workloads.beam_neo_12x4x4()
    sg = workloads.make_scene(ctx, wl, use_graph=True)
    se = workloads.make_scene(ctx, wl, use_graph=False)
    for _ in range(4):
        sg.step(wl["substeps"], float(wl["dt"]))
        se.step(wl["substeps"], float(wl["dt"]))
        xg, vg = sg.get_state(0)
        xe, ve = se.get_state(0)
        assert np.array_equal(xg, xe) and np.array_equal(vg, ve)
    sg.close()
    se.close()


def test_minstd_colouring_bit_exact(ctx):
    wl = workloads.beam_neo_12x4x4()
    x, v, order = run_gpu(ctx, wl, wl["substeps"], colour_order=xf.COLOUR_MINSTD)
    ob = run_oracle(wl, order, wl["substeps"])
    assert np.array_equal(x, ob.X) and np.array_equal(v, ob.V)


@pytest.mark.parametrize("block", [64, 128, 256])
def test_block_sizes_bit_exact(ctx, block):
    wl = workloads.beam_neo_12x4x4(energy=7)
    sc = workloads.make_scene(ctx, wl, block_size=block)
    sc.step(wl["substeps"], float(wl["dt"]))
    x, v = sc.get_state(0)
    order, _ = sc.order(0)
    ob = run_oracle(wl, order, wl["substeps"])
    assert np.array_equal(x, ob.X) and np.array_equal(v, ob.V)
    sc.close()


def test_big_yeoh_beam_one_frame_bit_exact(ctx):
    """The 96x24x24 Yeoh beam (331,776 tets) with ground and moving sphere."""
    wl = workloads.beam_yeoh_96x24x24()
    x, v, order = run_gpu(ctx, wl, wl["substeps"])
    ob = run_oracle(wl, order, wl["substeps"])
    print(f"\nbig beam: max|dx| vs oracle {np.abs(x - ob.X).max():.3e}")
    assert np.array_equal(x, ob.X)
    assert np.array_equal(v, ob.V)


@pytest.mark.parametrize("energy,sim", [(4, True), (7, True), (4, False), (3, True)])
@pytest.mark.parametrize("blocks", [0, 64])
def test_persistent_launch_bit_exact(ctx, energy, sim, blocks):
    wl = workloads.beam_neo_12x4x4(energy=energy, simultaneous=sim)
    x, v, order = run_gpu(ctx, wl, wl["substeps"], chunks=2, launch_mode=xf.LAUNCH_PERSISTENT,
                          persistent_blocks=blocks)
    ob = run_oracle(wl, order, 2 * wl["substeps"])
    assert np.array_equal(x, ob.X) and np.array_equal(v, ob.V)


def test_persistent_big_beam_bit_exact(ctx):
    wl = workloads.beam_yeoh_96x24x24()
    x, v, order = run_gpu(ctx, wl, wl["substeps"], launch_mode=xf.LAUNCH_PERSISTENT)
    ob = run_oracle(wl, order, wl["substeps"])
    assert np.array_equal(x, ob.X) and np.array_equal(v, ob.V)


@pytest.mark.parametrize("energy,sim", [(4, True), (7, True), (4, False), (3, True), (7, False)])
@pytest.mark.parametrize("passes", [1, 2])
@pytest.mark.parametrize("mode", [xf.LAUNCH_PER_COLOUR, xf.LAUNCH_PERSISTENT])
def test_volume_passes_bit_exact(ctx, energy, sim, passes, mode):
    """Settings::volumePasses sweeps after the main colour sweep, with ground
    contact on: GPU == oracle (itself == the compiled reference)."""
    wl = workloads.beam_neo_12x4x4(energy=energy, simultaneous=sim)
    wl["settings"] = dict(wl["settings"], volume_passes=passes)
    x, v, order = run_gpu(ctx, wl, wl["substeps"], chunks=2, launch_mode=mode)
    ob = run_oracle(wl, order, 2 * wl["substeps"])
    assert np.array_equal(x, ob.X) and np.array_equal(v, ob.V)


def test_user_colouring_validated_and_bit_exact(ctx):
    wl = workloads.beam_neo_12x4x4()
    bad = np.zeros(wl["tets"].size // 4, np.uint32)  # one colour for all: shares nodes
    sc = xf.Scene(ctx)
    with pytest.raises(xf.XfError):
        sc.add_body(wl["xyz"], wl["tets"], wl["density"], pinned=wl["pinned"], tet_colour=bad)
    sc.close()
    wl["tet_colour"] = xf.tet_block_colours(12, 4, 4)
    x, v, order = run_gpu(ctx, wl, wl["substeps"])
    ob = run_oracle(wl, order, wl["substeps"])
    assert np.array_equal(x, ob.X) and np.array_equal(v, ob.V)


def check_slab_order(sc, wl):
    """XCD-slab schedule: still colour-major; each colour cut into eighths
    [n r / 8, n (r + 1) / 8) that are slabs in centroid x (no tet of eighth r
    lies right of one of eighth r + 1), node-index ordered inside."""
    order, colour = sc.order(0)
    t = wl["tets"].reshape(-1, 4).astype(np.int64)
    x = wl["xyz"].reshape(-1, 3)[:, 0].astype(np.float32)[t]
    cx = ((x[:, 0] + x[:, 1]) + x[:, 2]) + x[:, 3]
    first = t.min(axis=1)
    co = colour[order]
    assert np.all(np.diff(co.astype(np.int64)) >= 0)
    for c in np.unique(co):
        seg = order[co == c]
        n = seg.size
        cuts = [(n * r) >> 3 for r in range(9)]
        for r in range(8):
            part = seg[cuts[r]:cuts[r + 1]]
            assert np.all(np.diff(first[part]) >= 0)
            if r < 7 and part.size and cuts[r + 2] > cuts[r + 1]:
                assert cx[part].max() <= cx[seg[cuts[r + 1]:cuts[r + 2]]].min()


@pytest.mark.parametrize("energy,sim", [(4, True), (7, True), (4, False), (3, True)])
@pytest.mark.parametrize("block", [64, 256])
@pytest.mark.parametrize("passes", [0, 1])
def test_xcd_slabs_bit_exact(ctx, energy, sim, block, passes):
    """Per-colour launches with the XCD-slab block mapping (workgroup b on
    eighth b mod 8 of the colour; colours of 40 tets leave eighths of 5, with
    256-thread blocks almost all lanes idle): same results as the oracle's
    sweep in the slab schedule's order."""
    wl = workloads.beam_neo_12x4x4(energy=energy, simultaneous=sim)
    wl["settings"] = dict(wl["settings"], volume_passes=passes)
    sc = workloads.make_scene(ctx, wl, xcd_slabs=True, block_size=block)
    check_slab_order(sc, wl)
    sc.step(wl["substeps"], float(wl["dt"]))
    sc.step(wl["substeps"], float(wl["dt"]))
    x, v = sc.get_state(0)
    order, _ = sc.order(0)
    sc.close()
    ob = run_oracle(wl, order, 2 * wl["substeps"])
    assert np.array_equal(x, ob.X) and np.array_equal(v, ob.V)


def test_xcd_slabs_big_beam_bit_exact(ctx):
    wl = workloads.beam_yeoh_96x24x24()
    sc = workloads.make_scene(ctx, wl, xcd_slabs=True)
    check_slab_order(sc, wl)
    sc.step(wl["substeps"], float(wl["dt"]))
    x, v = sc.get_state(0)
    order, _ = sc.order(0)
    sc.close()
    ob = run_oracle(wl, order, wl["substeps"])
    assert np.array_equal(x, ob.X) and np.array_equal(v, ob.V)


def test_xcd_slabs_flag_validated(ctx):
    wl = workloads.beam_neo_12x4x4()
    sc = xf.Scene(ctx)
    sc.add_body(wl["xyz"], wl["tets"], wl["density"], pinned=wl["pinned"])
    with pytest.raises(xf.XfError):
        sc.finalize(xcd_slabs=2)
    sc.close()
