"""Sweep-block launches with multiply-add contraction (xf_scene_opts.contract):
the sweep-block kernels compiled with -ffp-contract=fast, so the element
chain and the node phase use fused multiply-adds. Not bit-exact with the
reference's serial FP32 sweep (the oracle's ref64 variant, bit-identical to
the reference), so the bar is a tolerance, in the order the scene reports:

  * relative position error, |x - x_ref| / |x_ref| over all nodes (2-norms),
    at most 1e-4 after the frames run;
  * the largest node error, relative to the largest displacement of the run,
    at most ten times that of the FP64 element model (the oracle's f64
    variant) from the same reference. Contraction moves the result by about
    as much as the reference's own FP32 rounding does: measured ratios 0.97
    to 4.0 over the cases below, the largest on the 3x3x3 box, whose few
    nodes all touch the ground. A dropped term or a skipped phase moves
    nodes by a sizeable part of their displacement, tens to hundreds of
    times this bound's base.

The contracted kernels must have run (the results differ from the exact
ones)."""
import numpy as np
import pytest

import pyoracle as po
import workloads
import xpbdfem as xf

pytestmark = pytest.mark.gpu

POS_REL = 1e-4  # |x - x_ref| / |x_ref|, 2-norms over all nodes
FP64_FACTOR = 10.0  # max node error / max displacement, against the FP64 element model's


def errors(x, xr, x0):
    """(relative position error over all nodes, largest node error over the
    largest displacement)"""
    x, xr, x0 = (np.asarray(a, np.float64).reshape(-1, 3) for a in (x, xr, x0))
    pos = float(np.linalg.norm(x - xr) / np.linalg.norm(xr))
    disp = float(np.linalg.norm(x - xr, axis=1).max() / max(np.linalg.norm(xr - x0, axis=1).max(), 1e-30))
    return pos, disp


def check(wl, x, x0, order, substeps, label):
    """Bounds of the module docstring; returns the errors."""
    params = po.params_from_settings(wl["dt"], wl["settings"])
    ref = po.body_from_workload(wl, "ref64")
    ref.substep(params, order, count=substeps)
    f64 = po.body_from_workload(wl, "f64")
    f64.substep(params, order, count=substeps)
    pos, disp = errors(x, ref.X, x0)
    pos64, disp64 = errors(f64.X, ref.X, x0)
    print(f"\n{label}: contracted vs reference {pos:.2e} (rel. position), {disp:.2e} (of max displacement); "
          f"FP64 element model vs reference {pos64:.2e}, {disp64:.2e}")
    assert not np.array_equal(x, ref.X), "the contracted kernels did not run"
    assert pos <= POS_REL
    assert disp <= FP64_FACTOR * disp64
    return pos, disp


def run(ctx, wl, frames, **opts):
    opts.setdefault("sweep_blocks", True)
    opts.setdefault("launch_mode", xf.LAUNCH_BLOCKS)
    sc = workloads.make_scene(ctx, wl, **opts)
    x0, _ = sc.get_state(0)
    for _ in range(frames):
        sc.step(wl["substeps"], float(wl["dt"]))
    x, v = sc.get_state(0)
    order, _ = sc.order(0)
    info = sc.info()
    sc.close()
    return x0, x, v, order, info


def oracle(wl, order, substeps):
    ob = po.body_from_workload(wl, "ref64")
    ob.substep(po.params_from_settings(wl["dt"], wl["settings"]), order, count=substeps)
    return ob


@pytest.mark.parametrize("energy", [4, 7])
@pytest.mark.parametrize("block_size", [64, 256])
def test_contract_within_tolerance(ctx, energy, block_size):
    """12x4x4 beam, two frames with ground contact; the last block launch of
    each substep ends it (fused node phase, contracted too)."""
    wl = workloads.beam_neo_12x4x4(energy=energy)
    x0, x, v, order, info = run(ctx, wl, 2, block_size=block_size, contract=True)
    assert info["contract"] == 1 and info["fused_finish"] == 1
    check(wl, x, x0, order, 2 * wl["substeps"], f"energy {energy}, {block_size} threads")
    assert np.all(np.isfinite(v))


@pytest.mark.parametrize("block_size", [128, 256])
def test_contract_bench_beam(ctx, block_size):
    """The 96x24x24 Yeoh beam with ground and moving sphere (raised 5 cm so
    contacts start at once), the bench's workgroup sizes, graph replay, one
    frame."""
    wl = workloads.beam_yeoh_96x24x24()
    s = dict(wl["settings"])
    s["sphere_c0"] = (s["sphere_c0"][0], s["sphere_c0"][1] + 0.05, s["sphere_c0"][2])
    wl["settings"] = s
    x0, x, _, order, info = run(ctx, wl, 1, block_size=block_size, use_graph=True, contract=True)
    assert info["contract"] == 1 and info["n_colours"] == 24
    check(wl, x, x0, order, wl["substeps"], f"96x24x24, {block_size} threads, one frame")


def test_contract_batch(ctx):
    """Six boxes, MixedSel and YeohSkinFast mixed, per-tet material slots."""
    b = workloads.boxes_batch(n_bodies=6, seed=5, cells=(3, 6), frames=2)
    dt, S = float(b["dt"]), b["substeps"]
    sc = workloads.make_batch_scene(ctx, b, sweep_blocks=True, launch_mode=xf.LAUNCH_BLOCKS, block_size=128,
                                    contract=True)
    assert sc.info()["contract"] == 1
    x0 = [sc.get_state(i)[0] for i in range(len(b["bodies"]))]
    sc.step(S, dt)
    sc.step(S, dt)
    for i, wl in enumerate(b["bodies"]):
        order, _ = sc.order(i)
        x, _ = sc.get_state(i)
        check(dict(wl, dt=b["dt"]), x, x0[i], order, 2 * S, wl["name"])
    sc.close()


def test_contract_leaves_per_colour_exact(ctx):
    """Per-colour launches do not use the contracted kernels: bit-exact."""
    wl = workloads.beam_neo_12x4x4(energy=7)
    _, x, v, order, info = run(ctx, wl, 1, sweep_blocks=False, launch_mode=xf.LAUNCH_PER_COLOUR, contract=True)
    assert info["contract"] == 0
    ob = oracle(wl, order, wl["substeps"])
    assert np.array_equal(x, ob.X) and np.array_equal(v, ob.V)


def test_contract_rejected_by_adjoint(ctx):
    wl = workloads.beam_neo_12x4x4(energy=7)
    sc = workloads.make_scene(ctx, wl, sweep_blocks=True, launch_mode=xf.LAUNCH_BLOCKS, contract=True)
    n = sc.get_state(0)[0].size
    with pytest.raises(xf.XfError, match="contract"):
        sc.adjoint(4, float(wl["dt"]), np.array([4], np.uint32), np.zeros(n), window=2)
    sc.close()
