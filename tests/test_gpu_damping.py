"""Damping on the GPU: in-projection Rayleigh (Paper, Limit), Rayleigh post
passes (Post, PostAmortized) and PBD damping (GeoLinear3d::Damp,
Geo.cpp:790-811; Xpbd.h:87-348), bit for bit against the oracle (itself
bit-exact with the compiled reference, test_oracle_damping.py)."""
import os
import subprocess

import numpy as np
import pytest

import pyoracle as po
import workloads
import xpbdfem as xf

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEMO = os.path.join(ROOT, "tests", "cpp", "_build", "geo_adapter_demo")


def damped(energy, sim, rtype, damping=0.02, pbd=0.003, pbd_amortized=0.012):
    wl = workloads.beam_neo_12x4x4(energy=energy, simultaneous=sim)
    wl["settings"] = dict(wl["settings"], damping=damping, rayleigh_type=rtype, pbd_damping=pbd,
                          pbd_damping_amortized=pbd_amortized)
    return wl


def oracle_run(wl, order, substeps, tick=0):
    ob = po.body_from_workload(wl, "ref64")
    ob.tick = tick
    ob.substep(po.params_from_settings(wl["dt"], wl["settings"]), order, count=substeps)
    return ob


@pytest.mark.parametrize("rtype", [0, 1, 2, 3])
@pytest.mark.parametrize("energy,sim", [(4, True), (7, True), (4, False), (3, True), (7, False)])
@pytest.mark.parametrize("mode", [xf.LAUNCH_PER_COLOUR, xf.LAUNCH_PERSISTENT])
def test_damping_bit_exact(ctx, rtype, energy, sim, mode):
    wl = damped(energy, sim, rtype)
    sc = workloads.make_scene(ctx, wl, launch_mode=mode)
    sc.set_tick(5)
    S = wl["substeps"]
    sc.step(S, float(wl["dt"]))
    sc.step(S, float(wl["dt"]))
    x, v = sc.get_state(0)
    assert sc.tick() == 5 + 2 * S
    order, _ = sc.order(0)
    sc.close()
    ob = oracle_run(wl, order, 2 * S, tick=5)
    print(f"\nrtype {rtype} energy {energy} sim {sim} mode {mode}: max|dx| {np.abs(x - ob.X).max():.3e}")
    assert np.array_equal(x, ob.X) and np.array_equal(v, ob.V)


@pytest.mark.parametrize("rtype", [0, 2, 3])
def test_damping_xcd_slabs_bit_exact(ctx, rtype):
    """Post-damping colour kernels under the XCD-slab block mapping."""
    wl = damped(7, True, rtype)
    sc = workloads.make_scene(ctx, wl, xcd_slabs=True)
    S = wl["substeps"]
    sc.step(S, float(wl["dt"]))
    x, v = sc.get_state(0)
    order, _ = sc.order(0)
    sc.close()
    ob = oracle_run(wl, order, S)
    assert np.array_equal(x, ob.X) and np.array_equal(v, ob.V)


@pytest.mark.parametrize("kind", ["rayleigh", "pbd"])
@pytest.mark.parametrize("rtype", [2, 3])
def test_single_damping_pass_bit_exact(ctx, kind, rtype):
    wl = damped(4, True, rtype, damping=0.02 if kind == "rayleigh" else 0.0,
                pbd=0.003 if kind == "pbd" else 0.0)
    x, v, order = None, None, None
    sc = workloads.make_scene(ctx, wl)
    sc.step(wl["substeps"], float(wl["dt"]))
    x, v = sc.get_state(0)
    order, _ = sc.order(0)
    sc.close()
    ob = oracle_run(wl, order, wl["substeps"])
    assert np.array_equal(x, ob.X) and np.array_equal(v, ob.V)


def test_batched_mixed_damping_equals_single_scenes(ctx):
    """Bodies with different damping settings in one scene: each equals its
    own scene (the amortized 1/8 is per body, the tick is shared)."""
    specs = [(4, 3, 0.02, 0.003), (7, 2, 0.05, 0.0), (4, 0, 0.01, 0.002), (7, 1, 0.03, 0.0)]
    bodies = []
    for k, (energy, rtype, d, pbd) in enumerate(specs):
        wl = damped(energy, True, rtype, damping=d, pbd=pbd)
        if energy == 7:
            wl["settings"]["compliance"] = 0.1095 / 1e5
        xyz = wl["xyz"].reshape(-1, 3).copy()
        xyz[:, 2] += 2.0 * k
        wl["xyz"] = xyz.reshape(-1)
        bodies.append(wl)
    batch = dict(bodies=bodies, dt=bodies[0]["dt"], substeps=20)
    both = workloads.make_batch_scene(ctx, batch)
    both.set_tick(3)
    both.step(20, float(batch["dt"]))
    both.step(20, float(batch["dt"]))
    for i, wl in enumerate(bodies):
        one = workloads.make_scene(ctx, wl)
        one.set_tick(3)
        one.step(20, float(batch["dt"]))
        one.step(20, float(batch["dt"]))
        x1, v1 = one.get_state(0)
        xb, vb = both.get_state(i)
        assert np.array_equal(x1, xb) and np.array_equal(v1, vb), i
        ob = oracle_run(wl, one.order(0)[0], 40, tick=3)
        assert np.array_equal(x1, ob.X) and np.array_equal(v1, ob.V), i
        one.close()
    both.close()


def test_adjoint_rejects_damping(ctx):
    wl = damped(4, True, 3)
    sc = workloads.make_scene(ctx, wl)
    n = wl["xyz"].size // 3
    with pytest.raises(xf.XfError):
        sc.adjoint(20, float(wl["dt"]), np.array([20], np.uint32), np.zeros((1, 3 * n)))
    sc.close()


@pytest.mark.parametrize("rtype", [0, 3])
def test_adapter_with_damping_bit_exact(tmp_path, rtype):
    """The Geo-shaped adapter maps Settings flag bits 20-21, damping, the PBD
    factors and tickId (advanced once per Substep call)."""
    assert os.path.exists(DEMO), "build the adapter demo first (__graft_entry__.build())"
    out = tmp_path / "state.bin"
    frames, substeps = 2, 20
    print(f"\nrunning {DEMO} with rayleigh type {rtype}", flush=True)
    r = subprocess.run([DEMO, str(out), str(frames), str(substeps), "4", str(rtype), "0.02", "0.003", "0.012"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    wl = damped(4, True, rtype)
    n = wl["xyz"].size // 3
    data = np.fromfile(out, np.float64)
    order, _, _ = xf.mesh_colour(n, wl["tets"], xf.COLOUR_NATURAL)
    ob = oracle_run(wl, order, frames * substeps)
    assert np.array_equal(data[:3 * n], ob.X) and np.array_equal(data[3 * n:], ob.V)
