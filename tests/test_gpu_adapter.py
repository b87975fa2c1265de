"""The C++ drop-in adapter (xpbd-fem_amd/csrc/xf_geo_gpu.h), driven with a
Settings-shaped struct and one Substep call per substep like the
reference's Sim::Update, equals the oracle's serial sweep bit-for-bit."""
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


@pytest.mark.parametrize("energy", [4, 7])
def test_adapter_substep_loop_bit_exact(tmp_path, energy):
    assert os.path.exists(DEMO), "build the adapter demo first (__graft_entry__.build())"
    out = tmp_path / "state.bin"
    frames, substeps = 2, 20
    print(f"\nrunning {DEMO} energy {energy}", flush=True)
    r = subprocess.run([DEMO, str(out), str(frames), str(substeps), str(energy)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    wl = workloads.beam_neo_12x4x4(energy=energy)
    n = wl["xyz"].size // 3
    data = np.fromfile(out, np.float64)
    X, V = data[:3 * n], data[3 * n:]
    order, _, _ = xf.mesh_colour(n, wl["tets"], xf.COLOUR_NATURAL)
    ob = po.body_from_workload(wl, "ref64")
    ob.substep(po.params_from_settings(np.float32(1.0) / np.float32(1200.0), wl["settings"]), order,
               count=frames * substeps)
    assert np.array_equal(X, ob.X)
    assert np.array_equal(V, ob.V)


@pytest.mark.parametrize("energy,rtype,damp", [(7, 0, 0.0), (4, 0, 0.02), (7, 2, 0.02)])
def test_adapter_sweep_blocks_bit_exact(ctx, tmp_path, energy, rtype, damp):
    """GeoGpuLinear3d::SetSweepBlocks: the adapter's scene with sweep blocks
    and block launches (with in-projection and post Rayleigh damping); a scene
    of the same mesh and options gives the order the oracle replays."""
    assert os.path.exists(DEMO), "build the adapter demo first (__graft_entry__.build())"
    out = tmp_path / "state.bin"
    frames, substeps = 2, 20
    print(f"\nrunning {DEMO} energy {energy} rtype {rtype} with sweep blocks", flush=True)
    r = subprocess.run([DEMO, str(out), str(frames), str(substeps), str(energy), str(rtype), str(damp), "0", "0", "0",
                        "1"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    wl = workloads.beam_neo_12x4x4(energy=energy)
    wl["settings"] = dict(wl["settings"], damping=damp, rayleigh_type=rtype)
    n = wl["xyz"].size // 3
    data = np.fromfile(out, np.float64)
    X, V = data[:3 * n], data[3 * n:]
    sc = workloads.make_scene(ctx, wl, sweep_blocks=True, launch_mode=xf.LAUNCH_BLOCKS)
    order = sc.order(0)[0]
    sc.close()
    ob = po.body_from_workload(wl, "ref64")
    ob.substep(po.params_from_settings(np.float32(1.0) / np.float32(1200.0), wl["settings"]), order,
               count=frames * substeps)
    assert np.array_equal(X, ob.X)
    assert np.array_equal(V, ob.V)


@pytest.mark.parametrize("energy", [9, 10])
def test_adapter_ring_energies_bit_exact(ctx, tmp_path, energy):
    """Settings energy bits 9 / 10 through the adapter run the vertex-ring
    sweep; the ring order of a scene with the same mesh and options (vertex
    visit order, no slabs) replays it on the oracle."""
    assert os.path.exists(DEMO), "build the adapter demo first (__graft_entry__.build())"
    out = tmp_path / "state.bin"
    frames, substeps = 2, 20
    print(f"\nrunning {DEMO} energy {energy}", flush=True)
    r = subprocess.run([DEMO, str(out), str(frames), str(substeps), str(energy)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    wl = workloads.beam_neo_12x4x4(energy=energy)
    n = wl["xyz"].size // 3
    data = np.fromfile(out, np.float64)
    X, V = data[:3 * n], data[3 * n:]
    sc = workloads.make_scene(ctx, wl)
    order, rorder = sc.order(0)[0], sc.ring_order(0)[0]
    sc.close()
    ob = po.body_from_workload(wl, "ref64")
    ob.substep(po.params_from_settings(np.float32(1.0) / np.float32(1200.0), wl["settings"]), order,
               count=frames * substeps, ring_order=rorder)
    assert np.abs(X - wl["xyz"]).max() > 1e-4
    assert np.array_equal(X, ob.X)
    assert np.array_equal(V, ob.V)


def _autoresize_golden():
    return np.load(os.path.join(ROOT, "tests", "golden", "ref_autoresize.npz"))


def test_scene_fp64_positions_autoresize_bit_exact(ctx):
    """A body given FP64 (autoresized) positions: element constants and node
    state from them reproduce the reference's Init(autoResize) trajectory."""
    g = _autoresize_golden()
    sc = xf.Scene(ctx)
    sc.add_body(None, g["tets"], 1000.0, pinned=g["pinned"], xyz64=xf.autoresize(g["xyz"]))
    sc.finalize()
    sc.set_settings(0, energy=4, simultaneous=1, compliance=1e-5, poissons_ratio=0.5)
    x0, _ = sc.get_state(0)
    assert np.array_equal(x0, g["X0"])
    assert np.array_equal(sc.order(0)[0], g["order"])
    sc.step(40, float(np.float32(1.0) / np.float32(1200.0)))
    x, v = sc.get_state(0)
    sc.close()
    assert np.array_equal(x, g["X"]) and np.array_equal(v, g["V"])


def test_adapter_autoresize_bit_exact(tmp_path):
    """GeoGpuLinear3d::Init(..., autoResize = true) driven like Sim::Update."""
    g = _autoresize_golden()
    out = tmp_path / "state.bin"
    print(f"\nrunning {DEMO} with autoResize", flush=True)
    r = subprocess.run([DEMO, str(out), "2", "20", "4", "0", "0", "0", "0", "1"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    n = g["xyz"].size // 3
    data = np.fromfile(out, np.float64)
    assert np.array_equal(data[:3 * n], g["X"]) and np.array_equal(data[3 * n:], g["V"])
