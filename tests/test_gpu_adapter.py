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
