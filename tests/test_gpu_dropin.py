"""GeoGpu (xpbd-fem_amd/csrc/GeoGpu.h), the Geo subclass a maintainer drops
into the reference's tree, against the reference's own GeoLinear3d: the
compiled test (tests/cpp/geo_dropin_test.cpp) holds both in one Geo* array,
transforms them as Sim::FinishAddingBlocks does under
Settings_Rotate90Degrees, and steps them as Sim::Update does, with the
reference's first-substep pins, Settings_LockRight (fixed and rotating),
Settings_LockLeft cleared mid-run, the manipulator and the damping passes.
Bar: X and V bit-identical, CalculateVolume and Pick equal. The kinematic,
manipulator and pin entry points of the C ABI reject what they do not
cover."""
import os
import subprocess

import numpy as np
import pytest

import workloads
import xpbdfem as xf

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DROPIN = os.path.join(ROOT, "tests", "cpp", "_build", "geo_dropin_test")


def test_geo_gpu_in_the_reference_sim():
    if not os.path.exists(DROPIN):
        pytest.skip("geo_dropin_test is built only where the reference sources are present")
    r = subprocess.run([DROPIN, "400"], capture_output=True, text=True, timeout=240)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("bit-identical") == 7 and "ALL OK" in r.stdout


def test_kinematic_and_manip_arguments(ctx):
    wl = workloads.beam_neo_12x4x4()
    sc = workloads.make_scene(ctx, wl)
    n = wl["xyz"].size // 3
    with pytest.raises(xf.XfError, match="out of range"):
        sc.set_kinematic(0, [n], [0.0, 0.0, 0.0])
    with pytest.raises(xf.XfError, match="twice"):
        sc.set_kinematic(0, [3, 3], np.zeros(6))
    with pytest.raises(xf.XfError, match="non-finite"):
        sc.set_kinematic(0, [3], [0.0, np.nan, 0.0])
    with pytest.raises(xf.XfError, match="out of range"):
        sc.set_manip(0, n, (0.0, 0.0, 0.0))
    # held nodes stay exactly at their targets and rest
    tgt = wl["xyz"].reshape(-1, 3)[[5, 9]].astype(np.float64) + 0.01
    sc.set_kinematic(0, [5, 9], tgt)
    sc.step(wl["substeps"], float(wl["dt"]))
    x, v = sc.get_state(0)
    assert np.array_equal(x.reshape(-1, 3)[[5, 9]], tgt) and not np.any(v.reshape(-1, 3)[[5, 9]])
    # the adjoint covers neither
    L = np.array([1], np.uint32)
    with pytest.raises(xf.XfError, match="kinematic"):
        sc.adjoint(1, float(wl["dt"]), L, np.zeros((1, 3 * n)))
    sc.set_kinematic(0, [], [])
    sc.set_manip(0, 7, (0.0, 0.0, 0.0))
    with pytest.raises(xf.XfError, match="manipulator"):
        sc.adjoint(1, float(wl["dt"]), L, np.zeros((1, 3 * n)))
    sc.close()
