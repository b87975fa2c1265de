"""The rest of TSolveElement's energy switch on the GPU (Pixar 0/1/2, Cube
11/12 -- no constraint on tets --, the default for 13..31), with every
Rayleigh type and PBD for Pixar: GPU == the oracle's serial sweep (itself
bit-exact with the compiled reference, test_oracle_dispatch), bit for bit."""
import numpy as np
import pytest

import pyoracle as po
import workloads
import xpbdfem as xf

pytestmark = pytest.mark.gpu


def beam(energy, sim=True, nu=0.45, rtype=0, damp=0.0, pbd=0.0, pbda=0.0, passes=0):
    wl = workloads.beam_neo_12x4x4(energy=energy, simultaneous=sim)
    wl["settings"] = dict(wl["settings"], compliance=1e-6 if energy <= 2 else 2e-5, poissons_ratio=nu,
                          damping=damp, rayleigh_type=rtype, pbd_damping=pbd, pbd_damping_amortized=pbda,
                          volume_passes=passes)
    return wl


def check(ctx, wl, tick=3, **opts):
    sc = workloads.make_scene(ctx, wl, **opts)
    sc.set_tick(tick)
    S = wl["substeps"]
    sc.step(S, float(wl["dt"]))
    sc.step(S, float(wl["dt"]))
    x, v = sc.get_state(0)
    order = sc.order(0)[0]
    sc.close()
    ob = po.body_from_workload(wl, "ref64")
    ob.tick = tick
    ob.substep(po.params_from_settings(wl["dt"], wl["settings"]), order, count=2 * S)
    print(f"\nenergy {wl['settings']['energy']}: max|dx| vs oracle {np.abs(x - ob.X).max():.3e}, "
          f"moved {np.abs(x - wl['xyz']).max():.3e}")
    assert np.array_equal(x, ob.X) and np.array_equal(v, ob.V)


@pytest.mark.parametrize("energy", [0, 1, 2, 11, 12, 13, 31])
@pytest.mark.parametrize("sim", [True, False])
def test_energy_dispatch_bit_exact(ctx, energy, sim):
    check(ctx, beam(energy, sim))


@pytest.mark.parametrize("energy", [0, 2])
@pytest.mark.parametrize("rtype,damp,pbd,pbda", [(0, 0.02, 0.0, 0.0), (1, 0.02, 0.0, 0.0), (2, 0.02, 0.003, 0.0),
                                                 (3, 0.02, 0.003, 0.012)])
def test_pixar_damping_bit_exact(ctx, energy, rtype, damp, pbd, pbda):
    check(ctx, beam(energy, True, 0.45, rtype, damp, pbd, pbda))


@pytest.mark.parametrize("energy", [2, 12])
def test_dispatch_with_volume_passes_and_slabs(ctx, energy):
    check(ctx, beam(energy, True, 0.45, 3, 0.02, 0.003, 0.012, passes=1), xcd_slabs=True, block_size=128)


def test_dispatch_rejections(ctx):
    wl = beam(0)
    sc = workloads.make_scene(ctx, wl, launch_mode=xf.LAUNCH_PERSISTENT)
    with pytest.raises(xf.XfError):
        sc.step(1, float(wl["dt"]))
    sc.close()
    sc = workloads.make_scene(ctx, dict(wl, tet_mu=np.full(wl["tets"].size // 4, 1e5, np.float32)))
    with pytest.raises(xf.XfError):
        sc.step(1, float(wl["dt"]))
    sc.close()
    sc = workloads.make_scene(ctx, wl)
    with pytest.raises(xf.XfError):
        sc.set_settings(0, energy=32)
    sc.close()
