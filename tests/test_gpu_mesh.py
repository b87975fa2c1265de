"""The GPU substep on irregular tet meshes read from tetgen files (jittered
Delaunay meshes: irregular node valence, greedy colourings of 30+ colours),
bit for bit against the oracle (itself == the compiled reference on the
reference's Armadillo mesh, test_mesh_io.py)."""
import numpy as np
import pytest

import pyoracle as po
import workloads
import xpbdfem as xf

pytestmark = pytest.mark.gpu

scipy_spatial = pytest.importorskip("scipy.spatial")


def delaunay_workload(tmp_path, seed, energy):
    rng = np.random.default_rng(seed)
    g = np.stack(np.meshgrid(np.arange(12), np.arange(5), np.arange(5), indexing="ij"), -1).reshape(-1, 3) * 0.1
    p = g + rng.uniform(-0.02, 0.02, g.shape)
    p[:, 0] -= p[:, 0].min()
    p[:, 1] -= 0.2
    p[:, 2] -= 0.2
    d = scipy_spatial.Delaunay(p)
    node, ele = xf.write_tetgen(tmp_path / f"mesh{seed}", p, d.simplices)
    xyz, tets = xf.read_tetgen(node, ele, orient=True)
    wl = workloads.beam_neo_12x4x4(energy=energy)
    wl.update(xyz=xyz, tets=tets, pinned=(xyz[0::3] < 0.03).astype(np.uint8), tet_colour=None)
    if energy == 7:
        wl["settings"] = dict(wl["settings"], compliance=0.1095 / 1e5)
    return wl


@pytest.mark.parametrize("seed", [0, 1])
@pytest.mark.parametrize("energy", [4, 7])
@pytest.mark.parametrize("mode", [xf.LAUNCH_PER_COLOUR, xf.LAUNCH_PERSISTENT])
def test_irregular_mesh_bit_exact(ctx, tmp_path, seed, energy, mode):
    wl = delaunay_workload(tmp_path, seed, energy)
    sc = workloads.make_scene(ctx, wl, launch_mode=mode)
    S = wl["substeps"]
    sc.step(S, float(wl["dt"]))
    sc.step(S, float(wl["dt"]))
    x, v = sc.get_state(0)
    order, _ = sc.order(0)
    ncol = sc.info()["n_colours"]
    sc.close()
    ob = po.body_from_workload(wl, "ref64")
    ob.substep(po.params_from_settings(wl["dt"], wl["settings"]), order, count=2 * S)
    print(f"\nseed {seed}: {wl['tets'].size // 4} tets, {ncol} colours; max|dx| {np.abs(x - ob.X).max():.3e}")
    assert np.all(np.isfinite(x))
    assert np.array_equal(x, ob.X) and np.array_equal(v, ob.V)


@pytest.mark.parametrize("energy", [4, 7])
def test_irregular_mesh_sweep_blocks_bit_exact(ctx, tmp_path, energy):
    """Sweep blocks on the jittered Delaunay mesh: irregular blocks (coarse
    breadth-first levels, up to ~34 phases per launch colour) still sweep in
    the reported order."""
    wl = delaunay_workload(tmp_path, 0, energy)
    sc = workloads.make_scene(ctx, wl, sweep_blocks=True, launch_mode=xf.LAUNCH_BLOCKS)
    S = wl["substeps"]
    sc.step(S, float(wl["dt"]))
    x, v = sc.get_state(0)
    order, colour = sc.order(0)
    info = sc.info()
    sc.close()
    t = wl["tets"].reshape(-1, 4)
    pairs = np.stack([np.repeat(colour.astype(np.int64), 4), t.ravel().astype(np.int64)], 1)
    assert len(np.unique(pairs, axis=0)) == len(pairs)  # colours node-disjoint
    ob = po.body_from_workload(wl, "ref64")
    ob.substep(po.params_from_settings(wl["dt"], wl["settings"]), order, count=S)
    print(f"\n{t.shape[0]} tets: {info['n_blocks']} blocks, {info['n_colours']} phases per substep")
    assert np.array_equal(x, ob.X) and np.array_equal(v, ob.V)
