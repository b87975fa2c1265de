"""xf_scene_adjoint_noalloc against xf_scene_adjoint through the C ABI: the
same loss, per-body loss terms, dL/dlog(mu) and end state, bit for bit, on
two calls in a row that share the device scratch, the second on a smaller
scene, and a clear error for a problem beyond the scratch."""
import numpy as np
import pytest

import workloads
import xpbdfem as xf

pytestmark = pytest.mark.gpu


def fit_scenes(ctx, wl):
    """Two identical scenes of a workload with a uniform guess of mu."""
    guess = dict(wl, tet_mu=np.full(wl["tets"].size // 4, 5e4, np.float32))
    return workloads.make_scene(ctx, guess), workloads.make_scene(ctx, guess)


def targets(wl, loss_at, seed):
    """Targets near the rest positions (any targets define a loss)."""
    rng = np.random.default_rng(seed)
    x0 = wl["xyz"].astype(np.float64)
    return np.array([x0 + rng.normal(0.0, 0.01, x0.size) for _ in loss_at])


def run_pair(ctx, wl, substeps, loss_at, window, seed):
    a, b = fit_scenes(ctx, wl)
    loss_at = np.array(loss_at, np.uint32)
    tg = targets(wl, loss_at, seed)
    dt = float(wl["dt"])
    La, ga, ba = a.adjoint(substeps, dt, loss_at, tg, window=window, per_body=True)
    Lb, gb, bb = b.adjoint(substeps, dt, loss_at, tg, window=window, per_body=True, noalloc=True)
    xa, va = a.get_state(0)
    xb, vb = b.get_state(0)
    ta, tb = a.time(), b.time()
    a.close()
    b.close()
    print(f"\n{wl['name']}: {substeps} substeps, loss {La:.12e} / {Lb:.12e}, |grad| {np.linalg.norm(ga):.4e}")
    assert La == Lb and np.array_equal(ba, bb)
    assert np.array_equal(ga, gb) and np.linalg.norm(ga) > 0
    assert np.array_equal(xa, xb) and np.array_equal(va, vb) and ta == tb


def test_noalloc_equals_adjoint_two_calls(ctx):
    big = workloads.beam_fit_16x4x4()
    small = workloads.beam_neo_12x4x4()
    assert small["tets"].size < big["tets"].size and small["xyz"].size < big["xyz"].size
    run_pair(ctx, big, 120, (40, 80, 120), 20, seed=1)
    run_pair(ctx, small, 45, (20, 45), 16, seed=2)


def test_noalloc_rejects_beyond_scratch(ctx):
    wl = workloads.beam_yeoh(48, 12, 12)
    wl = dict(wl, settings=dict(wl["settings"], collide=xf.COLLIDE_GROUND))
    sc = workloads.make_scene(ctx, wl)
    loss_at = np.array([60], np.uint32)
    with pytest.raises(xf.XfError, match="bytes of scratch"):
        sc.adjoint(60, float(wl["dt"]), loss_at, np.zeros((1, wl["xyz"].size)), window=20, noalloc=True)
    sc.close()
