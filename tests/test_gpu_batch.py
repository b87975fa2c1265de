"""Batched scenes: many bodies swept together (colour c of every body in one
launch, per-body energy, material and ground) must equal each body run in a
scene of its own, bit-for-bit, and the oracle's sweep of that body."""
import numpy as np
import pytest

import pyoracle as po
import workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small_batch():
    return workloads.boxes_batch(n_bodies=6, seed=5, cells=(3, 6), frames=2)


def test_batched_equals_individual_scenes(ctx, small_batch):
    b = small_batch
    dt, S = float(b["dt"]), b["substeps"]
    batched = workloads.make_batch_scene(ctx, b)
    assert batched.info()["n_bodies"] == len(b["bodies"])
    batched.step(S, dt)
    batched.step(S, dt)
    for i, wl in enumerate(b["bodies"]):
        single = workloads.make_batch_scene(ctx, b, bodies=[i])
        single.step(S, dt)
        single.step(S, dt)
        xs, vs = single.get_state(0)
        xb, vb = batched.get_state(i)
        assert np.array_equal(xs, xb), wl["name"]
        assert np.array_equal(vs, vb), wl["name"]
        single.close()
    batched.close()


def test_batched_body_equals_oracle(ctx, small_batch):
    b = small_batch
    dt, S = float(b["dt"]), b["substeps"]
    sc = workloads.make_batch_scene(ctx, b)
    sc.step(S, dt)
    for i, wl in enumerate(b["bodies"]):
        order, _ = sc.order(i)
        ob = po.body_from_workload(wl, "ref64")
        ob.substep(po.params_from_settings(b["dt"], wl["settings"]), order, count=S)
        x, v = sc.get_state(i)
        assert np.array_equal(x, ob.X), wl["name"]
        assert np.array_equal(v, ob.V), wl["name"]
        assert x[1::3].min() >= -1e-12  # ground plane y = 0
    sc.close()


def test_full_batch_runs_finite(ctx):
    b = workloads.boxes_batch()
    sc = workloads.make_batch_scene(ctx, b)
    info = sc.info()
    assert info["n_bodies"] == 256 and info["n_tets"] > 3_000_000
    sc.step(b["substeps"], float(b["dt"]))
    sc.sync()
    assert sc.count_nonfinite() == 0
    sc.close()


def test_autotune_restores_state_and_keeps_results(ctx, small_batch):
    b = small_batch
    dt, S = float(b["dt"]), b["substeps"]
    tuned = workloads.make_batch_scene(ctx, b)
    plain = workloads.make_batch_scene(ctx, b)
    before = [tuned.get_state(i) for i in range(len(b["bodies"]))]
    t0 = tuned.time()
    mode, size, ms = tuned.autotune(S, dt, reps=2)
    assert mode in (0, 1) and size in (64, 128, 256, 512) and ms > 0
    assert tuned.time() == t0
    for i in range(len(b["bodies"])):
        x, v = tuned.get_state(i)
        assert np.array_equal(x, before[i][0]) and np.array_equal(v, before[i][1])
    tuned.step(S, dt)
    plain.step(S, dt)
    for i in range(len(b["bodies"])):
        xa, va = tuned.get_state(i)
        xb, vb = plain.get_state(i)
        assert np.array_equal(xa, xb) and np.array_equal(va, vb)
    tuned.close()
    plain.close()


def test_persistent_batch_equals_per_colour(ctx, small_batch):
    b = small_batch
    dt, S = float(b["dt"]), b["substeps"]
    pc = workloads.make_batch_scene(ctx, b)
    pe = workloads.make_batch_scene(ctx, b, launch_mode=1)
    pc.step(S, dt)
    pe.step(S, dt)
    for i in range(len(b["bodies"])):
        xa, va = pc.get_state(i)
        xb, vb = pe.get_state(i)
        assert np.array_equal(xa, xb) and np.array_equal(va, vb)
    pc.close()
    pe.close()


def test_xcd_slabs_batch_equals_oracle(ctx, small_batch):
    """Batched per-colour launches with the XCD-slab mapping: every body equals
    the oracle's sweep in its reported (slab) order."""
    b = small_batch
    dt, S = float(b["dt"]), b["substeps"]
    sc = workloads.make_batch_scene(ctx, b, xcd_slabs=True, block_size=128)
    sc.step(S, dt)
    for i, wl in enumerate(b["bodies"]):
        order, _ = sc.order(i)
        ob = po.body_from_workload(wl, "ref64")
        ob.substep(po.params_from_settings(b["dt"], wl["settings"]), order, count=S)
        x, v = sc.get_state(i)
        assert np.array_equal(x, ob.X), wl["name"]
        assert np.array_equal(v, ob.V), wl["name"]
    sc.close()
