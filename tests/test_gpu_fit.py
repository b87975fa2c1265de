"""Batched differentiable fitting (64-beam Yeoh workload, reduced sizes).

* The batched GPU adjoint of Yeoh (YeohSkinFast form with per-body C) bodies
  equals, body by body, the oracle's adjoint at the same precision: loss
  term and dL/dlog(mu) bit for bit.
* A few Adam iterations on log mu lower the loss and the relative mu error.
* At the full size (64 beams, 400,668 tets, 30 frames of 40 substeps, every
  frame in the loss), where the oracle would take minutes: each body's loss
  share and gradient from the batched adjoint equal, bit for bit, those of
  the same body alone in a scene of its own (bodies share no state).
"""
import numpy as np
import pytest

import fitting
import pyoracle as po
import workloads

pytestmark = pytest.mark.gpu


def small_batch(n_bodies=3, frames=2):
    return workloads.fit_batch(n_bodies=n_bodies, seed=9, cells=(4, 6), frames=frames)


def test_batched_yeoh_adjoint_bit_exact_vs_oracle(ctx):
    batch = small_batch()
    S = batch["substeps"]
    loss_at = np.array([S, 2 * S], np.uint32)
    dt = float(batch["dt"])
    bodies = batch["bodies"]
    guess = [dict(wl, tet_mu=wl["mu_guess"]) for wl in bodies]
    sc = workloads.make_batch_scene(ctx, dict(batch, bodies=guess))
    orders = [sc.order(b)[0] for b in range(len(bodies))]
    # targets: each body's own hidden-mu rollout (oracle, reference precision)
    tg_parts = []
    for wl, order in zip(bodies, orders):
        p = po.params_from_settings(dt, wl["settings"])
        b = po.body_from_workload(wl, "ref64")
        out, done = [], 0
        for la in loss_at:
            b.substep(p, order, count=int(la) - done)
            done = int(la)
            out.append(b.X.copy())
        tg_parts.append(np.array(out))
    targets = np.concatenate(tg_parts, axis=1)
    L, g = sc.adjoint(2 * S, dt, loss_at, targets, window=S)
    off, total = 0, 0.0
    for wl, tg, order in zip(guess, tg_parts, orders):
        t = wl["tets"].size // 4
        ob = po.body_from_workload(wl, "ref64")
        Lo, glog = po.rollout_adjoint(ob, po.params_from_settings(dt, wl["settings"]), order, 2 * S, loss_at, tg,
                                      window=S)
        assert np.array_equal(g[off:off + t], -glog), wl["name"]
        total += Lo
        off += t
    print(f"\nbatched loss {L:.12e}, sum of oracle losses {total:.12e}")
    assert np.isclose(L, total, rtol=1e-14)
    sc.close()


def test_adam_lowers_loss_and_mu_error(ctx):
    batch = small_batch(n_bodies=4, frames=3)
    fit = fitting.BatchFit(ctx, batch, loss_frames=(1, 2, 3))
    e0 = fit.mu_error()
    losses = [fit.step() for _ in range(12)]
    e1 = fit.mu_error()
    print(f"\nloss {losses[0]:.4e} -> {losses[-1]:.4e}; median rel mu err {np.median(e0):.4f} -> {np.median(e1):.4f}")
    assert all(np.isfinite(losses))
    assert losses[-1] < 0.5 * losses[0]
    assert np.median(e1) < np.median(e0)
    fit.close()


def test_full_batch_adjoint_equals_single_bodies(ctx):
    batch = workloads.fit_batch()
    fit = fitting.BatchFit(ctx, batch)
    L, g, body_loss = fit.loss_and_grad(per_body=True)
    assert np.isfinite(L) and np.all(np.isfinite(g))
    S = batch["substeps"]
    loss_at = fit.loss_at
    nodes = [wl["xyz"].size // 3 for wl in batch["bodies"]]
    node_off = np.concatenate([[0], np.cumsum(nodes)])
    for b in (0, 31, 63):
        wl = dict(batch["bodies"][b], tet_mu=batch["bodies"][b]["mu_guess"])
        one = workloads.make_batch_scene(ctx, dict(batch, bodies=[wl]))
        tg = fit.targets[:, 3 * node_off[b]:3 * node_off[b + 1]]
        Lb, gb = one.adjoint(fit.n_sub, fit.dt, loss_at, tg, window=fit.window)
        one.close()
        sl = slice(fit.offsets[b], fit.offsets[b + 1])
        print(f"\nbody {b}: batched share {body_loss[b]:.12e}, alone {Lb:.12e}")
        assert Lb == body_loss[b]
        assert np.array_equal(gb, g[sl])
    assert S == 40 and fit.n_sub == 1200 and len(loss_at) == 30
    fit.close()
