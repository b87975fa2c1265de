"""Gradient check through the thin torch.autograd.Function binding.

1. The backward equals the C ABI adjoint (xf_scene_adjoint) bit for bit, and
   the forward returns its loss: the binding adds nothing to the numbers.
2. Directional derivatives of the backward (hand-written GPU adjoint, FP32
   element math) against central finite differences of the FP64 oracle
   rollout, eps = 1e-2 in log mu. Finite differences of the FP32 forward
   itself are rounding noise at every usable eps (measured on this problem:
   8% to 320% off at eps 1e-2, sign flips at 1e-3), so the FP64 model is the
   only meaningful FD reference. The FP64 oracle adjoint matches these
   differences to about 1e-3 (test_oracle_adjoint.py); the FP32 gradient sits
   within DIR_TOL of them (measured 3.6% and 6.8% in the two directions below).
3. The same with the FP64 element model (fp64=True, xf_scene_adjoint_fp64):
   equal to the FP64 oracle adjoint and within 1e-3 of the FP64 differences.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import pyoracle as po  # noqa: E402
import workloads  # noqa: E402
import xpbdfem as xf  # noqa: E402
import xpbdfem_torch  # noqa: E402

DIR_TOL = 0.08


def test_rollout_loss_backward(ctx):
    truth = workloads.beam_fit_16x4x4()
    t = truth["tets"].size // 4
    n = truth["xyz"].size // 3
    S, dt = truth["substeps"], float(truth["dt"])
    n_sub = 10 * S
    loss_at = np.array([5 * S, 10 * S], np.uint32)
    ref = workloads.make_scene(ctx, truth)
    targets, done = [], 0
    for la in loss_at:
        ref.step(int(la) - done, dt)
        done = int(la)
        targets.append(ref.get_state(0)[0])
    ref.close()
    targets = np.array(targets)

    guess = dict(truth)
    guess["tet_mu"] = np.full(t, 5e4, np.float32)
    sc = workloads.make_scene(ctx, guess)
    x0, v0 = sc.get_state(0)
    log_mu = torch.full((t,), float(np.log(5e4)), dtype=torch.float64, requires_grad=True)
    L = xpbdfem_torch.RolloutLoss.apply(log_mu, sc, [x0], [v0], n_sub, dt, loss_at, targets, 20)
    L.backward()
    g = log_mu.grad.detach().numpy().copy()

    direct = workloads.make_scene(ctx, guess)
    Ld, gd = direct.adjoint(n_sub, dt, loss_at, targets, window=20)
    direct.close()
    assert float(L.detach()) == Ld
    assert np.array_equal(g, gd)

    order, _, _ = xf.mesh_colour(n, truth["tets"])
    p = po.params_from_settings(dt, truth["settings"])

    def loss64(lm):
        wl = dict(truth)
        wl["tet_mu"] = np.exp(lm).astype(np.float32)
        b = po.body_from_workload(wl, "f64")
        Lo, _ = po.rollout_adjoint(b, p, order, n_sub, loss_at, targets, window=20)
        return Lo

    rng = np.random.default_rng(4)
    lm0 = np.full(t, np.log(5e4))
    eps = 1e-2
    for _ in range(2):
        r = rng.normal(size=t)
        an = float(g @ r)
        fd = (loss64(lm0 + eps * r) - loss64(lm0 - eps * r)) / (2 * eps)
        print(f"\nL {Ld:.4e}: directional backward {an:.6e}, FP64 FD {fd:.6e}, rel {abs(an - fd) / abs(fd):.3e}")
        assert abs(an - fd) <= DIR_TOL * abs(fd)

    # FP64 element model through the same binding
    log_mu64 = torch.full((t,), float(np.log(5e4)), dtype=torch.float64, requires_grad=True)
    L64 = xpbdfem_torch.RolloutLoss.apply(log_mu64, sc, [x0], [v0], n_sub, dt, loss_at, targets, 20, True)
    L64.backward()
    g64 = log_mu64.grad.detach().numpy().copy()
    wl = dict(truth)
    wl["tet_mu"] = np.full(t, 5e4, np.float32)
    Lo, glog = po.rollout_adjoint(po.body_from_workload(wl, "f64"), p, order, n_sub, loss_at, targets, window=20)
    assert float(L64.detach()) == Lo and np.array_equal(g64, -glog)
    eps = 1e-3
    for _ in range(2):
        r = rng.normal(size=t)
        an = float(g64 @ r)
        fd = (loss64(lm0 + eps * r) - loss64(lm0 - eps * r)) / (2 * eps)
        print(f"FP64 model: directional backward {an:.6e}, FP64 FD {fd:.6e}, rel {abs(an - fd) / abs(fd):.3e}")
        assert abs(an - fd) <= 1e-3 * abs(fd)
    sc.close()
