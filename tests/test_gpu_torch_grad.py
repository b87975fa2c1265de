"""Gradient check through the thin torch.autograd.Function binding: the
backward (hand-written GPU adjoint) against central finite differences of
the GPU forward along random directions in log mu."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import workloads  # noqa: E402
import xpbdfem as xf  # noqa: E402
import xpbdfem_torch  # noqa: E402


def test_rollout_loss_backward_matches_fd(ctx):
    truth = workloads.beam_fit_16x4x4()
    guess = dict(truth)
    t = truth["tets"].size // 4
    guess["tet_mu"] = np.full(t, 5e4, np.float32)
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
    sc = workloads.make_scene(ctx, guess)
    x0, v0 = sc.get_state(0)
    log_mu = torch.full((t,), float(np.log(5e4)), dtype=torch.float64, requires_grad=True)
    args = (sc, [x0], [v0], n_sub, dt, loss_at, targets, 20)
    L = xpbdfem_torch.RolloutLoss.apply(log_mu, *args)
    L.backward()
    g = log_mu.grad.clone()
    rng = np.random.default_rng(4)
    for _ in range(2):
        r = torch.from_numpy(rng.normal(size=t))
        an = float(g @ r)
        eps = 1e-2
        with torch.no_grad():
            Lp = xpbdfem_torch.RolloutLoss.apply(log_mu + eps * r, *args)
            Lm = xpbdfem_torch.RolloutLoss.apply(log_mu - eps * r, *args)
        fd = float((Lp - Lm) / (2 * eps))
        print(f"\nL {float(L):.4e}: directional backward {an:.6e}, FD {fd:.6e}")
        assert abs(an - fd) <= 2e-2 * abs(fd)
    sc.close()
