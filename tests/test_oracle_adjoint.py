"""The oracle's FP64 adjoint against central finite differences.

The adjoint has no reference counterpart; it is pinned by differentiation:
  * one projection: vector-Jacobian products in positions and in log(comp0)
    against central FD of the FP64 forward;
  * whole rollouts of the 16x4x4 fitting beam (per-tet mu, ground contact,
    pinned face): a directional derivative along a random direction in
    log mu, and per-tet derivatives on 32 seeded tets with relative step
    1e-4, against central FD of the FP64 forward.
"""
import numpy as np
import pytest

import pyoracle as po
import workloads
import xpbdfem as xf

DT = float(np.float32(1.0) / np.float32(1200.0))


def rest_tet(rng):
    r = np.eye(3) * 0.1 + rng.normal(0, 0.02, (3, 3))
    if np.linalg.det(r) < 0:
        r[0] *= -1
    return np.vstack([r, np.zeros((1, 3))]).astype(np.float32)


@pytest.mark.parametrize("energy", [4, 7])
@pytest.mark.parametrize("seed", range(4))
def test_tet_adjoint_matches_fd(energy, seed):
    rng = np.random.default_rng(seed)
    rest = rest_tet(rng)
    el = po.init_elements(rest.reshape(-1), np.arange(4, dtype=np.uint32), 1000.0)
    F = np.eye(3) + rng.normal(0, 0.08, (3, 3))
    if energy == 7:
        F *= 1.05
    X = (rest.astype(np.float64) @ F.T).reshape(-1)
    w = rng.uniform(0.5, 2.0, 4).astype(np.float32)
    w[seed % 4] = 0.0 if seed == 3 else w[seed % 4]  # one case with a pinned node
    comp0, comp1, a = 0.02, 0.0, 1.0
    p = po.make_params(DT, energy=energy)

    def fwd(Xin, c0=comp0):
        Xo = Xin.copy()
        po.lib().xo_solve_tet_f64(p, Xo, w, el["qi"], el["qq"], el["qr"], c0, comp1, a)
        return Xo

    xbar = rng.normal(size=12)
    xb = xbar.copy()
    glog = po.lib().xo_adjoint_tet_f64(p, X, xb, w, el["qi"], el["qq"], el["qr"], comp0, comp1, a)
    d = rng.normal(size=12)
    fd = xbar @ (fwd(X + 1e-6 * d) - fwd(X - 1e-6 * d)) / 2e-6
    assert abs(xb @ d - fd) <= 1e-6 * abs(fd) + 1e-12
    e = 1e-3  # comp0 crosses the interface as FP32: keep the step well above its resolution
    fdc = xbar @ (fwd(X, comp0 * np.exp(e)) - fwd(X, comp0 * np.exp(-e))) / (2 * e)
    assert abs(glog - fdc) <= 1e-4 * abs(fdc)


@pytest.fixture(scope="module")
def fit():
    wl = workloads.beam_fit_16x4x4()
    n = wl["xyz"].size // 3
    order, _, _ = xf.mesh_colour(n, wl["tets"])
    p = po.params_from_settings(wl["dt"], wl["settings"])

    def make(mu):
        w2 = dict(wl)
        w2["tet_mu"] = np.asarray(mu, np.float32)
        return po.body_from_workload(w2, "f64")

    return wl, order, p, make


def targets_for(fit, loss_at):
    wl, order, p, make = fit
    b = make(wl["tet_mu"])
    out, done = [], 0
    for la in loss_at:
        b.substep(p, order, count=int(la) - done)
        done = int(la)
        out.append(b.X.copy())
    return np.array(out)


def test_rollout_adjoint_directional_fd(fit):
    wl, order, p, make = fit
    t = wl["tets"].size // 4
    loss_at = np.array([200], np.uint32)  # 10 frames of 20 substeps
    targets = targets_for(fit, loss_at)
    mu0 = np.full(t, 5e4, np.float32)
    L, glog = po.rollout_adjoint(make(mu0), p, order, 200, loss_at, targets)
    r = np.random.default_rng(1).normal(size=t)
    an = -(glog * r).sum()
    eps = 1e-3
    Lp, _ = po.rollout_adjoint(make(mu0 * np.exp(eps * r)), p, order, 200, loss_at, targets)
    Lm, _ = po.rollout_adjoint(make(mu0 * np.exp(-eps * r)), p, order, 200, loss_at, targets)
    fd = (Lp - Lm) / (2 * eps)
    print(f"\nL {L:.4e}: directional adjoint {an:.6e}, FD {fd:.6e}")
    assert abs(an - fd) <= 1e-3 * abs(fd)


def per_tet_fd(fit, mu0, loss_at, targets, tets, n_sub=600, rel_step=1e-4):
    """Central finite differences dL/dmu_k of the FP64 oracle's forward loss
    for each tet k (relative step rel_step, the perturbed mu rounded to FP32
    as the material stores it)."""
    wl, order, p, make = fit
    out = []
    for k in tets:
        mp, mm = mu0.copy(), mu0.copy()
        mp[k] = np.float32(mu0[k] * (1 + rel_step))
        mm[k] = np.float32(mu0[k] * (1 - rel_step))
        Lp = po.rollout_loss(make(mp), p, order, n_sub, loss_at, targets)
        Lm = po.rollout_loss(make(mm), p, order, n_sub, loss_at, targets)
        out.append((Lp - Lm) / (float(mp[k]) - float(mm[k])))
    return np.array(out)


def fd_tets(t):
    """The 32 seeded random tets of the gradient check."""
    return np.random.default_rng(32).choice(t, 32, replace=False)


def test_rollout_loss_matches_adjoint_loss(fit):
    wl, order, p, make = fit
    t = wl["tets"].size // 4
    loss_at = np.array([200, 400, 600], np.uint32)
    targets = targets_for(fit, loss_at)
    mu0 = np.full(t, 5e4, np.float32)
    L, _ = po.rollout_adjoint(make(mu0), p, order, 600, loss_at, targets)
    assert po.rollout_loss(make(mu0), p, order, 600, loss_at, targets) == L


def test_rollout_adjoint_per_tet_fd(fit):
    """30 frames, loss at frames 10, 20, 30; 32 seeded tets, relative step
    1e-4, central differences of the FP64 forward."""
    wl, order, p, make = fit
    t = wl["tets"].size // 4
    loss_at = np.array([200, 400, 600], np.uint32)
    targets = targets_for(fit, loss_at)
    mu0 = np.full(t, 5e4, np.float32)
    _, glog = po.rollout_adjoint(make(mu0), p, order, 600, loss_at, targets)
    an = -glog / mu0
    ks = fd_tets(t)
    fd = per_tet_fd(fit, mu0, loss_at, targets, ks)
    rel = np.abs(an[ks] - fd) / np.abs(fd)
    print("\nper-tet relative errors:", " ".join(f"{x:.1e}" for x in rel))
    assert np.median(rel) <= 1e-3
    assert rel.max() <= 5e-3
