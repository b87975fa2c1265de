"""Batched differentiable fitting of per-tet stiffness (no reference
counterpart; the reference has no gradients).

All bodies of a batch live in one scene, so every forward substep and every
adjoint sweep advances all of them in the same launches. Each iteration
resets the bodies to their initial state, runs xf_scene_adjoint (forward
rollout with checkpoints, then the reverse sweep) and takes one Adam step in
log mu on the host (a few hundred thousand scalars; the rollouts dominate),
projected onto the batch's mu range (the prior range of the hidden fields).

An optional smoothness prior adds, per body, smooth * mean over face-adjacent
tet pairs of (log mu_i - log mu_j)^2 to the loss (per-tet stiffness from a
few position snapshots is poorly determined where the body barely deforms;
the prior lets such tets follow their neighbours).

A body whose rollout goes non-finite at some iterate (the reference's solve
can turn unstable for strongly heterogeneous stiffness) is stepped back to
its previous iterate and continues with half its step size; the other
bodies are unaffected, since bodies share no state.

    python fitting.py --iters 200 --out fit.json
"""
import argparse
import json
import sys
import time

import numpy as np

import workloads
import xpbdfem as xf


class BatchFit:
    """Targets from the hidden-mu rollout, then Adam on log mu."""

    def __init__(self, ctx, batch, loss_frames=None, window=None, lr=0.05, betas=(0.9, 0.999), eps=1e-8,
                 smooth=0.0):
        self.batch = batch
        self.dt = float(batch["dt"])
        S = batch["substeps"]
        self.n_sub = batch["frames"] * S
        if loss_frames is None:
            loss_frames = range(1, batch["frames"] + 1)
        self.loss_frames = tuple(loss_frames)
        self.loss_at = np.array([f * S for f in loss_frames], np.uint32)
        self.window = int(window or S)
        self.lr, self.betas, self.eps = lr, betas, eps
        bodies = batch["bodies"]
        self.tets_per_body = [b["tets"].size // 4 for b in bodies]
        self.mu_true = np.concatenate([b["tet_mu"] for b in bodies]).astype(np.float64)

        truth = workloads.make_batch_scene(ctx, batch)
        nodes = [b["xyz"].size // 3 for b in bodies]
        tg = np.zeros((self.loss_at.size, 3 * sum(nodes)))
        done = 0
        for j, la in enumerate(self.loss_at):
            truth.step(int(la) - done, self.dt)
            done = int(la)
            off = 0
            for b, n in enumerate(nodes):
                tg[j, 3 * off:3 * (off + n)] = truth.get_state(b)[0]
                off += n
        truth.close()
        self.targets = tg

        guess = dict(batch)
        guess["bodies"] = [dict(b, tet_mu=b["mu_guess"]) for b in bodies]
        self.scene = workloads.make_batch_scene(ctx, guess)
        self.x0 = [self.scene.get_state(b) for b in range(len(bodies))]
        self.log_mu = np.log(np.concatenate([b["mu_guess"] for b in bodies]).astype(np.float64))
        lo, hi = batch.get("mu_range", (0.0, np.inf))
        self.log_bounds = (np.log(lo) if lo > 0 else -np.inf, np.log(hi))
        self.m = np.zeros_like(self.log_mu)
        self.v = np.zeros_like(self.log_mu)
        self.it = 0
        self.offsets = np.concatenate([[0], np.cumsum(self.tets_per_body)])
        self.body_of_tet = np.repeat(np.arange(len(bodies)), self.tets_per_body)
        self.lr_scale = np.ones(len(bodies))
        self.prev_log_mu = None
        self.failures = 0
        self.smooth = smooth
        pairs, weights = [], []
        for b, wl in enumerate(bodies):
            p = face_pairs(wl["tets"]) + self.offsets[b]
            pairs.append(p)
            weights.append(np.full(len(p), 1.0 / max(len(p), 1)))
        self.pairs = np.concatenate(pairs)
        self.pair_w = np.concatenate(weights)

    def _load(self, log_mu):
        mu = np.exp(log_mu).astype(np.float32)
        off = 0
        for b, t in enumerate(self.tets_per_body):
            self.scene.set_tet_mu(b, mu[off:off + t])
            self.scene.set_state(b, *self.x0[b])
            off += t
        self.scene.set_time(0.0)

    def loss_and_grad(self, per_body=False):
        self._load(self.log_mu)
        return self.scene.adjoint(self.n_sub, self.dt, self.loss_at, self.targets, window=self.window,
                                  per_body=per_body)

    def forward(self):
        """Forward rollout only (the adjoint's forward part), for timing."""
        self._load(self.log_mu)
        self.scene.step(self.n_sub, self.dt)
        self.scene.sync()

    def step(self):
        """One Adam iteration; returns the summed loss of the bodies whose
        rollout stayed finite."""
        _, g, body_loss = self.loss_and_grad(per_body=True)
        if self.smooth > 0.0:
            i, j = self.pairs[:, 0], self.pairs[:, 1]
            d = self.log_mu[i] - self.log_mu[j]
            gs = 2.0 * self.smooth * self.pair_w * d
            g = g + np.bincount(i, gs, g.size) - np.bincount(j, gs, g.size)
            body_loss = body_loss + np.bincount(self.body_of_tet[i], self.smooth * self.pair_w * d * d,
                                                body_loss.size)
        ok = np.isfinite(body_loss)
        for b in np.flatnonzero(ok):
            ok[b] = np.all(np.isfinite(g[self.offsets[b]:self.offsets[b + 1]]))
        if not ok.all():
            if self.prev_log_mu is None:
                raise FloatingPointError("non-finite rollout at the initial guess")
            for b in np.flatnonzero(~ok):
                sl = slice(self.offsets[b], self.offsets[b + 1])
                self.log_mu[sl] = self.prev_log_mu[sl]
                self.lr_scale[b] *= 0.5
                g[sl] = 0.0
            self.failures += int((~ok).sum())
        self.prev_log_mu = self.log_mu.copy()
        self.it += 1
        b1, b2 = self.betas
        live = ok[self.body_of_tet]
        self.m[live] = b1 * self.m[live] + (1 - b1) * g[live]
        self.v[live] = b2 * self.v[live] + (1 - b2) * g[live] * g[live]
        mh = self.m / (1 - b1 ** self.it)
        vh = self.v / (1 - b2 ** self.it)
        lr = self.lr * self.lr_scale[self.body_of_tet]
        self.log_mu[live] -= (lr * mh / (np.sqrt(vh) + self.eps))[live]
        np.clip(self.log_mu, *self.log_bounds, out=self.log_mu)
        return float(body_loss[ok].sum())

    def mu_error(self):
        """Per-tet relative error |mu / mu_true - 1| of the current estimate."""
        return np.abs(np.exp(self.log_mu) / self.mu_true - 1.0)

    def close(self):
        self.scene.close()


def frames_text(frames):
    f = list(frames)
    if len(f) > 3 and f == list(range(f[0], f[-1] + 1)):
        return f"{f[0]}-{f[-1]}"
    return "/".join(str(x) for x in f)


def face_pairs(tets):
    """Pairs (i, j), i < j, of tets sharing a face."""
    t = np.asarray(tets, np.int64).reshape(-1, 4)
    faces = np.sort(np.concatenate([t[:, [0, 1, 2]], t[:, [0, 1, 3]], t[:, [0, 2, 3]], t[:, [1, 2, 3]]]), axis=1)
    owner = np.tile(np.arange(len(t)), 4)
    key = (faces[:, 0] * (1 << 42)) + (faces[:, 1] * (1 << 21)) + faces[:, 2]
    o = np.argsort(key, kind="stable")
    same = key[o][1:] == key[o][:-1]
    a, b = owner[o][:-1][same], owner[o][1:][same]
    return np.stack([np.minimum(a, b), np.maximum(a, b)], axis=1)


def run(iters=200, n_bodies=64, seed=4, timing_forward=True, log=None, smooth=0.0, loss_frames=None):
    ctx = xf.Context(0)
    batch = workloads.fit_batch(n_bodies=n_bodies, seed=seed)
    t0 = time.perf_counter()
    fit = BatchFit(ctx, batch, loss_frames=loss_frames, smooth=smooth)
    setup_s = time.perf_counter() - t0
    err0 = fit.mu_error()
    fwd_s = None
    if timing_forward:
        fit.forward()
        t0 = time.perf_counter()
        fit.forward()
        fwd_s = time.perf_counter() - t0
    losses, t_grad = [], 0.0
    for i in range(iters):
        t0 = time.perf_counter()
        L = fit.step()
        t_grad += time.perf_counter() - t0
        losses.append(L)
        if log:
            e = fit.mu_error()
            log(f"iter {i + 1}/{iters}: loss {L:.6e}, median rel mu err {np.median(e):.4f}, mean {e.mean():.4f}, "
                f"body step-backs {fit.failures}")
    err = fit.mu_error()
    tets = int(sum(fit.tets_per_body))
    out = dict(
        workload=f"{batch['name']}: {n_bodies} beams, {tets} tets, {fit.n_sub} substeps "
                 f"({batch['frames']} frames x {batch['substeps']}), loss at frames {frames_text(fit.loss_frames)}, "
                 f"window {fit.window}",
        iterations=iters, smooth=smooth, setup_s=setup_s,
        adjoint_s_per_iter=t_grad / max(iters, 1),
        body_gradients_per_s=n_bodies * iters / t_grad if iters else None,
        batch_gradients_per_s=iters / t_grad if iters else None,
        forward_s=fwd_s,
        adjoint_over_forward=(t_grad / iters) / fwd_s if (iters and fwd_s) else None,
        reverse_over_forward=(t_grad / iters) / fwd_s - 1.0 if (iters and fwd_s) else None,
        loss_first=losses[0] if losses else None, loss_last=losses[-1] if losses else None,
        rel_mu_err_initial=dict(median=float(np.median(err0)), mean=float(err0.mean())),
        rel_mu_err_final=dict(median=float(np.median(err)), mean=float(err.mean()),
                              p90=float(np.quantile(err, 0.9))),
        body_step_backs=fit.failures,
    )
    fit.close()
    ctx.close()
    return out


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--iters", type=int, default=200)
    ap.add_argument("--bodies", type=int, default=64)
    ap.add_argument("--smooth", type=float, default=0.0, help="weight of the log-mu smoothness prior")
    ap.add_argument("--loss-frames", default="all",
                    help="frames whose node positions enter the loss: comma list, or 'all' for every frame")
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    frames = None if a.loss_frames == "all" else tuple(int(x) for x in a.loss_frames.split(","))
    res = run(a.iters, a.bodies, log=lambda s: print(s, flush=True), smooth=a.smooth, loss_frames=frames)
    print(json.dumps(res), flush=True)
    if a.out:
        with open(a.out, "w") as f:
            json.dump(res, f, indent=1)
    sys.exit(0)
