"""Thin torch.autograd.Function over xf_scene_adjoint, for gradient checks.

RolloutLoss.apply(log_mu, scene, x0, v0, substeps, dt, loss_at, targets,
window, fp64) sets every body's per-tet mu = exp(log_mu) (bodies
concatenated), resets the bodies to (x0, v0), rolls the scene forward and
returns the trajectory loss as a 0-d float64 tensor; backward returns
grad_output * dL/dlog(mu) from the hand-written adjoint (fp64: of the FP64
element model, xf_scene_adjoint_fp64). All work happens in the HIP library;
torch only carries the values.
"""
import numpy as np
import torch

import xpbdfem as xf

xf.lib()
if len(xf.loaded_hip_runtimes()) > 1:
    raise ImportError("two HIP runtimes are loaded (" + ", ".join(xf.loaded_hip_runtimes()) + "): import torch "
                      "before xpbdfem so that libxpbdfem.so binds to torch's runtime")


class RolloutLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, log_mu, scene, x0, v0, substeps, dt, loss_at, targets, window=20, fp64=False):
        mu = torch.exp(log_mu.detach().to("cpu", torch.float64)).numpy().astype(np.float32)
        off = 0
        for b, (n, t) in enumerate(scene.bodies):
            scene.set_tet_mu(b, mu[off:off + t])
            scene.set_state(b, x0[b], v0[b])
            off += t
        L, g = scene.adjoint(substeps, dt, loss_at, targets, window, fp64=fp64)
        ctx.save_for_backward(torch.from_numpy(g).to(log_mu.device))
        return torch.tensor(L, dtype=torch.float64, device=log_mu.device)

    @staticmethod
    def backward(ctx, grad_output):
        (g,) = ctx.saved_tensors
        return (grad_output * g, None, None, None, None, None, None, None, None, None)
