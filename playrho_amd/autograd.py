"""torch.autograd binding of the adjoint kernel (prx_adjoint_*).

    loss = BilliardsLoss.apply(v0, friction, restitution, descs, targets, nsteps, every)
    loss.sum().backward()       # v0.grad, friction.grad, restitution.grad

The forward pass steps the worlds (with the given initial velocities and
materials) K steps on the device and returns the per-world loss
sum |x_K - target|^2; the gradients come from the hand-written adjoint, which
the forward call runs at the same time (recomputing from its checkpoints).
The tensors are plumbing: the computation is the library's.

Parameters per world w: v0[w] (bodies, 3) initial velocities (vx, vy, w) of
every body (static bodies' entries are ignored), friction[w] and
restitution[w] (shapes,) per shape.
"""
import numpy as np
import torch

from . import Adjoint, Batch, StepConf, World, desc_with


def _descs_with(descs, v0, friction, restitution):
    out = []
    for w, d in enumerate(descs):
        vel = {b: v0[w, b] for b in range(v0.shape[1])}
        fr = {s: friction[w, s] for s in range(friction.shape[1])}
        re = {s: restitution[w, s] for s in range(restitution.shape[1])}
        out.append(desc_with(d, velocities=vel, friction=fr, restitution=re))
    return out


class BilliardsLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, v0, friction, restitution, descs, targets, nsteps, every):
        v = v0.detach().cpu().numpy().astype(np.float64)
        f = friction.detach().cpu().numpy().astype(np.float64)
        e = restitution.detach().cpu().numpy().astype(np.float64)
        worlds = [World.from_desc(d) for d in _descs_with(descs, v, f, e)]
        batch = Batch(worlds, use_graph=False)
        adj = Adjoint(batch, int(nsteps), int(every), StepConf())
        out = adj.run(np.asarray(targets, dtype=np.float32).reshape(-1, 2))
        nw, nb, ns = v.shape[0], v.shape[1], f.shape[1]
        ctx.save_for_backward(
            torch.from_numpy(out["grad_velocity"].reshape(nw, nb, 3)).to(v0.device),
            torch.from_numpy(out["grad_friction"].reshape(nw, ns)).to(v0.device),
            torch.from_numpy(out["grad_restitution"].reshape(nw, ns)).to(v0.device))
        return torch.from_numpy(out["loss"].astype(np.float32)).to(v0.device)

    @staticmethod
    def backward(ctx, gloss):
        gv, gf, ge = ctx.saved_tensors
        g = gloss.to(gv.dtype)
        return g[:, None, None] * gv, g[:, None] * gf, g[:, None] * ge, None, None, None, None
