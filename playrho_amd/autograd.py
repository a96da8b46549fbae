"""torch binding of the adjoint: the libtorch operators
torch.ops.playrho_amd.simulate, billiards_loss and articulated_loss, C++
torch::autograd::Functions over the library's adjoint
(csrc/torch_adjoint.cpp, built into libplayrho_amd_torch.so next to this
file).

    loss = billiards_loss(v0, friction, restitution, descs, targets, nsteps, every)
    loss.sum().backward()       # v0.grad, friction.grad, restitution.grad

The forward call steps the worlds (with the given initial velocities and
materials) nsteps steps on the device and returns the per-world loss
sum |x_K - target|^2 over dynamic bodies; the gradients come from the
hand-written adjoint, which the same call runs (recomputing each segment of
`every` steps from its checkpoint).

Parameters per world w: v0[w] (bodies, 3) initial velocities (vx, vy, w)
of every body (static bodies' entries are ignored, as SetVelocity ignores
them), friction[w] and restitution[w] (shapes,) per shape; descs: the
worlds' descriptions (World.save_desc()), all with the same body and shape
counts; targets: (worlds, bodies, 2).
"""
import os

import numpy as np
import torch

_LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libplayrho_amd_torch.so")
_loaded = False


def _ops():
    global _loaded
    if not _loaded:
        if not os.path.exists(_LIB):
            raise RuntimeError("libplayrho_amd_torch.so is not built: run __graft_entry__.build()")
        torch.ops.load_library(_LIB)
        _loaded = True
    return torch.ops.playrho_amd


def _pack(descs, v0, targets):
    blob = b"".join(descs)
    offsets = np.cumsum([0] + [len(d) for d in descs]).astype(np.int64)
    t = torch.as_tensor(np.asarray(targets, dtype=np.float32)).reshape(v0.shape[0], v0.shape[1], 2)
    return torch.frombuffer(bytearray(blob), dtype=torch.uint8), torch.from_numpy(offsets), t


def simulate(v0, friction, restitution, descs, nsteps, every, torque=None):
    """The final state (worlds, bodies, 6): x, y, angle, vx, vy, w of every
    body after nsteps steps from initial velocities v0 (worlds, bodies, 3),
    per-shape friction and restitution (worlds, shapes) and, for worlds with
    revolute joints, per-joint maximum motor torques (worlds, joints).  The
    operator torch.ops.playrho_amd.simulate is a torch::autograd::Function:
    any loss of the state back-propagates to v0, friction, restitution and
    torque through the adjoint (its backward steps the worlds again from
    checkpoints every `every` steps and reverses the steps)."""
    blob = b"".join(descs)
    offsets = torch.from_numpy(np.cumsum([0] + [len(d) for d in descs]).astype(np.int64))
    if torque is None:
        torque = torch.zeros((v0.shape[0], 0), dtype=v0.dtype)
    return _ops().simulate(v0, friction, restitution, torque, torch.frombuffer(bytearray(blob), dtype=torch.uint8),
                           offsets, int(nsteps), int(every))


def billiards_loss(v0, friction, restitution, descs, targets, nsteps, every):
    blob, offsets, t = _pack(descs, v0, targets)
    return _ops().billiards_loss(v0, friction, restitution, blob, offsets, t, int(nsteps), int(every))


def articulated_loss(v0, friction, restitution, torque, descs, targets, nsteps, every):
    """As billiards_loss, for worlds with revolute joints: torque (worlds,
    joints) holds every joint's maximum motor torque (each world is built as
    if its joints were created with these torques) and is differentiated
    too."""
    blob, offsets, t = _pack(descs, v0, targets)
    return _ops().articulated_loss(v0, friction, restitution, torque, blob, offsets, t, int(nsteps), int(every))
