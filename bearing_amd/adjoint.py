"""The adjoint as a torch.autograd.Function: a held trajectory (native.Adjoint) stepped from
an initial state with given masses, frictions and restitutions, differentiable in all four.

The arithmetic runs in the library (bm_adjoint_forward / bm_adjoint_backward: the device's
forward and reverse sweeps); this module only moves the tensors across. Static bodies are
part of the scene, not of the state: their rows of the state's gradient are 0."""
import numpy as np
import torch

from . import native


class HeldTrajectory(torch.autograd.Function):
    """out = the final state (bodies x 13: p, q, v, w) of `adj` stepped from `state0`."""

    @staticmethod
    def forward(ctx, adj, state0, mass, friction, restitution):
        ins = [t.detach().to("cpu", torch.float64).contiguous().numpy() for t in (state0, mass, friction, restitution)]
        out = adj.forward(*ins)
        ctx.adj = adj
        ctx.ins = ins
        ctx.token = adj.forwards
        ctx.device = state0.device
        return torch.from_numpy(out).to(state0.device)

    @staticmethod
    def backward(ctx, grad):
        adj = ctx.adj
        # The reverse pass walks the last forward pass's checkpoints: step forward again if
        # another forward pass ran since this one.
        if adj.forwards != ctx.token:
            adj.forward(*ctx.ins)
        g = adj.backward(grad.detach().to("cpu", torch.float64).contiguous().numpy())
        to = lambda a: torch.from_numpy(a).to(ctx.device)
        return None, to(g["state"]), to(g["mass"]), to(g["friction"]), to(g["restitution"])


class Held:
    """A native.Adjoint with the Function around it: held(state0, mass, friction, restitution)."""

    def __init__(self, adj):
        self.adj = adj

    def recorded(self, device="cpu"):
        rec = self.adj.recorded()
        return {k: torch.from_numpy(np.array(v)).to(device) for k, v in rec.items()}

    def __call__(self, state0, mass, friction, restitution):
        return HeldTrajectory.apply(self.adj, state0, mass, friction, restitution)
