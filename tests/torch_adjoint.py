"""Thin torch.autograd.Function over the C-ABI adjoint, for tests only.

Tensors stay on the CPU (torch is loaded only for autograd bookkeeping; the forward and
the adjoint run on the GPU through mp_step / mp_adjoint).
"""
import numpy as np
import torch


class FrozenFieldProblem:
    """K MoveParticles steps of fixed initial particles under a per-element field E."""

    def __init__(self, mp, mover, pos0, u0, cell0, drift_dir, qm, dt, spwt, steps, den_target, checkpoint_every=10):
        self.mp, self.mover = mp, mover
        self.pos0, self.u0, self.cell0 = pos0, u0, cell0
        self.dir = np.asarray(drift_dir, np.float64)
        self.qm, self.dt, self.spwt, self.steps = qm, dt, spwt, steps
        self.tgt = np.asarray(den_target, np.float64)
        self.ckpt = checkpoint_every

    def forward(self, E, drift, mode=1):
        """Loss, final cells (by pid) and final density of the forward run (deterministic scatter)."""
        mv = self.mover
        mv.upload_field(E)
        mv.set_particles(self.pos0, self.u0 + drift * self.dir, self.cell0)
        for _ in range(self.steps):
            mv.step(self.qm, self.dt, self.spwt, mode)
        den = mv.density()
        _, _, cell, pid = mv.get_particles()
        return float(((den - self.tgt) ** 2).sum()), (pid, cell), den

    def adjoint(self, E, drift):
        self.mover.upload_field(E)
        return self.mover.adjoint(self.pos0, self.u0 + drift * self.dir, self.cell0, self.qm, self.dt, self.spwt,
                                  self.steps, self.tgt, self.dir, self.ckpt)


class AdjointLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, E, drift, problem):
        loss, gE, gd = problem.adjoint(E.detach().numpy(), float(drift))
        ctx.save_for_backward(torch.from_numpy(gE), torch.tensor(gd, dtype=torch.float64))
        return torch.tensor(loss, dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad_out):
        gE, gd = ctx.saved_tensors
        return grad_out * gE, grad_out * gd, None
