"""torch.autograd harness around the GPU discrete adjoint (harm_adj_*).

    J = MagneticEnergyAfterSteps.apply(p0, harm, nsteps, dt0, checkpoint_every, dts)
    J.backward()          # p0.grad = dJ/dp0 from the GPU reverse sweep

p0 is the primitive state [N1+6][N2+6][8] (float64, any device); the forward
runs `nsteps` steps of `harm` (a Harm context) on the GPU from p0 after
bound_prim, with the time steps `dts` imposed (or recorded when None, the
first step using dt0), and returns the magnetic energy sum 0.5 b^2 gdet dV.
StepsFunction returns the final state instead and takes any adjoint of it.
"""
import numpy as np
import torch

from . import P, Adjoint


def _start(p0, harm, dt0):
    harm.set_array(P, np.ascontiguousarray(p0.detach().to("cpu", torch.float64).numpy()))
    harm.bound_prim(P)
    harm.set_time(0.0, dt0, 1e30)


class MagneticEnergyAfterSteps(torch.autograd.Function):
    @staticmethod
    def forward(ctx, p0, harm, nsteps, dt0=1e-3, checkpoint_every=16, dts=None):
        _start(p0, harm, dt0 if dts is None else float(dts[0]))
        adj = Adjoint(harm, nsteps, checkpoint_every)
        adj.forward(dts)
        J = adj.seed_magnetic_energy()
        ctx.adj, ctx.device = adj, p0.device
        return torch.tensor(J, dtype=torch.float64, device=p0.device)

    @staticmethod
    def backward(ctx, gJ):
        lam = torch.from_numpy(ctx.adj.backward()).to(ctx.device)
        ctx.adj.close()
        return gJ * lam, None, None, None, None, None


class StepsFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, p0, harm, nsteps, dt0=1e-3, checkpoint_every=16, dts=None):
        _start(p0, harm, dt0 if dts is None else float(dts[0]))
        adj = Adjoint(harm, nsteps, checkpoint_every)
        adj.forward(dts)
        ctx.adj, ctx.device = adj, p0.device
        return torch.from_numpy(harm.get_array(P)).to(p0.device)

    @staticmethod
    def backward(ctx, g):
        ctx.adj.seed(np.ascontiguousarray(g.detach().to("cpu", torch.float64).numpy()))
        lam = torch.from_numpy(ctx.adj.backward()).to(ctx.device)
        ctx.adj.close()
        return lam, None, None, None, None, None
