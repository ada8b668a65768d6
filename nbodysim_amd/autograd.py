"""torch.autograd.Function over the hand-written adjoint of the step.

    from nbodysim_amd.autograd import trajectory
    sim = nbodysim_amd.Simulation(theta=0.5, dt=0.0333, precision="f32")
    xT, vT = trajectory(x0, v0, m, sim, nsteps=200, every=20)
    loss = ((xT - target) ** 2).sum(1).mean()
    loss.backward()          # x0.grad, v0.grad, m.grad from the GPU adjoint

The forward runs nsteps frames on the GPU keeping (x, v) checkpoints every
`every` steps; backward recomputes each segment from its checkpoint and
runs the reverse walk (cell decisions of the forward held fixed).  Inputs
and outputs are float64 tensors on any device; the state crosses to the
engine as host arrays (56 B per particle each way).
"""
import numpy as np
import torch


class _Trajectory(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x0, v0, m, sim, nsteps, every):
        dev = x0.device
        sim.set_state(x0.detach().cpu().numpy(), v0.detach().cpu().numpy(),
                      m.detach().cpu().numpy())
        sim.adjoint_forward(nsteps, every)
        xT, vT, _ = sim.state()
        ctx.sim = sim
        ctx.dev = dev
        return (torch.from_numpy(xT).to(dev), torch.from_numpy(vT).to(dev))

    @staticmethod
    def backward(ctx, gxT, gvT):
        gx = None if gxT is None else gxT.detach().cpu().numpy().astype(np.float64)
        gv = None if gvT is None else gvT.detach().cpu().numpy().astype(np.float64)
        gx0, gv0, gm = ctx.sim.adjoint_backward(gx, gv)
        t = lambda a: torch.from_numpy(a).to(ctx.dev)
        return t(gx0), t(gv0), t(gm), None, None, None


def trajectory(x0, v0, m, sim, nsteps, every=20):
    """(x_T, v_T) after nsteps frames, differentiable w.r.t. x0, v0, m."""
    return _Trajectory.apply(x0, v0, m, sim, int(nsteps), int(every))
