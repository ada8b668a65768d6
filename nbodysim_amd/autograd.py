"""torch.autograd.Function over the hand-written adjoint of the step.

    import torch                        # before nbodysim_amd (see below)
    from nbodysim_amd.autograd import trajectory
    sim = nbodysim_amd.Simulation(theta=0.5, dt=0.0333, precision="f32")
    xT, vT = trajectory(x0, v0, m, sim, nsteps=200, every=20)
    loss = ((xT - target) ** 2).sum(1).mean()
    loss.backward()          # x0.grad, v0.grad, m.grad from the GPU adjoint

The forward runs nsteps frames on the GPU keeping (x, v) checkpoints every
`every` steps; backward recomputes each segment from its checkpoint and
runs the reverse walk (cell decisions of the forward held fixed).  Inputs
and outputs are float64 tensors.

GPU tensors stay on the GPU: the state, the seeds and the gradients cross
as device arrays (nbs_set_state_device, nbs_get_state_device,
nbs_adjoint_backward_device), with no host copy.  That needs the library
and PyTorch on one HIP runtime: PyTorch ships its own, and the library
binds to whichever the process loaded first, so PyTorch must be imported
before the library is first used (this module imports it; a library
loaded earlier on its own runtime refuses the tensors' addresses, and the
error says so).  CPU tensors cross as host arrays (56 B per particle each
way).
"""
import numpy as np
import torch

from . import NbsError

_RUNTIME_HINT = ("the library is bound to another HIP runtime than PyTorch's: import torch "
                 "(or nbodysim_amd.autograd) before the first use of nbodysim_amd in this "
                 "process")


def _dev(t, shape):
    """float64, contiguous, on t's device, of the given shape."""
    t = t.detach().to(torch.float64).contiguous()
    if tuple(t.shape) != shape:
        raise ValueError("expected shape %s, got %s" % (shape, tuple(t.shape)))
    return t


def _device_call(fn, *args):
    try:
        fn(*args)
    except NbsError as e:
        if "another HIP runtime" in str(e):
            raise RuntimeError("%s (%s)" % (_RUNTIME_HINT, e)) from e
        raise


class _Trajectory(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x0, v0, m, sim, nsteps, every):
        n = m.shape[0]
        dev = x0.device
        ctx.sim, ctx.dev, ctx.n = sim, dev, n
        if dev.type == "cuda":
            x, v, mm = _dev(x0, (n, 3)), _dev(v0, (n, 3)), _dev(m, (n,))
            torch.cuda.current_stream(dev).synchronize()
            _device_call(sim.set_state_device, n, x.data_ptr(), v.data_ptr(), mm.data_ptr())
            sim.adjoint_forward(nsteps, every)
            xT = torch.empty((n, 3), dtype=torch.float64, device=dev)
            vT = torch.empty((n, 3), dtype=torch.float64, device=dev)
            _device_call(sim.state_device, xT.data_ptr(), vT.data_ptr(), 0)
            return xT, vT
        sim.set_state(x0.detach().cpu().numpy(), v0.detach().cpu().numpy(),
                      m.detach().cpu().numpy())
        sim.adjoint_forward(nsteps, every)
        xT, vT, _ = sim.state()
        return torch.from_numpy(xT).to(dev), torch.from_numpy(vT).to(dev)

    @staticmethod
    def backward(ctx, gxT, gvT):
        sim, dev, n = ctx.sim, ctx.dev, ctx.n
        if dev.type == "cuda":
            gx = None if gxT is None else _dev(gxT, (n, 3))
            gv = None if gvT is None else _dev(gvT, (n, 3))
            gx0 = torch.empty((n, 3), dtype=torch.float64, device=dev)
            gv0 = torch.empty((n, 3), dtype=torch.float64, device=dev)
            gm = torch.empty((n,), dtype=torch.float64, device=dev)
            torch.cuda.current_stream(dev).synchronize()
            _device_call(sim.adjoint_backward_device, 0 if gx is None else gx.data_ptr(),
                         0 if gv is None else gv.data_ptr(), gx0.data_ptr(), gv0.data_ptr(),
                         gm.data_ptr())
            return gx0, gv0, gm, None, None, None
        gx = None if gxT is None else gxT.detach().cpu().numpy().astype(np.float64)
        gv = None if gvT is None else gvT.detach().cpu().numpy().astype(np.float64)
        gx0, gv0, gm = sim.adjoint_backward(gx, gv)
        t = lambda a: torch.from_numpy(a).to(dev)
        return t(gx0), t(gv0), t(gm), None, None, None


def trajectory(x0, v0, m, sim, nsteps, every=20):
    """(x_T, v_T) after nsteps frames, differentiable w.r.t. x0, v0, m."""
    return _Trajectory.apply(x0, v0, m, sim, int(nsteps), int(every))
