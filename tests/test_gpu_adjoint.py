"""The hand-written adjoint of the step against FP64 central finite
differences.  At theta = 0 the walk is the direct sum (a smooth map); at
theta = 0.5 the finite differences run through nbs_frozen_forward, which
holds the tree topology and cell decisions of the recorded trajectory (the
map the adjoint differentiates).  Bar: 1e-3 relative per checked
coordinate (plus a small absolute floor for near-zero gradients)."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

ROT30 = np.array([[np.cos(np.pi / 6), -np.sin(np.pi / 6), 0.0],
                  [np.sin(np.pi / 6), np.cos(np.pi / 6), 0.0],
                  [0.0, 0.0, 1.0]])


def sim(**kw):
    import nbodysim_amd
    return nbodysim_amd.Simulation(**kw)


def loss(xT, tgt):
    return float(np.mean(np.sum((xT - tgt) ** 2, axis=1)))


def loss_diff(xp, xm, tgt):
    return float(np.mean(np.sum((xp - xm) * (xp + xm - 2 * tgt), axis=1)))


def check_fd(adj, fd, floor_rel=1e-6):
    adj, fd = np.asarray(adj), np.asarray(fd)
    floor = floor_rel * np.abs(fd).max()
    err = np.abs(adj - fd) / (np.abs(fd) + floor)
    assert err.max() < 1e-3, (err.max(), adj, fd)
    return err.max()


def run_case(n, steps, every, theta, frozen, seed=3, ncoord=12, precision="f64"):
    pos, vel, mass = O.generate(n, "sphere", seed=seed)
    tgt = pos @ ROT30.T
    with sim(theta=theta, dt=0.0333, precision=precision) as s:
        s.set_state(pos, vel, mass)
        s.adjoint_forward(steps, every, record_all=frozen)
        xT, vT, _ = s.state()
        gx0, gv0, gm = s.adjoint_backward(2.0 * (xT - tgt) / n, None)
        if precision == "f32":
            # finite differences of the FP64 frozen map on the trajectory
            # the FP32-force forward recorded
            s.configure(precision="f64")
            xT = s.frozen_forward(pos, vel, mass)[0]

        def fwd(p, v, m):
            if frozen:
                return s.frozen_forward(p, v, m)[0]
            with sim(theta=theta, dt=0.0333, precision="f64") as t:
                t.set_state(p, v, m)
                t.step(steps)
                return t.state()[0]

        # the frozen replay of the unperturbed inputs is the trajectory
        if frozen and precision == "f64":
            np.testing.assert_array_equal(fwd(pos, vel, mass), xT)
        rng = np.random.default_rng(seed)
        adj, fd = [], []
        for k in range(ncoord):
            kind = ("x", "v", "m")[k % 3]
            i = int(rng.integers(n))
            c = int(rng.integers(3))
            p, v, m = pos.copy(), vel.copy(), mass.copy()
            if kind == "x":
                h = 1e-5
                p[i, c] += h
                xp = fwd(p, v, m)
                p[i, c] -= 2 * h
                xm = fwd(p, v, m)
                adj.append(gx0[i, c])
            elif kind == "v":
                h = 1e-6
                v[i, c] += h
                xp = fwd(p, v, m)
                v[i, c] -= 2 * h
                xm = fwd(p, v, m)
                adj.append(gv0[i, c])
            else:
                h = 1e-6 * mass[i]
                m[i] += h
                xp = fwd(p, v, m)
                m[i] -= 2 * h
                xm = fwd(p, v, m)
                adj.append(gm[i])
            # L(+) - L(-) formed from the trajectories themselves,
            # mean((x+ - x-) . (x+ + x- - 2 t)), so the difference is not
            # lost to the cancellation of two large loss values
            fd.append(loss_diff(xp, xm, tgt) / (2 * h))
    # FP32 forces carry ~1e-7 relative rounding each, which moves every
    # component by ~1e-6 of the gradient's scale (measured: a component of
    # 6e-6 beside a largest one of 3.6e-2 is off by 1.7e-8); the FP32 case
    # is held to 1e-3 of max(|fd_i|, 1e-3 max_j |fd_j|)
    return check_fd(adj, fd, 1e-3 if precision == "f32" else 1e-6), (gx0, gv0, gm)


def test_adjoint_matches_fd_direct_sum_dynamics():
    run_case(n=384, steps=12, every=5, theta=0.0, frozen=False)


def test_adjoint_matches_fd_frozen_theta05():
    # Horizon 10 steps: the FD of the frozen map converges there (FD at h,
    # h/4, h/16 agree to ~1e-5).  Over ~20+ steps close encounters (the
    # reference has no softening) make the map chaotic and FP64 central
    # differences stop converging, while the adjoint stays exact.
    run_case(n=2048, steps=10, every=4, theta=0.5, frozen=True)


def test_adjoint_fp32_forces_matches_fd_of_frozen_map():
    """The adjoint at config 5's precision (FP32 walk forces, FP64 state and
    adjoint): its gradient against FP64 central differences of the frozen
    map on the trajectory the FP32 forward recorded: within 1e-3 of
    max(|fd_i|, 1e-3 max_j |fd_j|) per coordinate (measured: 5e-5 relative
    or better on every component above 1e-3 of the largest; the two maps
    differ by the FP32 rounding of the forces along 10 steps)."""
    run_case(n=2048, steps=10, every=4, theta=0.5, frozen=True, precision="f32", seed=5)


def test_checkpoint_interval_does_not_change_the_gradient():
    pos, vel, mass = O.generate(3000, "sphere", seed=9)
    out = []
    for every in (1, 7, 20):
        with sim(theta=0.5, dt=0.0333, precision="f32") as s:
            s.set_state(pos, vel, mass)
            s.adjoint_forward(20, every)
            xT, _, _ = s.state()
            out.append(s.adjoint_backward(xT, None))
    # Segments are recomputed bit-exactly from their checkpoints; the only
    # run-to-run freedom is the order of the FP64 atomic adds that gather
    # node adjoints from many waves (~1e-14 relative, measured); a recompute
    # taking a different tree or cell decision moves the gradient by O(1).
    for k in (1, 2):
        for a, b in zip(out[0], out[k]):
            err = np.abs(a - b).max() / np.abs(b).max()
            assert err < 1e-11, (k, err)


def test_tree_gradient_close_to_direct_sum_gradient():
    # theta = 0.5 adjoint vs the theta = 0 (direct sum) adjoint at small N
    pos, vel, mass = O.generate(1024, "sphere", seed=11)
    tgt = pos @ ROT30.T
    g = []
    for theta in (0.5, 0.0):
        with sim(theta=theta, dt=0.0333, precision="f64") as s:
            s.set_state(pos, vel, mass)
            s.adjoint_forward(10, 5)
            xT, _, _ = s.state()
            g.append(s.adjoint_backward(2.0 * (xT - tgt) / len(pos), None))
    for a, b in zip(g[0], g[1]):
        rel = np.linalg.norm(a - b) / np.linalg.norm(b)
        assert rel < 5e-2, rel


def test_torch_autograd_function():
    torch = pytest.importorskip("torch")
    import nbodysim_amd
    from nbodysim_amd.autograd import trajectory
    pos, vel, mass = O.generate(1000, "sphere", seed=5)
    tgt = torch.from_numpy(pos @ ROT30.T)
    x0 = torch.from_numpy(pos).requires_grad_(True)
    v0 = torch.from_numpy(vel).requires_grad_(True)
    m = torch.from_numpy(mass).requires_grad_(True)
    with nbodysim_amd.Simulation(theta=0.5, dt=0.0333, precision="f32") as s:
        xT, vT = trajectory(x0, v0, m, s, nsteps=10, every=5)
        L = ((xT - tgt) ** 2).sum(1).mean()
        L.backward()
    with sim(theta=0.5, dt=0.0333, precision="f32") as s:
        s.set_state(pos, vel, mass)
        s.adjoint_forward(10, 5)
        xT2, _, _ = s.state()
        gx0, gv0, gm = s.adjoint_backward(2.0 * (xT2 - tgt.numpy()) / len(pos), None)
    # same computation through two entry points; the FP64 atomic adds that
    # gather node adjoints may sum in another order (~1e-14 of the scale)
    for got, want in ((x0.grad.numpy(), gx0), (v0.grad.numpy(), gv0), (m.grad.numpy(), gm)):
        np.testing.assert_allclose(got, want, rtol=1e-10, atol=1e-12 * np.abs(want).max())
