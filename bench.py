"""Benchmark of the Barnes-Hut step on one MI355X.

Workload: gen_type sphere, N = 1,048,576 uniform in a ball of r = 100,
masses uniform in [5e10, 5e11], rigid rotation 0.1 about z, theta = 0.5,
dt = 0.0333 (root from the bounding box + 2, the reference's gen_root).
One step = root, octree build, mass/COM, walk and kick-drift update on the
GPU with the state resident in HBM.

Prints ONE JSON line:
  value     particle-steps/s of the FP32-force / FP64-state path
  fp64      the all-FP64 path on the same workload, and the speed ratio
  accuracy  rms relative acceleration errors at the initial state on 4096
            seeded samples: e_prec (FP32 tree vs FP64 tree), e_trunc (FP64
            tree vs FP64 direct sum), e_total (FP32 tree vs direct sum)
  roofline  the walk kernel: 20 FLOP per body-node interaction (interactions
            counted by the kernel) over its HIP-event time, against the FP32
            vector peak
  cpu_baseline  the oracle (CPU restatement of the reference's step) on the
            host cores: full tree build + mass/COM of all N particles and the
            walk of a seeded sample, extrapolated to one step of N
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_PEAK_TFLOPS = 157.3   # MI355X vector FP32 (AMD spec)
FP64_PEAK_TFLOPS = 78.6    # MI355X vector FP64
HBM_PEAK_GBS = 8000.0
FLOP_PER_INTERACTION = 20
# interaction counting of the timed walks (--counting)
COUNTING = "total"
# pieces per walk traversal (--walk-split; 0: the library's choice)
WALK_SPLIT = 0
TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "walk_traffic_sphere1M.md")


def measured_traffic(key):
    """HBM bytes per walk launch for `key` (precision and shape, e.g.
    "f32 64,8,1") from the committed rocprofv3 counter summary, whose lines
    read `traffic_bytes_per_launch[<key>]: <bytes>`; None if absent."""
    try:
        with open(TRAFFIC_SUMMARY) as f:
            for line in f:
                line = line.strip().strip("`")
                if line.startswith("traffic_bytes_per_launch[%s]:" % key):
                    return int(line.split(":")[1])
    except OSError:
        pass
    return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def time_path(sim, pos, vel, mass, precision, steps, warmup, shape, graph, order=False,
              theta=None, offsets=None):
    """Timed region: `steps` graph-replayed steps (nbs_step_graph; eager
    launches with graph=False) after `warmup` steps.  Before it (untimed)
    the walk's launch shape is tuned on the initial state (nbs_autotune,
    which restores the state) unless `shape` fixes it.  Then the same
    window again: the state at the window's start is set back and the same
    `steps` steps run with eager launches and HIP events around each phase
    on the context's stream (kernel times for the phases and the roofline,
    and the interactions those walks counted).  The trajectory depends only
    on positions and caller order, so both runs take the same steps; the
    end states are compared bit for bit.  `offsets` batches the state into
    systems (nbs_set_systems) for both runs.  Both runs count interactions
    as COUNTING says (by default the running total only, as the frame loop
    steps: per-particle counts are a diagnostic the reference's step does
    not have)."""
    kw = {"precision": precision}
    if theta is not None:
        kw["theta"] = theta
    sim.configure(**kw)
    sim.set_counting(COUNTING)
    sim.set_walk_split(WALK_SPLIT)
    sim.set_state(pos, vel, mass)
    if offsets is not None:
        sim.set_systems(offsets)
    sim.set_walk_order(order)
    if shape is None:
        shape, _ = sim.autotune(reps=3)
    else:
        sim.set_walk_shape(*shape)
    if warmup:
        sim.step(warmup, graph=graph)
    start_pos, start_vel, _ = sim.state()
    sim.synchronize()
    t0 = time.perf_counter()
    sim.step(steps, graph=graph)
    sim.synchronize()
    t1 = time.perf_counter()
    info = sim.tree_info()
    sim_time = float(sim.time()[0])
    end_pos, end_vel, _ = sim.state()
    # the same window, eager and profiled
    sim.set_state(start_pos, start_vel, mass)
    if offsets is not None:
        sim.set_systems(offsets)
    sim.interactions_total(reset=True)
    sim.profile(True)
    sim.profile_read(reset=True)
    sim.synchronize()
    t2 = time.perf_counter()
    sim.step(steps)
    sim.synchronize()
    t3 = time.perf_counter()
    prof = sim.profile_read(reset=True)
    sim.profile(False)
    inter = sim.interactions_total(reset=True)
    replay_pos, replay_vel, _ = sim.state()
    return {"ms_per_step": (t1 - t0) * 1e3 / steps, "prof": prof, "interactions": inter,
            "prof_steps": steps, "eager_ms_per_step": (t3 - t2) * 1e3 / steps,
            "same_window_bit_identical": bool(np.array_equal(replay_pos, end_pos) and
                                              np.array_equal(replay_vel, end_vel)),
            "tree": info, "walk_shape": list(shape), "end_state": (end_pos, end_vel),
            "sim_time": sim_time}


def accuracy_gated(sim, pos, vel, mass, ids, ref, rms, args, shape, graph):
    """The largest theta whose FP32-force and FP64 walks both stay within
    the rms error gate against the FP64 direct sum at the initial state
    (a descending scan, then bisection to 0.0025), and both paths' speed
    there: the reference's algorithm with the MAC tightened until it meets
    the accuracy the metric is quoted at."""
    def err(theta):
        out = []
        for prec in ("f32", "f64"):
            sim.configure(theta=theta, precision=prec)
            sim.set_state(pos, vel, mass)
            out.append(rms(sim.forces()[ids], ref))
        return out
    scan = []
    hi, lo = None, None  # hi fails the gate, lo meets it
    for theta in (0.3, 0.2, 0.15, 0.12, 0.1, 0.08, 0.06, 0.04):
        e = err(theta)
        scan.append({"theta": theta, "e_total_f32": e[0], "e_total_f64": e[1]})
        if max(e) <= args.gate:
            lo = (theta, e)
            break
        hi = theta
    if lo is None:
        return {"gate": args.gate, "scan": scan, "theta": None}
    while hi is not None and hi - lo[0] > 0.0025:
        mid = round((hi + lo[0]) / 2, 5)
        e = err(mid)
        scan.append({"theta": mid, "e_total_f32": e[0], "e_total_f64": e[1]})
        if max(e) <= args.gate:
            lo = (mid, e)
        else:
            hi = mid
    theta = lo[0]
    res = {"gate": args.gate, "theta": theta, "e_total_f32": lo[1][0], "e_total_f64": lo[1][1],
           "steps": args.gated_steps, "scan": scan}
    n = len(mass)
    for prec in ("f32", "f64"):
        r = time_path(sim, pos, vel, mass, prec, args.gated_steps, args.warmup, shape, graph,
                      theta=theta)
        walk_ms = r["prof"]["walk_ms"] / max(r["prof"]["walk_launches"], 1)
        res[prec] = {"value": n * 1e3 / r["ms_per_step"], "ms_per_step": r["ms_per_step"],
                     "walk_ms": walk_ms,
                     "interactions_per_particle": r["interactions"] / r["prof_steps"] / n,
                     "walk_tflops": FLOP_PER_INTERACTION * r["interactions"] / r["prof_steps"] /
                     (walk_ms * 1e-3) / 1e12}
    res["speedup_fp32_over_fp64"] = res["f32"]["value"] / res["f64"]["value"]
    sim.configure(theta=0.5)
    return res


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(pos, mass, theta, sample, threads, prec="f64"):
    """The oracle's step (gen_root + octree + mass/COM, then the BFS walk)
    on the host cores, at the particle state the GPU's timed region ended
    in, so both sides do the same interactions.  prec "f32" is the
    reference's FLOAT build (every value and operation in float).  With
    sample < N the walk covers a seeded subset and is extrapolated to N."""
    import oracle as O
    n = len(pos)
    t0 = time.perf_counter()
    tree = O.Tree(pos, mass, prec=prec)  # gen_root + add_particle + mass + COM
    t1 = time.perf_counter()
    if sample >= n:
        _, cnt = tree.forces(theta=theta, threads=threads)
        walked = n
    else:
        rng = np.random.default_rng(0xBEEF)
        idx = np.sort(rng.choice(n, sample, replace=False))
        _, cnt = tree.forces_subset(idx, theta=theta, threads=threads)
        walked = sample
    t2 = time.perf_counter()
    t_step = (t1 - t0) + (t2 - t1) * n / walked
    return {
        "value": n / t_step, "unit": "particle-steps/s", "cores": threads, "kind": "port",
        "precision": prec, "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
        "sample": "oracle (C++ restatement of the reference step, %s) at the state the "
                  "timed GPU steps ended in: tree build + mass/COM of all %d particles "
                  "(%.2f s, serial like gen_root) and the BFS walk of %d particles on %d "
                  "threads (%.2f s)%s" % (prec.upper(), n, t1 - t0, walked, threads, t2 - t1,
                                          "" if walked == n else ", extrapolated to N"),
        "interactions_per_particle": float(cnt.mean()),
        "seconds": t2 - t0,
    }


def copy_calibration(mib, reps):
    """A plain device-to-device copy of known size, the yardstick for the
    FETCH_SIZE / WRITE_SIZE counters (counter KiB per copied KiB)."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.arange(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration": {"bytes_read": n * 4, "bytes_written": n * 4,
                                           "launches": reps}}), flush=True)


# ----------------------------------------------------------------------------
# The other workloads of the step, each printing its own JSON line
# (`--config 3|4|5`); the default run above is the headline.

SPHERE = dict(r_sphere=100.0, rotation_magnitude=0.1, rotation_vector=(0.0, 0.0, 1.0),
              min_mass=5e10, max_mass=5e11)


def config_shell(args):
    """Shell of N = 4,194,304 at r = 100 with 1 % radial jitter, rotation
    0.1 about z, FP32 forces, adaptive dt (min 0.00333, max position /
    velocity change 3), theta 0.3 / 0.5 / 0.7, from the same initial
    state each: 2 warm-up steps, `--steps3` steps timed as a HIP graph
    (eager launches with --eager), then the same window replayed eagerly
    with HIP events per phase (time_path)."""
    import nbodysim_amd
    from nbodysim_amd import host
    n = args.n3
    pos, vel, mass = host.generate(n, gen_type="shell", seed=0xDEADBEEF, shell_jitter=0.01,
                                   **SPHERE)
    res = []
    for theta in (0.3, 0.5, 0.7):
        sim = nbodysim_amd.Simulation(theta=theta, dt=0.0333, precision="f32", adaptive=True,
                                      min_adaptive_dt=0.00333, max_pos_change=3.0,
                                      max_vel_change=3.0)
        r = time_path(sim, pos, vel, mass, "f32", args.steps3, 2, (64, 8, 1),
                      not args.eager)
        prof, steps = r["prof"], r["prof_steps"]
        walk_ms = prof["walk_ms"] / max(prof["walk_launches"], 1)
        res.append({"theta": theta, "value": n * 1e3 / r["ms_per_step"],
                    "ms_per_step": r["ms_per_step"],
                    "eager_ms_per_step": r["eager_ms_per_step"],
                    "simulated_time": r["sim_time"], "steps": steps, "warmup": 2,
                    "interactions_per_particle": r["interactions"] / (n * steps),
                    "walk_ms": walk_ms, "build_ms": prof["build_ms"] / steps,
                    "update_ms": prof["update_ms"] / steps,
                    "same_window_bit_identical": r["same_window_bit_identical"],
                    "walk_tflops": FLOP_PER_INTERACTION * r["interactions"] / steps /
                                   (walk_ms * 1e-3) / 1e12,
                    "tree_end": r["tree"]})
        log("shell theta", theta, res[-1])
        sim.close()
    print(json.dumps({
        "metric": "particle-steps/s, shell N=%d, FP32 forces, adaptive dt, theta sweep" % n,
        "unit": "particle-steps/s", "higher_is_better": True, "n_gpus": 1, "dtype": "f32",
        "data": "synthetic (seeded gen_type shell, 1% radial jitter)",
        "config": {"workload": "gen_type shell N=%d r=100 jitter 1%% rotation 0.1 about z, "
                               "masses U[5e10,5e11], adaptive dt (dt 0.0333, min 0.00333, "
                               "max_pos_change 3, max_vel_change 3), %d timed steps per theta "
                               "after 2 warm-up; phases from the same window replayed eagerly"
                               % (n, args.steps3),
                   "launch": "eager" if args.eager else "graph"},
        "results": res}), flush=True)


def batched_systems(count, seed=0xDEADBEEF):
    """`count` systems: N_s = round(exp(U(ln 1024, ln 65536))), gen_type
    uniform over cube / sphere / shell (numpy default_rng(seed) draws N_s
    and the type), system s generated with seed + s.  Cubes: side 512 at
    rest; spheres and shells: r = 100, rotation 0.1 about z."""
    from nbodysim_amd import host
    rng = np.random.default_rng(seed)
    sizes = np.rint(np.exp(rng.uniform(np.log(1024), np.log(65536), count))).astype(int)
    kinds = rng.integers(0, 3, count)
    parts = []
    for s_, (n_s, k) in enumerate(zip(sizes, kinds)):
        g = ("cube", "sphere", "shell")[k]
        kw = dict(size=512.0, min_mass=5e10, max_mass=5e11) if g == "cube" else SPHERE
        parts.append(host.generate(int(n_s), gen_type=g, seed=seed + s_, **kw))
    return parts


def config_batched(args):
    """256 systems stepped as one batched context (systems keyed into one
    sort, one walk launch per step, HIP-graph replay, autotuned walk shape)
    against one context and one launch sequence per system."""
    import nbodysim_amd
    t0 = time.perf_counter()
    parts = batched_systems(args.systems)
    ic_host_ms = (time.perf_counter() - t0) * 1e3
    pos = np.concatenate([p[0] for p in parts])
    vel = np.concatenate([p[1] for p in parts])
    mass = np.concatenate([p[2] for p in parts])
    off = np.cumsum([0] + [len(p[2]) for p in parts])
    n = int(off[-1])
    # the same systems generated on the device (nbs_generate_state)
    rng = np.random.default_rng(0xDEADBEEF)
    sizes = np.rint(np.exp(rng.uniform(np.log(1024), np.log(65536), args.systems))).astype(int)
    kinds = rng.integers(0, 3, args.systems)
    systems = [(int(n_s), ("cube", "sphere", "shell")[k], 0xDEADBEEF + s_)
               for s_, (n_s, k) in enumerate(zip(sizes, kinds))]
    gen = nbodysim_amd.Simulation(theta=0.5, dt=0.0333, precision="f32")
    gen.generate_state(systems=systems[:4], size=512.0, **SPHERE)  # first call: set-up
    gen.synchronize()
    t0 = time.perf_counter()
    gen.generate_state(systems=systems, size=512.0, **SPHERE)
    gen.synchronize()
    ic_dev_ms = (time.perf_counter() - t0) * 1e3
    ic_masses_equal = bool(np.array_equal(gen.masses(), mass))
    gen.close()
    log("ICs host %.1f ms, device %.1f ms" % (ic_host_ms, ic_dev_ms))
    sim = nbodysim_amd.Simulation(theta=0.5, dt=0.0333, precision="f32")
    r = time_path(sim, pos, vel, mass, "f32", args.steps4, 2, None, not args.eager,
                  offsets=off)
    sim.close()
    shape, ms_b, prof = r["walk_shape"], r["ms_per_step"], r["prof"]
    phases = {"build": prof["build_ms"] / args.steps4, "walk": prof["walk_ms"] / args.steps4,
              "update": prof["update_ms"] / args.steps4,
              "eager_step": r["eager_ms_per_step"],
              "same_window_bit_identical": r["same_window_bit_identical"]}
    log("batched", shape, ms_b, phases)
    # one context per system, each system's steps launched on its own
    sims = []
    for p in parts:
        t = nbodysim_amd.Simulation(theta=0.5, dt=0.0333, precision="f32")
        t.set_counting(COUNTING)
        t.set_state(*p)
        t.set_walk_shape(*shape)
        t.step(2)
        sims.append(t)
    for t in sims:
        t.synchronize()
    steps_s = max(2, args.steps4 // 10)
    t0 = time.perf_counter()
    for _ in range(steps_s):
        for t in sims:
            t.step(1)
    for t in sims:
        t.synchronize()
    ms_s = (time.perf_counter() - t0) * 1e3 / steps_s
    for t in sims:
        t.close()
    log("per-system", ms_s)
    print(json.dumps({
        "metric": "particle-steps/s, %d batched systems (total N=%d)" % (len(parts), n),
        "value": n * 1e3 / ms_b, "unit": "particle-steps/s", "higher_is_better": True,
        "n_gpus": 1, "dtype": "f32", "ms_per_step": ms_b, "steps": args.steps4,
        "data": "synthetic (seeded per-system cube/sphere/shell, the reference's generator)",
        "config": {"workload": "%d systems, N_s log-uniform in [1024, 65536], theta 0.5, "
                               "dt 0.0333, FP32 forces" % len(parts),
                   "walk_shape": list(shape),
                   "launch": ("eager" if args.eager else "graph") + ", one batched context"},
        "phases_ms_per_step": phases,
        "interactions_per_particle": r["interactions"] / (n * args.steps4),
        "per_system_launches": {"value": n * 1e3 / ms_s, "ms_per_step": ms_s,
                                "steps": steps_s},
        "speedup_batched": ms_s / ms_b,
        "initial_conditions": {"host_ms": ic_host_ms, "device_ms": ic_dev_ms,
                               "device_includes": "upload of per-system seeds/offsets, "
                                                  "generation, context set-up",
                               "masses_bit_identical": ic_masses_equal}}), flush=True)


def config_adjoint(args):
    """Reverse mode on the N = 65536 sphere: loss = mean |x_T - R_z(30) x_0|^2
    over 200 steps with checkpoints every 20; gradients w.r.t. x_0, v_0, m.
    Finite-difference checks (FP64, frozen tree topology and cell
    decisions) on 64 seeded coordinates at the horizon where central
    differences converge, plus the convergence study at 200 steps; and the
    tree gradient against the direct-sum (theta = 0) gradient at N = 2048."""
    import nbodysim_amd
    from nbodysim_amd import host
    c30, s30 = np.cos(np.pi / 6), np.sin(np.pi / 6)
    rot = np.array([[c30, -s30, 0.0], [s30, c30, 0.0], [0.0, 0.0, 1.0]])
    n = args.n5
    pos, vel, mass = host.generate(n, gen_type="sphere", seed=0xDEADBEEF, **SPHERE)
    tgt = pos @ rot.T
    out = {"metric": "adjoint of %d steps, sphere N=%d, checkpoints every %d" %
                     (args.steps5, n, args.every5),
           "unit": "ms", "higher_is_better": False, "n_gpus": 1, "dtype": "f64",
           "data": "synthetic (seeded gen_type sphere)"}
    # timing: FP32 forces, FP64 state and adjoint
    with nbodysim_amd.Simulation(theta=0.5, dt=0.0333, precision="f32") as sim:
        sim.set_walk_split(WALK_SPLIT)
        sim.set_state(pos, vel, mass)
        t0 = time.perf_counter()
        sim.adjoint_forward(args.steps5, args.every5)
        xT, _, _ = sim.state()
        t1 = time.perf_counter()
        gx, gv, gm = sim.adjoint_backward(2.0 * (xT - tgt) / n, None)
        t2 = time.perf_counter()
    out.update({"forward_ms": (t1 - t0) * 1e3, "backward_ms": (t2 - t1) * 1e3,
                "loss": float(np.mean(np.sum((xT - tgt) ** 2, axis=1))),
                "grad_norms": {"x0": float(np.linalg.norm(gx)), "v0": float(np.linalg.norm(gv)),
                               "m": float(np.linalg.norm(gm))}})
    log("adjoint timing", out)
    if args.adjoint_timing_only:
        out["value"] = out["forward_ms"] + out["backward_ms"]
        print(json.dumps(out), flush=True)
        return

    def fd_study(steps, ncoord, ndir):
        """FP64 central differences of the frozen map (tree topology and
        cell decisions of the recorded trajectory held).  Close encounters
        (no softening) amplify FP64 roundoff into a floor of ~1e-8 on the
        loss difference, so a single coordinate's difference converges only
        where its gradient lifts the loss change above that floor: each of
        `ncoord` seeded coordinates is checked where its differences at h
        and h/4 agree to 1e-3, and the rest are counted.  Directional
        derivatives along `ndir` seeded unit directions over all 7N inputs
        and along the gradient itself carry the whole gradient's signal and
        converge (checked at h, h/2, h/4)."""
        with nbodysim_amd.Simulation(theta=0.5, dt=0.0333, precision="f64") as sim:
            sim.set_state(pos, vel, mass)
            sim.adjoint_forward(steps, args.every5, record_all=True)
            xT, _, _ = sim.state()
            gx, gv, gm = sim.adjoint_backward(2.0 * (xT - tgt) / n, None)

            def dloss(dp, dv, dm, h):
                xp = sim.frozen_forward(pos + h * dp, vel + h * dv, mass * (1 + h * dm))[0]
                xm = sim.frozen_forward(pos - h * dp, vel - h * dv, mass * (1 - h * dm))[0]
                return float(np.mean(np.sum((xp - xm) * (xp + xm - 2 * tgt), axis=1))) / (2 * h)

            rng = np.random.default_rng(0xDEADBEEF)
            coords, skipped = [], 0
            for k in range(ncoord):
                kind = ("x", "v", "m")[k % 3]
                i, c = int(rng.integers(n)), int(rng.integers(3))
                dp, dv, dm = np.zeros_like(pos), np.zeros_like(vel), np.zeros_like(mass)
                if kind == "x":
                    dp[i, c] = 1.0
                    adj = gx[i, c]
                elif kind == "v":
                    dv[i, c] = 1.0
                    adj = gv[i, c]
                else:
                    dm[i] = 1.0
                    adj = gm[i] * mass[i]  # relative mass perturbation
                f1, f2 = dloss(dp, dv, dm, 1e-5), dloss(dp, dv, dm, 2.5e-6)
                if abs(f1 - f2) <= 1e-3 * max(abs(f1), abs(f2)):
                    coords.append({"kind": kind, "i": i, "c": c, "adjoint": float(adj), "fd": f2,
                                   "rel_err": abs(adj - f2) / abs(f2)})
                else:
                    skipped += 1
            dirs = []
            for k in range(ndir + 1):
                if k < ndir:
                    dp = rng.standard_normal(pos.shape)
                    dv = rng.standard_normal(vel.shape)
                    dm = rng.standard_normal(mass.shape)
                    name = "random"
                else:  # along the gradient (masses relative)
                    dp, dv, dm = gx.copy(), gv.copy(), gm * mass
                    name = "gradient"
                nrm = np.sqrt(np.sum(dp ** 2) + np.sum(dv ** 2) + np.sum(dm ** 2))
                dp, dv, dm = dp / nrm, dv / nrm, dm / nrm
                adj = float(np.sum(gx * dp) + np.sum(gv * dv) + np.sum(gm * mass * dm))
                fds = [dloss(dp, dv, dm, h) for h in (1e-5, 5e-6, 2.5e-6)]
                dirs.append({"direction": name, "adjoint": adj, "fd": fds,
                             "rel_err": abs(adj - fds[-1]) / abs(fds[-1]),
                             "fd_spread": (max(fds) - min(fds)) / abs(fds[-1])})
            return coords, skipped, dirs

    coords, skipped, dirs = fd_study(args.fd_steps5, 64, 8)
    out["fd_check"] = {
        "steps": args.fd_steps5, "coords": 64, "coords_converged": len(coords),
        "coords_not_converged": skipped,
        "max_rel_err_converged": float(max(c["rel_err"] for c in coords)) if coords else None,
        "median_rel_err_converged": float(np.median([c["rel_err"] for c in coords])) if coords else None,
        "directional": dirs,
        "max_rel_err_directional": float(max(d["rel_err"] for d in dirs))}
    log("fd check", out["fd_check"])
    # at the full horizon the map is chaotic: FD at three h for one
    # direction shows the non-convergence
    _, _, dirs200 = fd_study(args.steps5, 0, 1)
    out["fd_at_horizon"] = {"steps": args.steps5, "directional": dirs200}
    log("fd at horizon", out["fd_at_horizon"])
    # tree gradient vs direct-sum gradient at reduced N
    p2, v2, m2 = host.generate(2048, gen_type="sphere", seed=0xDEADBEEF, **SPHERE)
    t2g = p2 @ rot.T
    g = []
    for theta in (0.5, 0.0):
        with nbodysim_amd.Simulation(theta=theta, dt=0.0333, precision="f64") as sim:
            sim.set_state(p2, v2, m2)
            sim.adjoint_forward(10, 5)
            x2, _, _ = sim.state()
            g.append(sim.adjoint_backward(2.0 * (x2 - t2g) / len(p2), None))
    out["direct_sum_check"] = {"n": 2048, "steps": 10, "rel_diff": {
        k: float(np.linalg.norm(a - b) / np.linalg.norm(b))
        for k, a, b in zip(("x0", "v0", "m"), g[0], g[1])}}
    out["value"] = out["forward_ms"] + out["backward_ms"]
    print(json.dumps(out), flush=True)


def config_frames(args):
    """Frame output at N = 1,048,576 (sphere, FP32 forces), the reference's
    frame loop (src/main.cpp:1420-1492): per frame one step, then the state
    downloaded and written as a binary frame (80 B per particle), the
    previous frame's file removed once the new one is in place.  Variants:
    steps only; download only, synchronous (pageable get_state) or through
    the asynchronous pinned snapshot; download + write in the stepping
    thread; synchronous download + background writer (dump_threaded);
    asynchronous snapshot + background writer (what nbs_run_frames does)."""
    import shutil
    import tempfile
    import threading
    import nbodysim_amd
    from nbodysim_amd import host
    n, frames = args.n, args.frames
    pos, vel, mass = host.generate(n, gen_type="sphere", seed=0xDEADBEEF, **SPHERE)
    d = tempfile.mkdtemp(prefix="nbs_frames_", dir="/tmp")
    sim = nbodysim_amd.Simulation(theta=0.5, dt=0.0333, precision="f32")
    sim.set_counting(COUNTING)
    sim.set_state(pos, vel, mass)
    sim.set_walk_shape(64, 8, 1)
    sim.step(2)

    def write(f, p, v, a):
        host.write_binary(host.frame_filename(d, f, True), p, v, a, mass)
        prev = host.frame_filename(d, f - 1, True)
        if f > 0 and os.path.exists(prev):
            os.remove(prev)

    def run(mode):
        writer, pending = None, None
        t0 = time.perf_counter()
        for f in range(frames):
            sim.step(1)
            if mode == "steps":
                continue
            if mode in ("download_sync", "sync_write"):
                p, v, a = sim.state()
                if mode == "sync_write":
                    write(f, p, v, a)
            elif mode == "threaded_writer":
                if writer:
                    writer.join()
                p, v, a = sim.state()
                writer = threading.Thread(target=write, args=(f, p, v, a))
                writer.start()
            else:  # download_async, async_snapshot_writer
                if pending is not None:
                    p, v, a = sim.snapshot_end(copy=False)
                    if writer:
                        writer.join()
                        writer = None
                    if mode == "async_snapshot_writer":
                        writer = threading.Thread(target=write, args=(pending, p, v, a))
                        writer.start()
                sim.snapshot_begin()
                pending = f
        if pending is not None:
            p, v, a = sim.snapshot_end(copy=False)
            if writer:
                writer.join()
                writer = None
            if mode == "async_snapshot_writer":
                write(pending, p, v, a)
        if writer:
            writer.join()
        sim.synchronize()
        return (time.perf_counter() - t0) * 1e3 / frames

    res = {}
    try:
        for mode in ("steps", "download_sync", "download_async", "sync_write",
                     "threaded_writer", "async_snapshot_writer"):
            res[mode] = run(mode)
            log("frames", mode, res[mode])
        p, v, a = sim.state()
        t0 = time.perf_counter()
        for k in range(3):
            write(frames + k, p, v, a)
        write_ms = (time.perf_counter() - t0) * 1e3 / 3
    finally:
        sim.close()
        shutil.rmtree(d, ignore_errors=True)
    print(json.dumps({
        "metric": "ms per frame (one step + a binary frame of N=%d particles)" % n,
        "unit": "ms", "higher_is_better": False, "n_gpus": 1, "dtype": "f32",
        "data": "synthetic (seeded gen_type sphere, the reference's generator)",
        "config": {"workload": "gen_type sphere N=%d r=100 rotation 0.1 about z, theta 0.5, "
                               "dt 0.0333, FP32 forces, eager steps, %d frames per variant, "
                               "frames written to /tmp" % (n, frames)},
        "ms_per_frame": res, "write_one_frame_ms": write_ms,
        "frame_bytes": 80 * n}), flush=True)


def config_collide(args):
    """Collision merging (config.collide, src/main.cpp:1442-1483) at
    N = 1,048,576 (sphere, theta 0.5, range 0.01): eager steps with and
    without merging, FP32 and FP64 forces, per-phase times from HIP events
    (the pair pass, merge, update and compaction count as the update
    phase).  The sphere's close pairs merge during the first steps."""
    import nbodysim_amd
    from nbodysim_amd import host
    n, steps = args.n, args.collide_steps
    pos, vel, mass = host.generate(n, gen_type="sphere", seed=0xDEADBEEF, **SPHERE)
    res = {}
    for prec in ("f32", "f64"):
        for coll in (False, True):
            with nbodysim_amd.Simulation(theta=0.5, dt=0.0333, precision=prec) as sim:
                sim.set_counting(COUNTING)
                sim.set_state(pos, vel, mass)
                sim.set_walk_shape(64, 8, 1)
                if coll:
                    sim.set_collisions(args.collide_range)
                sim.step(2)
                sim.synchronize()
                sim.profile(True)
                sim.profile_read(reset=True)
                t0 = time.perf_counter()
                sim.step(steps)
                sim.synchronize()
                ms = (time.perf_counter() - t0) * 1e3 / steps
                prof = sim.profile_read(reset=True)
                key = "%s_%s" % (prec, "collide" if coll else "plain")
                res[key] = {"ms_per_step": ms, "n_end": sim.n,
                            "build_ms": prof["build_ms"] / steps,
                            "walk_ms": prof["walk_ms"] / steps,
                            "update_ms": prof["update_ms"] / steps}
                if coll:
                    pairs, removed = sim.collision_stats()
                    res[key].update(pairs=pairs, absorbed=removed)
                log("collide", key, res[key])
    print(json.dumps({
        "metric": "ms per step with collision merging", "unit": "ms",
        "higher_is_better": False, "n_gpus": 1, "dtype": "f32/f64",
        "data": "synthetic (seeded gen_type sphere, the reference's generator)",
        "config": {"workload": "gen_type sphere N=%d r=100 rotation 0.1 about z, theta 0.5, "
                               "dt 0.0333, collision_range %g, eager steps, %d timed after 2"
                               % (n, args.collide_range, steps)},
        "results": res}), flush=True)


def config_image(args):
    """The frame image (write_image, src/main.cpp:556-678) of N = 1,048,576
    sphere particles at the reference's defaults (1920 x 1080, iso, bw,
    brightness 255): device render plus download of the RGB rows, against
    the oracle's restatement of the reference's CPU loop on one core, and
    the bytes compared."""
    import nbodysim_amd
    from nbodysim_amd import host
    import oracle as O
    n, reps = args.n, args.image_reps
    pos, vel, mass = host.generate(n, gen_type="sphere", seed=0xDEADBEEF, **SPHERE)
    res = {}
    with nbodysim_amd.Simulation(theta=0.5, dt=0.0333, precision="f32") as sim:
        sim.set_state(pos, vel, mass)
        for kw in (dict(), dict(color="heat", adaptive=True, nonlinear=True)):
            key = "bw" if not kw else "heat_adaptive_nonlinear"
            img = sim.render_image(**kw)
            t0 = time.perf_counter()
            for _ in range(reps):
                img = sim.render_image(**kw)
            dev_ms = (time.perf_counter() - t0) * 1e3 / reps
            t0 = time.perf_counter()
            ref = O.render_image(pos, **kw)
            cpu_ms = (time.perf_counter() - t0) * 1e3
            res[key] = {"device_ms": dev_ms, "cpu_oracle_ms": cpu_ms,
                        "bytes_identical": bool(np.array_equal(img, ref)),
                        "lit_pixels": int(img.any(axis=2).sum())}
            log("image", key, res[key])
    print(json.dumps({
        "metric": "ms per frame image (render + download)", "unit": "ms",
        "higher_is_better": False, "n_gpus": 1, "dtype": "f64",
        "data": "synthetic (seeded gen_type sphere, the reference's generator)",
        "config": {"workload": "gen_type sphere N=%d r=100, image 1920x1080 iso, "
                               "brightness 255, %d renders timed" % (n, reps)},
        "cpu_baseline": {"kind": "port", "cores": 1,
                         "sample": "the oracle's write_image restatement, whole image"},
        "results": res}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--fp64-steps", type=int, default=None,
                    help="timed steps of the all-FP64 path (default: --steps)")
    ap.add_argument("--gate", type=float, default=1e-4,
                    help="rms relative error bound of the accuracy-gated line")
    ap.add_argument("--gated-steps", type=int, default=10,
                    help="timed steps of each path at the accuracy-gated theta")
    ap.add_argument("--no-gated", action="store_true")
    ap.add_argument("--samples", type=int, default=4096)
    ap.add_argument("--cpu-sample", type=int, default=0,
                    help="particles the CPU baseline walks (0 = all)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-fp32", type=int, default=65536,
                    help="particles the FP32 (FLOAT build) CPU baseline walks (0 = none)")
    ap.add_argument("--walk-shape", default="64,8,1", metavar="BLOCK,GROUPS,PACK",
                    help="walk launch shape (default: the one nbs_autotune picks on this "
                         "workload, so the committed counter profiles apply)")
    ap.add_argument("--autotune", action="store_true",
                    help="let nbs_autotune pick the walk shape before the timed steps")
    ap.add_argument("--walk-order", action="store_true",
                    help="longest-first dispatch of the walk's workgroups (nbs_set_walk_order)")
    ap.add_argument("--counting", default="total", choices=("total", "particles", "none"),
                    help="interaction counting of the timed walks (nbs_set_counting); the "
                         "interaction figures need total or particles")
    ap.add_argument("--walk-split", type=int, default=0, choices=(0, 1, 2, 4, 8, 16, 32),
                    help="pieces per walk traversal (nbs_set_walk_split; 0: by wavefront count)")
    ap.add_argument("--eager", action="store_true",
                    help="timed steps launched one kernel at a time, not by graph replay")
    ap.add_argument("--copy-calib", type=int, default=0, metavar="MIB",
                    help="only run the counter yardstick: copy MIB MiB 5 times")
    ap.add_argument("--config", type=int, default=0, choices=(0, 3, 4, 5),
                    help="run one of the other workloads instead: 3 shell theta sweep, "
                         "4 batched systems, 5 adjoint")
    ap.add_argument("--frames-study", action="store_true",
                    help="frame output variants (download, snapshot, writer thread)")
    ap.add_argument("--frames", type=int, default=20)
    ap.add_argument("--collide-study", action="store_true",
                    help="steps with collision merging against plain steps")
    ap.add_argument("--collide-steps", type=int, default=10)
    ap.add_argument("--collide-range", type=float, default=0.01)
    ap.add_argument("--image-study", action="store_true",
                    help="frame image rendering on the device against the CPU oracle")
    ap.add_argument("--image-reps", type=int, default=10)
    ap.add_argument("--n3", type=int, default=1 << 22)
    ap.add_argument("--steps3", type=int, default=50)
    ap.add_argument("--systems", type=int, default=256)
    ap.add_argument("--steps4", type=int, default=200)
    ap.add_argument("--n5", type=int, default=65536)
    ap.add_argument("--steps5", type=int, default=200)
    ap.add_argument("--every5", type=int, default=20)
    ap.add_argument("--fd-steps5", type=int, default=10)
    ap.add_argument("--adjoint-timing-only", action="store_true",
                    help="--config 5: the forward and backward timing only (for profiling)")
    args = ap.parse_args()
    global COUNTING, WALK_SPLIT
    COUNTING = args.counting
    WALK_SPLIT = args.walk_split
    if args.fp64_steps is None:
        args.fp64_steps = args.steps
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.copy_calib:
        copy_calibration(args.copy_calib, 5)
        return
    if args.config:
        {3: config_shell, 4: config_batched, 5: config_adjoint}[args.config](args)
        return
    if args.frames_study:
        config_frames(args)
        return
    if args.collide_study:
        config_collide(args)
        return
    if args.image_study:
        config_image(args)
        return

    import nbodysim_amd
    from nbodysim_amd import host

    n, theta, dt = args.n, 0.5, 0.0333
    pos, vel, mass = host.generate(n, gen_type="sphere", seed=0xDEADBEEF, r_sphere=100.0,
                                   rotation_magnitude=0.1, rotation_vector=(0.0, 0.0, 1.0),
                                   min_mass=5e10, max_mass=5e11)
    log("generated %d particles" % n)
    sim = nbodysim_amd.Simulation(theta=theta, dt=dt, precision="f32")

    # ---- accuracy at the initial state
    sim.set_state(pos, vel, mass)
    a32 = sim.forces()
    sim.configure(precision="f64")
    a64 = sim.forces()
    ids = np.random.default_rng(0xDEADBEEF).choice(n, args.samples, replace=False)
    ref = sim.direct(ids)

    def rms(a, b):
        return float(np.sqrt(np.mean((np.linalg.norm(a - b, axis=1) /
                                      np.linalg.norm(b, axis=1)) ** 2)))
    acc = {"e_prec": rms(a32[ids], a64[ids]), "e_trunc": rms(a64[ids], ref),
           "e_total": rms(a32[ids], ref), "samples": args.samples,
           "where": "initial state, FP64 direct sum on the samples as reference"}
    log("accuracy", acc)

    # ---- timed paths
    shape = None if args.autotune else tuple(int(v) for v in args.walk_shape.split(","))
    graph = not args.eager
    r32 = time_path(sim, pos, vel, mass, "f32", args.steps, args.warmup, shape, graph,
                    args.walk_order)
    log("f32", {k: v for k, v in r32.items() if k != "end_state"})
    r64 = time_path(sim, pos, vel, mass, "f64", args.fp64_steps, args.warmup, shape, graph,
                    args.walk_order)
    log("f64", {k: v for k, v in r64.items() if k != "end_state"})
    gated = None
    if not args.no_gated:
        gated = accuracy_gated(sim, pos, vel, mass, ids, ref, rms, args, shape, graph)
        log("gated", gated)
    sim.close()

    value = n * 1e3 / r32["ms_per_step"]
    value64 = n * 1e3 / r64["ms_per_step"]
    walk_ms = r32["prof"]["walk_ms"] / max(r32["prof"]["walk_launches"], 1)
    inter_per_launch = r32["interactions"] / r32["prof_steps"]
    achieved = FLOP_PER_INTERACTION * inter_per_launch / (walk_ms * 1e-3) / 1e12
    walk64_ms = r64["prof"]["walk_ms"] / max(r64["prof"]["walk_launches"], 1)
    achieved64 = FLOP_PER_INTERACTION * (r64["interactions"] / r64["prof_steps"]) / \
        (walk64_ms * 1e-3) / 1e12
    build_ms = r32["prof"]["build_ms"] / r32["prof_steps"]
    nodes = r32["tree"]["n_nodes"]
    build_bytes = 56 * n + 32 * nodes
    out = {
        "metric": "particle-steps/s @N=1M theta=0.5, FP32 vs FP64 GPU at rms accel err <=1e-4",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r32["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded gen_type sphere, the reference's generator)",
        "config": {"workload": "gen_type sphere N=%d r=100 rotation 0.1 about z, masses "
                               "U[5e10,5e11], theta=0.5, dt=0.0333, bbox+2 root; FP32 walk "
                               "forces, FP64 state/keys/COM" % n,
                   "n": n, "theta": theta, "dt": dt, "parallelism": "single GPU",
                   "walk_shape": r32["walk_shape"], "launch": "graph" if graph else "eager",
                   "walk_order": "longest-first" if args.walk_order else "plain",
                   "counting": COUNTING},
        "interactions_per_s": inter_per_launch / (r32["ms_per_step"] * 1e-3),
        "interactions_per_particle": inter_per_launch / n,
        "fp64": {"value": value64, "ms_per_step": r64["ms_per_step"],
                 "steps": args.fp64_steps, "walk_ms": walk64_ms,
                 "walk_shape": r64["walk_shape"],
                 "walk_tflops": achieved64, "walk_frac_fp64_peak": achieved64 / FP64_PEAK_TFLOPS,
                 "speedup_fp32_over_fp64": value / value64},
        "accuracy": acc,
        "phases_ms_per_step": {"build": build_ms, "walk": walk_ms,
                               "update": r32["prof"]["update_ms"] / r32["prof_steps"],
                               "eager_step": r32["eager_ms_per_step"],
                               "graph_step": r32["ms_per_step"],
                               "window": "the %d timed steps, replayed from the same start "
                                         "state with eager launches and HIP events per phase "
                                         "(end state bit-identical: %s)"
                                         % (args.steps, r32["same_window_bit_identical"])},
        "tree_build": {"ms": build_ms, "algorithmic_bytes": build_bytes,
                       "gbps": build_bytes / (build_ms * 1e-3) / 1e9,
                       "frac_hbm": build_bytes / (build_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                       "nodes": nodes, "max_depth": r32["tree"]["max_depth"],
                       "buckets": r32["tree"]["n_buckets"]},
        "roofline": {"bound": "valu", "kernel": "walk_kernel<f32>",
                     "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP32_PEAK_TFLOPS,
                     "traffic": measured_traffic("f32 %s" % ",".join(map(str, r32["walk_shape"]))),
                     "flop_per_interaction": FLOP_PER_INTERACTION,
                     "avg_launch_ms": walk_ms},
    }
    if gated is not None:
        out["accuracy_gated"] = gated
    if not args.no_cpu:
        # the host cores this process may use, at most 16: one GPU's share of
        # the box's cores
        threads = max(1, min(16, len(os.sched_getaffinity(0))))
        end_pos = r32["end_state"][0]
        out["cpu_baseline"] = cpu_baseline(end_pos, mass, theta, args.cpu_sample or n, threads)
        log("cpu", out["cpu_baseline"])
        if args.cpu_fp32:
            fp32 = cpu_baseline(end_pos, mass, theta, args.cpu_fp32, threads, prec="f32")
            out["cpu_baseline"]["fp32"] = fp32
            log("cpu fp32", fp32)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
