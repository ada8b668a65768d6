"""One process of the device-tensor check (run by test_gpu_torch_device.py
in a fresh interpreter, so that PyTorch is imported before the library and
both use PyTorch's HIP runtime): the autograd trajectory with GPU tensors
against the same trajectory with CPU tensors, and the device state
round trip.  Prints one JSON line with the comparisons."""
import json
import os
import sys

import torch  # first: the library binds to the HIP runtime loaded first

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import numpy as np  # noqa: E402

import nbodysim_amd  # noqa: E402
from nbodysim_amd import host  # noqa: E402
from nbodysim_amd.autograd import trajectory  # noqa: E402


def run(device, pos, vel, mass, tgt, precision):
    x0 = torch.tensor(pos, device=device, requires_grad=True)
    v0 = torch.tensor(vel, device=device, requires_grad=True)
    m = torch.tensor(mass, device=device, requires_grad=True)
    t = torch.tensor(tgt, device=device)
    with nbodysim_amd.Simulation(theta=0.5, dt=0.0333, precision=precision) as s:
        xT, vT = trajectory(x0, v0, m, s, nsteps=10, every=5)
        loss = ((xT - t) ** 2).sum(1).mean()
        loss.backward()
    out = [a.detach().cpu().numpy() for a in (xT, vT, x0.grad, v0.grad, m.grad)]
    return out, {"xT": str(xT.device), "grad": str(x0.grad.device)}


def main():
    pos, vel, mass = host.generate(3000, gen_type="sphere", seed=21, r_sphere=100.0,
                                   rotation_magnitude=0.1, rotation_vector=(0.0, 0.0, 1.0))
    c, s_ = np.cos(np.pi / 6), np.sin(np.pi / 6)
    tgt = pos @ np.array([[c, -s_, 0.0], [s_, c, 0.0], [0.0, 0.0, 1.0]]).T
    res = {"cuda": torch.cuda.is_available()}
    maps = open("/proc/self/maps").read()
    res["hip_runtimes"] = sorted({l.split()[-1] for l in maps.splitlines()
                                  if "libamdhip64" in l})
    for prec in ("f64", "f32"):
        gpu, devs = run("cuda", pos, vel, mass, tgt, prec)
        cpu, _ = run("cpu", pos, vel, mass, tgt, prec)
        r = {"devices": devs,
             "xT_equal": bool(np.array_equal(gpu[0], cpu[0])),
             "vT_equal": bool(np.array_equal(gpu[1], cpu[1]))}
        for k, name in ((2, "gx0"), (3, "gv0"), (4, "gm")):
            scale = float(np.abs(cpu[k]).max())
            r[name + "_max_rel"] = float(np.abs(gpu[k] - cpu[k]).max() / scale)
        res[prec] = r
    # the device state round trip, and a host address refused
    with nbodysim_amd.Simulation(theta=0.5, dt=0.0333, precision="f32") as s:
        x = torch.tensor(pos, device="cuda")
        v = torch.tensor(vel, device="cuda")
        m = torch.tensor(mass, device="cuda")
        s.set_state_device(len(mass), x.data_ptr(), v.data_ptr(), m.data_ptr())
        s.step(3)
        p_dev = torch.empty_like(x)
        v_dev = torch.empty_like(v)
        a_dev = torch.empty_like(x)
        s.state_device(p_dev.data_ptr(), v_dev.data_ptr(), a_dev.data_ptr())
        p_h, v_h, a_h = s.state()
        res["state_round_trip_equal"] = bool(
            np.array_equal(p_dev.cpu().numpy(), p_h) and np.array_equal(v_dev.cpu().numpy(), v_h)
            and np.array_equal(a_dev.cpu().numpy(), a_h))
        host_arr = np.zeros((len(mass), 3))
        try:
            s.state_device(host_arr.ctypes.data, 0, 0)
            res["host_address_refused"] = False
        except nbodysim_amd.NbsError as e:
            res["host_address_refused"] = "another HIP runtime" in str(e)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
