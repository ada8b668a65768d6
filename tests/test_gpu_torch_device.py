"""PyTorch GPU tensors through the library without host copies
(nbs_set_state_device, nbs_get_state_device, nbs_adjoint_backward_device,
nbodysim_amd.autograd).  PyTorch ships its own HIP runtime and the library
binds to the runtime its process loaded first, so the check runs in a
fresh interpreter that imports torch first (tests/torch_device_case.py);
this process has already loaded the library on its own runtime.

Bars: the GPU-tensor trajectory equals the CPU-tensor one bit for bit
(same engine, same steps), gradients within 1e-10 of the largest
component (the adjoint gathers node adjoints with FP64 atomic adds, whose
order may differ between runs); one HIP runtime in the process; the
device state equals nbs_get_state's; a host address is refused."""
import json
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def test_torch_device_tensors_through_the_library():
    pytest.importorskip("torch")
    print("torch device case: fresh interpreter", flush=True)
    p = subprocess.run([sys.executable, os.path.join(HERE, "torch_device_case.py")],
                       capture_output=True, text=True, timeout=240)
    print(p.stdout[-3000:], p.stderr[-3000:], flush=True)
    assert p.returncode == 0
    res = json.loads(p.stdout.strip().splitlines()[-1])
    assert res["cuda"]
    assert len(res["hip_runtimes"]) == 1, res["hip_runtimes"]
    for prec in ("f64", "f32"):
        r = res[prec]
        assert r["devices"]["xT"].startswith("cuda") and r["devices"]["grad"].startswith("cuda")
        assert r["xT_equal"] and r["vT_equal"], (prec, r)
        for name in ("gx0", "gv0", "gm"):
            assert r[name + "_max_rel"] <= 1e-10, (prec, name, r)
    assert res["state_round_trip_equal"]
    assert res["host_address_refused"]
