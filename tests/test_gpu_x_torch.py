"""GPU checks of the torch.autograd.Function around the sampled TT contraction (xerus_amd.autograd).

Kept in its own module, collected after the torch-free GPU tests. The first test runs in this process, where libxtt
was loaded before torch (the fixture): torch's tensors stay on the CPU and are staged through libxtt buffers. The
second runs tests/torch_device_check.py in a process of its own with torch imported first, so that libxtt binds to
torch's HIP runtime and the Function takes device tensors in place.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from helpers import SEED
from test_gpu_tt import twin_random

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def test_autograd_adjoint(xe, oracle):
    """torch.autograd.Function around the sampled contraction: values vs the oracle's measure, backward vs the
    oracle's ADF gradient (Δ = −∂½Σ(x−b)²/∂G), which test_oracle pins to FP64 finite differences."""
    from xerus_amd.autograd import adf_loss, tt_sample  # loads libxtt before torch
    import torch
    print("autograd adjoint", flush=True)
    d, n, r, P = 6, 7, 5, 4000
    x, xo = twin_random(xe, oracle, [n] * d, [r] * (d - 1))
    truth, trutho = twin_random(xe, oracle, [n] * d, [r] * (d - 1), seed=SEED + 2)
    oracle.seed(SEED + 3)
    mo = oracle.Measurements.random(P, [n] * d)
    mo.measure(trutho)
    pos = torch.tensor(mo.positions(d).astype(np.int32))
    vals = torch.tensor(mo.values(), dtype=torch.float32)
    cores = [torch.tensor(c, dtype=torch.float32, requires_grad=True) for c in xo.cores()]
    sampled = tt_sample(pos, cores)
    mx = oracle.Measurements.from_arrays(mo.positions(d), np.zeros(P))
    mx.measure(xo)
    assert np.abs(sampled.detach().numpy() - mx.values()).max() < 1e-5 * np.abs(mx.values()).max()
    loss = adf_loss(pos, vals, cores)
    loss.backward()
    go, _ = oracle.adf_gradient(xo, mo)
    for k in range(d):
        g = cores[k].grad.numpy()
        assert np.abs(g + go[k]).max() < 3e-5 * np.abs(go[k]).max()


def test_autograd_device_tensors():
    """Device tensors in place on torch's runtime (a process of its own, torch imported first): values and gradients
    on the device against the oracle, central differences at d=3 through the torch path, CPU tensors moved there and
    back."""
    print("autograd on device tensors", flush=True)
    out = subprocess.run([sys.executable, os.path.join(HERE, "torch_device_check.py")], capture_output=True, text=True,
                         timeout=300)
    print(out.stdout, out.stderr[-3000:], flush=True)
    assert out.returncode == 0, out.stderr[-3000:]
    vals = {line.split()[0]: float(line.split()[1]) for line in out.stdout.splitlines() if len(line.split()) == 2}
    assert vals["torch_runtime"] == 1 and vals["values_on_device"] == 1 and vals["grads_on_device"] == 1
    assert vals["values_rel_err"] < 1e-5
    assert vals["grad_rel_err"] < 3e-5
    assert vals["fd_rel_err"] < 1e-3
    assert vals["cpu_roundtrip"] == 1
