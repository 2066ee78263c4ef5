"""GPU check of the torch.autograd.Function around the sampled TT contraction (xerus_amd.autograd).

Kept in its own module, collected after the torch-free GPU tests: it is the only GPU test that loads torch
into the same process as libxtt. libxtt is loaded first (the fixture), torch's tensors stay on the CPU and its
own HIP runtime is never initialised; all device work runs inside libxtt.
"""
import numpy as np
import pytest

from helpers import SEED
from test_gpu_tt import twin_random

pytestmark = pytest.mark.gpu


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
