"""Run by tests/test_gpu_x_torch.py in a process of its own, where torch is imported before xerus_amd.autograd, so
libxtt binds to torch's HIP runtime and the autograd Function takes device tensors in place. Prints one line per
check ("name value") for the test to assert on."""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import numpy as np  # noqa: E402
import torch  # noqa: E402  (first: its HIP runtime is the one libxtt binds to)

from xerus_amd import _lib  # noqa: E402
from xerus_amd.autograd import adf_loss, tt_sample  # noqa: E402
from helpers import SEED  # noqa: E402
from oracle import oracle as O  # noqa: E402

print("torch_runtime", int(_lib.on_torch_runtime()), flush=True)
dev = torch.device("cuda", 0)

# values and gradients on device tensors against the oracle (FP64 CPU restatement of adf.cpp)
d, n, r, P = 6, 7, 5, 4000
O.seed(SEED)
xo = O.TT.random([n] * d, [r] * (d - 1))
trutho = O.TT.random([n] * d, [r] * (d - 1))
mo = O.Measurements.random(P, [n] * d)
mo.measure(trutho)
pos = torch.tensor(mo.positions(d).astype(np.int32), device=dev)
vals = torch.tensor(mo.values(), dtype=torch.float32, device=dev)
cores = [torch.tensor(c, dtype=torch.float32, device=dev, requires_grad=True) for c in xo.cores()]
sampled = tt_sample(pos, cores)
print("values_on_device", int(sampled.is_cuda), flush=True)
mx = O.Measurements.from_arrays(mo.positions(d), np.zeros(P))
mx.measure(xo)
ref = mx.values()
print("values_rel_err", float(np.abs(sampled.detach().cpu().numpy() - ref).max() / np.abs(ref).max()), flush=True)
loss = adf_loss(pos, vals, cores)
loss.backward()
go, _ = O.adf_gradient(xo, mo)
err = max(float(np.abs(cores[k].grad.cpu().numpy() + go[k]).max() / np.abs(go[k]).max()) for k in range(d))
print("grads_on_device", int(all(c.grad.is_cuda for c in cores)), flush=True)
print("grad_rel_err", err, flush=True)

# central differences at d=3 through the torch path: the loss is quadratic in every single core entry, so the
# central difference with any step is its exact derivative (up to FP32 rounding of the loss values)
d, n, r, P = 3, 3, 2, 15
O.seed(SEED + 7)
xo = O.TT.random([n] * d, [r] * (d - 1))
trutho = O.TT.random([n] * d, [r] * (d - 1))
mo = O.Measurements.random(P, [n] * d)
mo.measure(trutho)
pos = torch.tensor(mo.positions(d).astype(np.int32), device=dev)
vals = torch.tensor(mo.values(), dtype=torch.float32, device=dev)
cores = [torch.tensor(c, dtype=torch.float32, device=dev, requires_grad=True) for c in xo.cores()]
grads = torch.autograd.grad(adf_loss(pos, vals, cores), cores)
h, worst = 0.25, 0.0
with torch.no_grad():
    for k in range(d):
        for idx in np.ndindex(tuple(cores[k].shape)):
            cp = [c.detach().clone() for c in cores]
            cm = [c.detach().clone() for c in cores]
            cp[k][idx] += h
            cm[k][idx] -= h
            fd = (adf_loss(pos, vals, cp).item() - adf_loss(pos, vals, cm).item()) / (2 * h)
            worst = max(worst, abs(fd - grads[k][idx].item()) / max(1.0, abs(fd)))
print("fd_rel_err", worst, flush=True)

# CPU tensors in the same process: moved to the device and back
cpu_cores = [c.detach().cpu().requires_grad_(True) for c in cores]
v = tt_sample(pos.cpu(), cpu_cores)
print("cpu_roundtrip", int(not v.is_cuda), flush=True)
