"""torch.autograd binding of the sampled-entry TT contraction and its hand-written adjoint kernel.

``tt_sample(positions, cores)`` returns x(position_i) for every row of ``positions`` (P×d, int32); its backward is
libxtt's adjoint kernel, Σ_i seed_i·∂x_i/∂G_k for every core at once. The ADF gradient of xerus
(adf.cpp:314-396) is the negative gradient of ½‖tt_sample(P, G) − b‖² with respect to G_k.

torch bundles its own HIP runtime, and libxtt must run on the ROCm runtime it was built for, so the binding keeps
the device side inside libxtt: tensors are CPU float32 (cores shaped (r_{k-1}, n_k, r_k), r_0 = r_d = 1, ranks
<= 16) and are staged through libxtt device buffers. Import xerus_amd (this module does) before torch.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import DeviceArray, call

_lib.lib()  # load libxtt (and its HIP runtime) before torch is imported below

import torch  # noqa: E402


def _meta(shapes):
    d = len(shapes)
    n = (C.c_int * d)(*[int(s[1]) for s in shapes])
    r = (C.c_int * (d + 1))(*([int(shapes[0][0])] + [int(s[2]) for s in shapes]))
    return d, n, r


def _check(positions, cores):
    if positions.dtype != torch.int32 or positions.dim() != 2 or positions.is_cuda:
        raise ValueError("positions must be a CPU int32 tensor of shape (P, d)")
    if positions.shape[1] != len(cores):
        raise ValueError("positions have a different order than the TT")
    for c in cores:
        if c.dtype != torch.float32 or c.is_cuda or c.dim() != 3:
            raise ValueError("cores must be CPU float32 tensors of shape (r_{k-1}, n_k, r_k)")
    # the kernels take every core offset from the rank chain, so it must describe the cores exactly
    if cores[0].shape[0] != 1 or cores[-1].shape[2] != 1:
        raise ValueError("the first core's left rank and the last core's right rank must be 1")
    for k in range(len(cores) - 1):
        if cores[k].shape[2] != cores[k + 1].shape[0]:
            raise ValueError(f"core {k} has right rank {cores[k].shape[2]} but core {k + 1} has left rank {cores[k + 1].shape[0]}")


class TTSample(torch.autograd.Function):
    @staticmethod
    def forward(ctx, positions, *cores):
        _check(positions, cores)
        shapes = [tuple(c.shape) for c in cores]
        d, n, r = _meta(shapes)
        flat = np.concatenate([c.detach().numpy().reshape(-1) for c in cores]).astype(np.float32)
        pos = np.ascontiguousarray(positions.numpy())
        P = int(pos.shape[0])
        dcores, dpos = DeviceArray.from_numpy(flat), DeviceArray.from_numpy(pos)
        dout = DeviceArray((max(P, 1),), np.float32)
        call("xt_tt_sample_f32", d, n, r, dcores.ptr, dpos.ptr, P, dout.ptr, None)
        ctx.save_for_backward(positions)
        ctx.dev = (dcores, dpos)
        ctx.shapes = shapes
        return torch.from_numpy(dout.numpy()[:P].copy())

    @staticmethod
    def backward(ctx, grad_values):
        (positions,) = ctx.saved_tensors
        dcores, dpos = ctx.dev
        shapes = ctx.shapes
        d, n, r = _meta(shapes)
        P = int(positions.shape[0])
        seed = DeviceArray.from_numpy(grad_values.detach().to(torch.float32).contiguous().numpy())
        total = sum(a * b * c for a, b, c in shapes)
        dgrad = DeviceArray((total,), np.float32)
        call("xt_tt_sample_adjoint_f32", d, n, r, dcores.ptr, dpos.ptr, P, seed.ptr, dgrad.ptr, None)
        g = dgrad.numpy()
        out, o = [], 0
        for s in shapes:
            k = s[0] * s[1] * s[2]
            out.append(torch.from_numpy(g[o: o + k].reshape(s).copy()))
            o += k
        return (None, *out)


def tt_sample(positions: torch.Tensor, cores) -> torch.Tensor:
    return TTSample.apply(positions, *cores)


def adf_loss(positions: torch.Tensor, values: torch.Tensor, cores) -> torch.Tensor:
    """½‖P_Ω(x) − b‖², differentiable in every core through the adjoint kernel."""
    return 0.5 * torch.sum((tt_sample(positions, cores) - values) ** 2)
