"""torch.autograd binding of the sampled-entry TT contraction and its hand-written adjoint kernel.

``tt_sample(positions, cores)`` returns x(position_i) for every row of ``positions`` (P×d, int32); its backward is
libxtt's adjoint kernel, Σ_i seed_i·∂x_i/∂G_k for every core at once. The ADF gradient of xerus
(adf.cpp:314-396) is the negative gradient of ½‖tt_sample(P, G) − b‖² with respect to G_k.

Tensors: cores float32 shaped (r_{k-1}, n_k, r_k) with r_0 = r_d = 1 and ranks <= 16, positions int32 (P, d).

Two ways the process can be set up, decided by what was loaded first:
- torch first (the usual case: ``import torch`` before xerus_amd, or this module imported before libxtt is loaded):
  libxtt binds to torch's HIP runtime, so device tensors are used in place on torch's current stream, with no host
  round trip; CPU tensors are moved to the current device and the results back.
- libxtt first (``xerus_amd.lib()`` before importing torch): torch then runs on the ROCm runtime without a usable
  device, so tensors must be on the CPU and are staged through libxtt device buffers.

Stacks: the adjoint keeps no left or right stacks in memory between the forward and the backward pass. Each
workgroup recomputes its samples' stacks from the cores in LDS (the limit of stack checkpointing with no stored
checkpoint): at r <= 16 a recomputed stack entry costs r² FMAs from LDS against r floats read back from HBM, and the
forward pass then saves only the positions and the cores.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import DeviceArray, call

import torch  # noqa: E402  (before libxtt is loaded, libxtt binds to torch's runtime; after it, torch has no device)

_lib.lib()


def _device_mode() -> bool:
    return _lib.on_torch_runtime()


def _meta(shapes):
    d = len(shapes)
    n = (C.c_int * d)(*[int(s[1]) for s in shapes])
    r = (C.c_int * (d + 1))(*([int(shapes[0][0])] + [int(s[2]) for s in shapes]))
    return d, n, r


def _check(positions, cores):
    if positions.dtype != torch.int32 or positions.dim() != 2:
        raise ValueError("positions must be an int32 tensor of shape (P, d)")
    if len(cores) == 0 or positions.shape[1] != len(cores):
        raise ValueError("positions have a different order than the TT")
    for c in cores:
        if c.dtype != torch.float32 or c.dim() != 3:
            raise ValueError("cores must be float32 tensors of shape (r_{k-1}, n_k, r_k)")
        if c.device != positions.device:
            raise ValueError("positions and cores must be on the same device")
    if positions.is_cuda and not _device_mode():
        raise ValueError("device tensors need libxtt on torch's runtime: import torch before xerus_amd.autograd and do "
                         "not call xerus_amd.lib() first")
    # the kernels take every core offset from the rank chain, so it must describe the cores exactly
    if cores[0].shape[0] != 1 or cores[-1].shape[2] != 1:
        raise ValueError("the first core's left rank and the last core's right rank must be 1")
    for k in range(len(cores) - 1):
        if cores[k].shape[2] != cores[k + 1].shape[0]:
            raise ValueError(f"core {k} has right rank {cores[k].shape[2]} but core {k + 1} has left rank {cores[k + 1].shape[0]}")


def _ptr(t) -> C.c_void_p:
    return C.c_void_p(t.data_ptr())


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


class TTSample(torch.autograd.Function):
    @staticmethod
    def forward(ctx, positions, *cores):
        _check(positions, cores)
        shapes = [tuple(c.shape) for c in cores]
        d, n, r = _meta(shapes)
        P = int(positions.shape[0])
        ctx.shapes = shapes
        if _device_mode():
            home = positions.device
            dev = positions.device if positions.is_cuda else torch.device("cuda", torch.cuda.current_device())
            pos = positions.to(dev).contiguous()
            flat = torch.cat([c.detach().to(dev).reshape(-1) for c in cores]).contiguous()
            out = torch.empty(max(P, 1), dtype=torch.float32, device=dev)
            call("xt_tt_sample_f32", d, n, r, _ptr(flat), _ptr(pos), P, _ptr(out), _stream())
            ctx.save_for_backward(pos, flat)
            ctx.home = home
            return out[:P].to(home)
        flat = np.concatenate([c.detach().numpy().reshape(-1) for c in cores]).astype(np.float32)
        pos = np.ascontiguousarray(positions.numpy())
        dcores, dpos = DeviceArray.from_numpy(flat), DeviceArray.from_numpy(pos)
        dout = DeviceArray((max(P, 1),), np.float32)
        call("xt_tt_sample_f32", d, n, r, dcores.ptr, dpos.ptr, P, dout.ptr, None)
        ctx.save_for_backward(positions)
        ctx.dev = (dcores, dpos)
        return torch.from_numpy(dout.numpy()[:P].copy())

    @staticmethod
    def backward(ctx, grad_values):
        shapes = ctx.shapes
        d, n, r = _meta(shapes)
        total = sum(a * b * c for a, b, c in shapes)
        if _device_mode():
            pos, flat = ctx.saved_tensors
            P = int(pos.shape[0])
            seed = grad_values.detach().to(device=flat.device, dtype=torch.float32).contiguous()
            grads = torch.empty(total, dtype=torch.float32, device=flat.device)
            call("xt_tt_sample_adjoint_f32", d, n, r, _ptr(flat), _ptr(pos), P, _ptr(seed), _ptr(grads), _stream())
            grads = grads.to(ctx.home)
        else:
            (positions,) = ctx.saved_tensors
            dcores, dpos = ctx.dev
            P = int(positions.shape[0])
            seed = DeviceArray.from_numpy(grad_values.detach().to(torch.float32).contiguous().numpy())
            dgrad = DeviceArray((total,), np.float32)
            call("xt_tt_sample_adjoint_f32", d, n, r, dcores.ptr, dpos.ptr, P, seed.ptr, dgrad.ptr, None)
            grads = torch.from_numpy(dgrad.numpy())
        out, o = [], 0
        for s in shapes:
            k = s[0] * s[1] * s[2]
            out.append(grads[o: o + k].reshape(s))
            o += k
        return (None, *out)


def tt_sample(positions: torch.Tensor, cores) -> torch.Tensor:
    return TTSample.apply(positions, *cores)


def adf_loss(positions: torch.Tensor, values: torch.Tensor, cores) -> torch.Tensor:
    """½‖P_Ω(x) − b‖², differentiable in every core through the adjoint kernel."""
    return 0.5 * torch.sum((tt_sample(positions, cores) - values) ** 2)
