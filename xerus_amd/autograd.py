"""torch.autograd binding of the sampled-entry TT contraction and its hand-written adjoint kernel.

``tt_sample(positions, cores)`` returns x(position_i) for every row of ``positions`` (P×d, int32/uint32 on the
GPU); its backward is libxtt's adjoint kernel, Σ_i seed_i·∂x_i/∂G_k for every core at once. The ADF
gradient of xerus (adf.cpp:314-396) is the negative gradient of ½‖tt_sample(P, G) − b‖² with respect to G_k.
Cores are FP32 CUDA tensors shaped (r_{k-1}, n_k, r_k) with r_0 = r_d = 1, ranks <= 16.
"""
from __future__ import annotations

import ctypes as C

import torch

from ._lib import call


def _meta(cores):
    d = len(cores)
    n = [int(c.shape[1]) for c in cores]
    r = [int(cores[0].shape[0])] + [int(c.shape[2]) for c in cores]
    return d, (C.c_int * d)(*n), (C.c_int * (d + 1))(*r)


class TTSample(torch.autograd.Function):
    @staticmethod
    def forward(ctx, positions, *cores):
        if positions.dtype not in (torch.int32, torch.uint32) or positions.dim() != 2 or not positions.is_cuda:
            raise ValueError("positions must be a CUDA int32/uint32 tensor of shape (P, d)")
        if positions.shape[1] != len(cores):
            raise ValueError("positions have a different order than the TT")
        for c in cores:
            if c.dtype != torch.float32 or not c.is_cuda or c.dim() != 3:
                raise ValueError("cores must be CUDA float32 tensors of shape (r_{k-1}, n_k, r_k)")
        d, n, r = _meta(cores)
        flat = torch.cat([c.detach().reshape(-1) for c in cores]).contiguous()
        pos = positions.contiguous()
        P = int(pos.shape[0])
        out = torch.empty(P, dtype=torch.float32, device=flat.device)
        stream = torch.cuda.current_stream(flat.device).cuda_stream
        call("xt_tt_sample_f32", d, n, r, flat.data_ptr(), pos.data_ptr(), P, out.data_ptr(), stream)
        ctx.save_for_backward(pos, flat)
        ctx.shapes = [tuple(c.shape) for c in cores]
        return out

    @staticmethod
    def backward(ctx, grad_values):
        pos, flat = ctx.saved_tensors
        shapes = ctx.shapes
        d = len(shapes)
        n = (C.c_int * d)(*[s[1] for s in shapes])
        r = (C.c_int * (d + 1))(*([shapes[0][0]] + [s[2] for s in shapes]))
        seed = grad_values.detach().to(torch.float32).contiguous()
        grads = torch.empty_like(flat)
        stream = torch.cuda.current_stream(flat.device).cuda_stream
        call("xt_tt_sample_adjoint_f32", d, n, r, flat.data_ptr(), pos.data_ptr(), int(pos.shape[0]), seed.data_ptr(),
             grads.data_ptr(), stream)
        out, o = [], 0
        for s in shapes:
            k = s[0] * s[1] * s[2]
            out.append(grads[o: o + k].view(s))
            o += k
        return (None, *out)


def tt_sample(positions: torch.Tensor, cores) -> torch.Tensor:
    return TTSample.apply(positions, *cores)


def adf_loss(positions: torch.Tensor, values: torch.Tensor, cores) -> torch.Tensor:
    """½‖P_Ω(x) − b‖², differentiable in every core through the adjoint kernel."""
    return 0.5 * torch.sum((tt_sample(positions, cores) - values) ** 2)
