"""torch.autograd wrapper of the S(k)-misfit loss and its adjoint kernel.

    loss = StructureFactorLoss.apply(pos, shells, first, last, length_scale, weights, target)

pos is a float32 CUDA tensor [N, 3] (one frame; coordinates need not be
wrapped, since the phase uses k = 2*pi*n/length_scale).  The forward pass computes
rho_k and S(k) on the device through the C ABI (amdat_sk_loss_grad).  The
backward pass calls it again with gradient outputs: the adjoint kernel
recomputes every phase exp(i k.r_j) instead of storing the N x K of them.
Gradients are FP64 on the device and are returned in pos's dtype.

Loss: L = sum_s w_s (S_s - T_s)^2, with S_s = sum_{v in s} |rho_v|^2 / (H_s N).
"""
import numpy as np
import torch

from . import _abi

_CTX = {}


def _context(device):
    idx = device.index if device.index is not None else torch.cuda.current_device()
    if idx not in _CTX:
        _CTX[idx] = _abi.Context(idx)
    ctx = _CTX[idx]
    ctx.set_stream(torch.cuda.current_stream(device).cuda_stream)
    return ctx


def _columns(pos):
    p = pos.detach().to(torch.float32)
    return [p[:, c].contiguous() for c in range(3)]


class StructureFactorLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pos, shells, first, last, length_scale, weights, target):
        if not pos.is_cuda or pos.dim() != 2 or pos.shape[1] != 3:
            raise ValueError("pos must be a CUDA tensor of shape [N, 3]")
        c = _context(pos.device)
        cols = _columns(pos)
        loss, S = _abi.sk_loss_grad(c, pos.shape[0], *[t.data_ptr() for t in cols], shells, first, last,
                                    float(length_scale), weights, target)
        ctx.shells, ctx.first, ctx.last, ctx.ell = shells, first, last, float(length_scale)
        ctx.weights = np.asarray(weights, np.float64)
        ctx.target = np.asarray(target, np.float64)
        ctx.save_for_backward(pos)
        ctx.S = S
        return pos.new_tensor(loss, dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad_out):
        (pos,) = ctx.saved_tensors
        c = _context(pos.device)
        cols = _columns(pos)
        N = pos.shape[0]
        g = [torch.empty(N, dtype=torch.float64, device=pos.device) for _ in range(3)]
        _abi.sk_loss_grad(c, N, *[t.data_ptr() for t in cols], ctx.shells, ctx.first, ctx.last, ctx.ell,
                          ctx.weights, ctx.target, grad_ptrs=[t.data_ptr() for t in g])
        grad = torch.stack(g, 1) * grad_out.to(torch.float64)
        return grad.to(pos.dtype), None, None, None, None, None, None


def structure_factor_loss(pos, shells, first, last, length_scale, weights, target):
    return StructureFactorLoss.apply(pos, shells, first, last, length_scale, weights, target)
