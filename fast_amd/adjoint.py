"""The transposed residual Jacobian behind ``torch.autograd``.

``flux_balance(W, ctx, zone)`` is the AUSM+(P) o3 Euler face balance R(W) of a zone's interior
cells (the fluxes the reference's RHS kernels accumulate into drodm), W the primitive state of
every cell of the zone (5 x NDIMDX, ghost cells included, float64).  Its backward is the
transposed Jacobian product of ``fm_residual_vjp`` (the reverse mode the reference holds for 2-D
zones as ADJOINT/AUSM/2d/fluausm_euler_o3_2d_b.for).  Both run on the GPU through the C ABI; the
zone's time level ``level`` (default 0) is the working state and is overwritten with W.
``flux_balance_zones(Ws, ctx)`` does the same for every zone of a context joined by injection
transfers, with the transfers' transpose in its backward (``fm_residual_vjp_zones``).  W and
the gradients are host (CPU) tensors: the library brings its own HIP runtime, so torch's CUDA
device is not used in the same process.
"""
import numpy as np
import torch

from . import _lib


class FluxBalance(torch.autograd.Function):
    @staticmethod
    def forward(ctx, W, fmctx, zone, level):
        if W.dtype != torch.float64:
            raise TypeError("W must be float64")
        w = W.detach().cpu().numpy()
        fmctx.set(zone, (_lib.FIELD_RO, _lib.FIELD_RO_M1, _lib.FIELD_RO_P1)[level], w)
        R = fmctx.flux_balance(zone, level)
        ctx.fm = (fmctx, zone, level)
        ctx.save_for_backward(W)
        return torch.from_numpy(R).to(W.device)

    @staticmethod
    def backward(ctx, gR):
        (W,) = ctx.saved_tensors
        fmctx, zone, level = ctx.fm
        fmctx.set(zone, (_lib.FIELD_RO, _lib.FIELD_RO_M1, _lib.FIELD_RO_P1)[level], W.detach().cpu().numpy())
        g = fmctx.residual_vjp(zone, np.ascontiguousarray(gR.detach().cpu().numpy()), level)
        return torch.from_numpy(g).to(W.device), None, None, None


def flux_balance(W, fmctx, zone, level=0):
    """R(W) with the transposed Jacobian product as its backward."""
    return FluxBalance.apply(W, fmctx, zone, level)


_LEVELS = (_lib.FIELD_RO, _lib.FIELD_RO_M1, _lib.FIELD_RO_P1)


def _load(fmctx, Ws, level):
    for z, W in enumerate(Ws):
        if W.dtype != torch.float64:
            raise TypeError("every W must be float64")
        fmctx.set(z, _LEVELS[level], W.detach().cpu().numpy())
    fmctx.transfer(level)


class FluxBalanceZones(torch.autograd.Function):
    @staticmethod
    def forward(ctx, fmctx, level, *Ws):
        _load(fmctx, Ws, level)
        Rs = [torch.from_numpy(fmctx.flux_balance(z, level)).to(W.device) for z, W in enumerate(Ws)]
        ctx.fm = (fmctx, level)
        ctx.save_for_backward(*Ws)
        return tuple(Rs)

    @staticmethod
    def backward(ctx, *gRs):
        Ws = ctx.saved_tensors
        fmctx, level = ctx.fm
        _load(fmctx, Ws, level)
        gs = fmctx.residual_vjp_zones([np.ascontiguousarray(g.detach().cpu().numpy()) for g in gRs], level)
        return (None, None) + tuple(torch.from_numpy(g).to(W.device) for g, W in zip(gs, Ws))


def flux_balance_zones(Ws, fmctx, level=0):
    """R_n(W) of every zone of ``fmctx`` (Ws[n]: 5 x NDIMDX of zone n, in the order the zones were
    added), the ghost cells of each W first filled by the context's transfers; the backward is
    fm_residual_vjp_zones, so the gradient of a raccord-filled ghost lands on its donor cell and
    the ghost itself gets zero (the forward overwrites it)."""
    return FluxBalanceZones.apply(fmctx, level, *Ws)
