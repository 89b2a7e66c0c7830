"""torch.autograd wrapper of the LSM moduli and their adjoint.

``LSMModuli.apply(ab, stiff, ind, gstrain, rtol, precision)`` returns the
tensor (K, G) of Voigt bulk and shear moduli; its backward calls the HIP
adjoint (psx_lsm_adjoint), giving d/d(alf1, bet1, alf2, bet2) and
d/d(per-voxel stiffness) in one extra batched solve.  Intended for
calibrating spring constants to target moduli.
"""
from __future__ import annotations

import numpy as np
import torch

from .lsm import LSM

__all__ = ["LSMModuli", "calibrate"]


class LSMModuli(torch.autograd.Function):
    @staticmethod
    def forward(ctx, ab, stiff, ind, gstrain=0.01, rtol=1e-10, precision="f64"):
        a = ab.detach().cpu().double().numpy()
        s = None if stiff is None else stiff.detach().cpu().double().numpy()
        lsm = LSM(ind, a[0], a[1], a[2], a[3], gstrain=gstrain, precision=precision, stiff=s)
        _, K, G = lsm.moduli(rtol=rtol)
        ctx.lsm = lsm
        ctx.rtol = rtol
        ctx.has_stiff = stiff is not None
        return torch.tensor([K, G], dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad):
        g = grad.detach().cpu().double().numpy()
        dab, ds = ctx.lsm.adjoint(float(g[0]), float(g[1]), rtol=ctx.rtol, want_stiff=ctx.has_stiff)
        ctx.lsm.close()
        gab = torch.from_numpy(dab)
        gs = torch.from_numpy(ds) if ctx.has_stiff else None
        return gab, gs, None, None, None, None


def calibrate(ind, ab0, target_KG, steps=50, lr=0.2, gstrain=0.01, rtol=1e-10, precision="f64"):
    """Fits (alf1, bet1, alf2, bet2) to target (K, G) by gradient descent on
    log-parameters; returns (ab history (steps+1, 4), loss history)."""
    t = torch.tensor(target_KG, dtype=torch.float64)
    logab = torch.tensor(np.log(np.asarray(ab0, float)), dtype=torch.float64, requires_grad=True)
    opt = torch.optim.Adam([logab], lr=lr)
    hist, losses = [logab.detach().exp().numpy().copy()], []
    for _ in range(steps):
        opt.zero_grad()
        KG = LSMModuli.apply(logab.exp(), None, ind, gstrain, rtol, precision)
        loss = (((KG - t) / t) ** 2).sum()
        loss.backward()
        opt.step()
        losses.append(float(loss))
        hist.append(logab.detach().exp().numpy().copy())
    return np.array(hist), np.array(losses)
