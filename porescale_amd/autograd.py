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
        # the handle stays open until the graph is freed: backward may run
        # more than once on one forward (retain_graph, gradcheck)
        dab, ds = ctx.lsm.adjoint(float(g[0]), float(g[1]), rtol=ctx.rtol, want_stiff=ctx.has_stiff)
        gab = torch.from_numpy(dab)
        gs = torch.from_numpy(ds) if ctx.has_stiff else None
        return gab, gs, None, None, None, None


def calibrate(ind, ab0, target_KG, steps=50, lr=0.2, gstrain=0.01, rtol=1e-10, precision="f64",
              method="gauss-newton", stiff=None):
    """Fits (alf1, bet1, alf2, bet2) to target (K, G) in log-parameters.

    method "gauss-newton": each step solves the six load cases, then two
    adjoint solves give the 2x4 Jacobian of (K, G); the step is the
    minimum-norm Gauss-Newton step (length-capped at 0.5 in log space).
    method "adam": first-order steps through ``LSMModuli`` (torch.autograd)
    on the relative squared error.
    Returns (ab history (steps+1, 4), loss history (steps,)), loss being the
    relative squared error at the parameters before each step.
    """
    t = np.asarray(target_KG, float)
    logab = np.log(np.asarray(ab0, float))
    hist, losses = [np.exp(logab)], []
    if method == "adam":
        tt = torch.tensor(t, dtype=torch.float64)
        x = torch.tensor(logab, dtype=torch.float64, requires_grad=True)
        opt = torch.optim.Adam([x], lr=lr)
        st = None if stiff is None else torch.as_tensor(stiff, dtype=torch.float64)
        for _ in range(steps):
            opt.zero_grad()
            KG = LSMModuli.apply(x.exp(), st, ind, gstrain, rtol, precision)
            loss = (((KG - tt) / tt) ** 2).sum()
            loss.backward()
            opt.step()
            losses.append(float(loss.detach()))
            hist.append(x.detach().exp().numpy().copy())
        return np.array(hist), np.array(losses)
    if method != "gauss-newton":
        raise ValueError(f"unknown method {method!r}")
    # one handle for the whole fit: each step replaces its spring constants
    # (psx_lsm_set_springs), which refreshes the multigrid operators in place
    with LSM(ind, *np.exp(logab), gstrain=gstrain, precision=precision, stiff=stiff) as s:
        for k in range(steps):
            ab = np.exp(logab)
            if k:
                s.set_springs(*ab)
            _, K, G = s.moduli(rtol=rtol)
            dK, _ = s.adjoint(1.0, 0.0, rtol=rtol, want_stiff=False)
            dG, _ = s.adjoint(0.0, 1.0, rtol=rtol, want_stiff=False)
            r = np.array([K, G]) / t - 1.0
            losses.append(float(r @ r))
            J = np.stack([dK / t[0], dG / t[1]]) * ab[None, :]
            delta = -np.linalg.lstsq(J, r, rcond=None)[0]
            n = np.linalg.norm(delta)
            if n > 0.5:
                delta *= 0.5 / n
            logab = logab + delta
            hist.append(np.exp(logab))
    return np.array(hist), np.array(losses)
