"""The synthetic workloads the benchmark and the parity tests run.

Each builder returns the voxel codes (``ind[k, j, i]``, 1 = pore, 0 =
solid), the INPUT_PARAMETERS values (:class:`io.LBMInput`), the collision
choice for :class:`lbm.LBM` and the initial blue voxels; ``initial_state``
turns the latter into the rest-equilibrium start (all red or all blue per
node) that ``initdistr`` writes (lsm_lbm_static_alpha_beta_multipore.f:
2834-2840), in the reference's node order, for ``set_state``.
"""
from __future__ import annotations

import numpy as np

from . import geometry as geo
from . import io as pio
from .lbm import mrt_rates, trt_rates

__all__ = ["twophase_trt_pack", "twophase_mrt_pack", "initial_state"]


def twophase_trt_pack(n: int = 64):
    """Two-phase TRT on a periodic sphere pack: radius 6, porosity 0.35,
    seed 0; tau = 1 (nu = 1/6) with magic parameter 1/4, A = 0.01, recolouring
    0.7, no wetting spec (wet 0), body force 1e-5 along z, blob start at
    blue saturation 0.5.  At n = 64 the reference method's own case."""
    ind = geo.sphere_pack((n, n, n), 0.35, 6.0, seed=0)
    inp = pio.LBMInput(B=0.7, sig=0.01, vir=1.0 / 6.0, vib=1.0 / 6.0, wet=0.0, idir=3, ncx=n, ncy=n, ncz=n,
                       dt=1.0, bfr=[0.0, 0.0, 1e-5], bfb=[0.0, 0.0, 1e-5])
    Sr = trt_rates(inp.vir, 0.25)
    blue = geo.phase_blobs(ind, 0.5, seed=0, radius=4.0)
    return ind, inp, ("trt", 0.25), (Sr, Sr), blue


def twophase_mrt_pack(n: int = 256, radii=(4, 10), seed: int = 0):
    """Two-phase MRT on a periodic sphere pack: radii uniform in 4..10,
    porosity 0.30; red viscosity 1/6, blue 1/60 (ratio 1:10), 45 degree
    contact angle (wet = cos 45), A = 0.01, recolouring 0.7, body force 1e-5
    along z, blob start at blue saturation 0.5.  Smaller ``n`` with smaller
    ``radii`` give the same physics at test sizes."""
    ind = geo.sphere_pack((n, n, n), 0.30, radii, seed=seed)
    inp = pio.LBMInput(B=0.7, sig=0.01, vir=1.0 / 6.0, vib=1.0 / 60.0, wet=0.70710678118654752, idir=3,
                       ncx=n, ncy=n, ncz=n, dt=1.0, bfr=[0.0, 0.0, 1e-5], bfb=[0.0, 0.0, 1e-5])
    Sr, Sb = mrt_rates(inp.vir), mrt_rates(inp.vib)
    blue = geo.phase_blobs(ind, 0.5, seed=seed, radius=4.0 if n >= 128 else 3.0)
    return ind, inp, ("mrt", Sr, Sb), (Sr, Sb), blue


def initial_state(inp: pio.LBMInput, ind: np.ndarray, blue: np.ndarray):
    """Rest equilibrium, all red or all blue per fluid node, (nl, 19) each,
    fluid nodes in the reference's order (x outer, z fastest)."""
    dt = inp.dt
    cb2 = inp.rr0 * inp.cr2 / inp.rb0
    tr = np.array([1 - 2 * inp.cr2 * dt * dt] + [inp.cr2 * dt * dt / 6] * 6 + [inp.cr2 * dt * dt / 12] * 12)
    tb = np.array([1 - 2 * cb2 * dt * dt] + [cb2 * dt * dt / 6] * 6 + [cb2 * dt * dt / 12] * 12)
    t_ind = np.transpose(ind, (2, 1, 0))
    bl = np.transpose(blue, (2, 1, 0))[t_ind == inp.iipore]
    c = np.where(bl, 0.0, 1.0)
    return tr[None, :] * c[:, None] * inp.rr0, tb[None, :] * (1 - c[:, None]) * inp.rb0
