"""Two-phase D3Q19 colour-gradient LBM on the GPU (host-side interface).

Mirrors the reference program ``twophaseflow``
(``LBM-LSM/lsm_lbm_static_alpha_beta_multipore.f``): the same inputs
(INPUT_PARAMETERS values, ``ind`` codes), the same state (``fr``/``fb`` of
``iteration2`` in the reference's node order) and the same diagnostics
(``check_state``, one ``output.m`` row).  The arithmetic runs in
libporescale's HIP kernels; this module only moves arrays.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np

from . import _lib
from .io import LBMInput, write_save_fields

__all__ = ["LBM", "trt_rates", "mrt_rates"]


def trt_rates(nu: float, lam: float, dt: float = 1.0) -> np.ndarray:
    """TRT relaxation vector in the reference's moment order (collmatrTRT,
    ``lsm_lbm_static_alpha_beta_multipore.f:2858-2893``) with magic parameter
    ``lam`` instead of the hard-wired 3/16."""
    sv = 1.0 / (3.0 * nu * dt + 0.5)
    sq = 1.0 / (lam / (1.0 / sv - 0.5) + 0.5)
    return np.array([0, sv, sv, 0, sq, 0, sq, 0, sq, sv, sv, sv, sv, sv, sv, sv, sq, sq, sq], float)


def mrt_rates(nu: float, dt: float = 1.0, se=1.19, seps=1.4, sq=1.2, spi=1.4, sm=1.98) -> np.ndarray:
    """MRT relaxation vector (d'Humieres basis order of collision,
    ``lsm_lbm_static_alpha_beta_multipore.f:2519-2559``): viscous rates from
    ``nu``, the others free."""
    sv = 1.0 / (3.0 * nu * dt + 0.5)
    return np.array([0, se, seps, 0, sq, 0, sq, 0, sq, sv, spi, sv, spi, sv, sv, sv, sm, sm, sm], float)


class LBM:
    """One LBM domain on the GPU.

    Parameters follow the reference's INPUT_PARAMETERS (``inp``).  ``ind`` is
    the voxel code array indexed ``[k, j, i]`` (what ``io.read_ind`` returns).
    ``collision``: "ref" (collmatrTRT), ("trt", lam) or ("mrt", Sr, Sb).
    ``precision``: "f64" (bit-identical to the reference), "f32" (FP32
    deviation + 25-bit colour fraction) or "f32x" (the fraction to 33 bits).
    """

    def __init__(self, inp: LBMInput, ind: np.ndarray, collision="ref", precision: str = "f64"):
        L = _lib.lib()
        self.inp = inp
        nz, ny, nx = ind.shape
        self.shape = (nz, ny, nx)
        p = _lib.LBMParams()
        for n in ("cr2", "rr0", "rb0", "B", "sig", "vir", "vib", "wet", "ee", "lsize", "dt"):
            setattr(p, n, float(getattr(inp, n)))
        p.idir, p.iipore = int(inp.idir), int(inp.iipore)
        p.bfr[:] = [float(x) for x in inp.bfr]
        p.bfb[:] = [float(x) for x in inp.bfb]
        if collision == "ref":
            p.collision = _lib.COLL_REF_TRT
        elif collision[0] == "trt":
            p.collision, p.lambda_ = _lib.COLL_TRT, float(collision[1])
        elif collision[0] == "mrt":
            p.collision = _lib.COLL_MRT
            p.Sr[:] = [float(x) for x in collision[1]]
            p.Sb[:] = [float(x) for x in collision[2]]
        else:
            raise ValueError(f"unknown collision {collision!r}")
        p.precision = {"f64": _lib.PREC_F64, "f32": _lib.PREC_F32, "f32x": _lib.PREC_F32X}[precision]
        self.precision = precision
        self._ind = _lib.i32(ind)
        h = C.c_void_p()
        _lib.check(L.psx_lbm_create(C.byref(p), nx, ny, nz, _lib.ptr(self._ind), C.byref(h)))
        self._h = h
        self._L = L
        c = np.zeros(5, np.int64)
        _lib.check(L.psx_lbm_counts(h, _lib.ptr(c)))
        self.nl, self.nps, self.npb2, self.npb4, self.ns = (int(x) for x in c)

    def close(self):
        if getattr(self, "_h", None):
            self._L.psx_lbm_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ------------------------------------------------------------------
    def tables(self):
        """Reference-format tables (1-based): icor (nl,3), ips (nps,3),
        ipb2 (npb2,4), ipb4 (npb4,6); rows are Fortran columns."""
        icor = np.zeros((self.nl, 3), np.int32)
        ips = np.zeros((self.nps, 3), np.int32)
        ipb2 = np.zeros((self.npb2, 4), np.int32)
        ipb4 = np.zeros((self.npb4, 6), np.int32)
        _lib.check(self._L.psx_lbm_tables(self._h, _lib.ptr(icor), _lib.ptr(ips), _lib.ptr(ipb2),
                                          _lib.ptr(ipb4)))
        return icor, ips, ipb2, ipb4

    def rates(self):
        Sr, Sb = np.zeros(19), np.zeros(19)
        _lib.check(self._L.psx_lbm_rates(self._h, _lib.ptr(Sr), _lib.ptr(Sb)))
        return Sr, Sb

    def initdistr(self, sat: float):
        """Random all-red / all-blue start of initdistr (icon = 0)."""
        _lib.check(self._L.psx_lbm_init_random(self._h, float(sat)))

    def set_state(self, fr: np.ndarray, fb: np.ndarray):
        """fr, fb shaped (nl, 19): the reference's (19, nl) populations."""
        fr, fb = _lib.f64(fr), _lib.f64(fb)
        if fr.shape != (self.nl, 19) or fb.shape != (self.nl, 19):
            raise ValueError(f"state must be ({self.nl}, 19)")
        _lib.check(self._L.psx_lbm_set_state(self._h, _lib.ptr(fr), _lib.ptr(fb)))

    def get_state(self):
        fr, fb = np.zeros((self.nl, 19)), np.zeros((self.nl, 19))
        _lib.check(self._L.psx_lbm_get_state(self._h, _lib.ptr(fr), _lib.ptr(fb)))
        return fr, fb

    def fields(self):
        rr, rb, cf = np.zeros(self.nl), np.zeros(self.nl), np.zeros(self.nl)
        _lib.check(self._L.psx_lbm_get_fields(self._h, _lib.ptr(rr), _lib.ptr(rb), _lib.ptr(cf)))
        return rr, rb, cf

    def save_fields(self):
        """The fluid fields save_results writes (L1:2067-2142), as arrays
        (ncz, ncy, ncx) of the voxel grid: ``velocity`` (.., 3), ``color``,
        ``pressure`` (float64) and ``phase`` (int32 ind_red_blue codes)."""
        nz, ny, nx = self.shape
        u = np.zeros((nz, ny, nx, 3))
        col, pre = np.zeros((nz, ny, nx)), np.zeros((nz, ny, nx))
        ph = np.zeros((nz, ny, nx), np.int32)
        _lib.check(self._L.psx_lbm_fields(self._h, _lib.ptr(u), _lib.ptr(col), _lib.ptr(pre), _lib.ptr(ph)))
        return {"velocity": u, "color": col, "pressure": pre, "phase": ph}

    def step(self, nsteps: int = 1):
        """``nsteps`` iterations of the majorityrule time loop body."""
        _lib.check(self._L.psx_lbm_step(self._h, int(nsteps)))

    def sync(self):
        _lib.check(self._L.psx_lbm_sync(self._h))

    def check_state(self, it1: int = 1, itw: int = 1) -> _lib.LBMDiag:
        d = _lib.LBMDiag()
        _lib.check(self._L.psx_lbm_check_state(self._h, int(it1), int(itw), C.byref(d)))
        return d

    def profile(self, nsteps: int):
        """(total ms, mean collide-kernel ms, mean density-kernel ms)."""
        ms = np.zeros(3)
        _lib.check(self._L.psx_lbm_profile(self._h, int(nsteps), _lib.ptr(ms)))
        return tuple(ms)

    def run(self, itm: int, itw: int, prestep: bool = True, icon: int = 0, save_dir=None, its: int = 1):
        """The majorityrule schedule: 2 pre-steps, then itm*itw steps with a
        check_state after every itw; returns the output.m rows.  With
        ``save_dir``, every its-th iteration writes the save_results files
        there (L1:3633, 3963-3967)."""
        if prestep:
            self.step(2)
        rows = []
        it0 = 1
        isv = 0
        for it1 in range(it0, itm + 1):
            isv += 1
            self.step(itw)
            if save_dir is not None and isv == its:
                isv = 0
                write_save_fields(save_dir, it1, itw, self.save_fields())
            rows.append(list(self.check_state(it1, itw).outm))
        return np.array(rows)


class Coupled:
    """The one-way coupled model of L1 on top of an :class:`LBM`: the solid
    lattice of the same ``ind`` moves under the fluid's momentum exchange
    each main-loop step (L1:3849-3888) and ``calc_ab`` gives Biot's alpha
    tensor and beta per pore family (L1:4008-4468).  Em, nu, rs, dl come
    from the LBM's INPUT_PARAMETERS (lines 26-30)."""

    def __init__(self, lbm: LBM, cube=None):
        from .geometry import cube_tables
        L = lbm._L
        inp = lbm.inp
        bon, ang, sb, ab = cube if cube is not None else cube_tables()
        self._cube = [np.ascontiguousarray(bon, np.int32), np.ascontiguousarray(ang, np.int32),
                      np.ascontiguousarray(sb, np.float64), np.ascontiguousarray(ab, np.float64)]
        h = C.c_void_p()
        _lib.check(L.psx_cpl_create(lbm._h, float(inp.Em), float(inp.nu), float(inp.dl), float(inp.rs),
                                    *(_lib.ptr(x) for x in self._cube), C.byref(h)))
        self._h, self._L, self.lbm = h, L, lbm
        c = np.zeros(4, np.int64)
        _lib.check(L.psx_cpl_counts(h, _lib.ptr(c)))
        self.ns = int(c[0])

    def close(self):
        if getattr(self, "_h", None):
            self._L.psx_cpl_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def nodes(self):
        """icors (ns, 3) 1-based and mijks (ncz, ncy, ncx) L1 solid ids."""
        nz, ny, nx = self.lbm.shape
        icors = np.zeros((self.ns, 3), np.int32)
        mijks = np.zeros((nz, ny, nx), np.int32)
        _lib.check(self._L.psx_cpl_nodes(self._h, _lib.ptr(icors), _lib.ptr(mijks)))
        return icors, mijks

    def step(self, nsteps: int = 1):
        """nsteps main-loop steps: LBM step, fluid force, lsm_step."""
        _lib.check(self._L.psx_cpl_step(self._h, int(nsteps)))

    def get_state(self):
        """rdat, vdat, adat, afdat as (ns, 3) in L1's solid numbering."""
        r, v, a, af = (np.zeros((self.ns, 3)) for _ in range(4))
        _lib.check(self._L.psx_cpl_get_state(self._h, *(_lib.ptr(x) for x in (r, v, a, af))))
        return r, v, a, af

    def set_state(self, r, v, a):
        arrs = [np.ascontiguousarray(x, np.float64).reshape(self.ns, 3) for x in (r, v, a)]
        _lib.check(self._L.psx_cpl_set_state(self._h, *(_lib.ptr(x) for x in arrs)))

    def calc_ab(self, npores=None, pres=0.0):
        """asigmap, asigmam, astrain (3,3 sums over the solid cubes, row =
        force component), epsi, alphai (3,3), ncub, betai (npores,)."""
        npores = int(self.lbm.inp.npores if npores is None else npores)
        out = np.zeros(38 + npores)
        _lib.check(self._L.psx_cpl_calc_ab(self._h, npores, float(pres), _lib.ptr(out)))
        t = lambda a: a.reshape(3, 3).T
        return {"asigmap": t(out[0:9]), "asigmam": t(out[9:18]), "astrain": t(out[18:27]), "epsi": out[27],
                "alphai": t(out[28:37]), "ncub": int(out[37]), "betai": out[38:]}
