"""ctypes binding of the CPU oracle (oracle/liboracle.so) and runners for the
flang-compiled reference programs (oracle/_ref/*).  Test infrastructure only."""
from __future__ import annotations

import ctypes as C
import os
import shutil
import subprocess
import tempfile

import numpy as np

from porescale_amd import io as pio
from porescale_amd import geometry as geo

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
REF_DIR = os.path.join(ORACLE_DIR, "_ref")

_P = np.ctypeslib.ndpointer
_D = C.POINTER(C.c_double)


class LBMParams(C.Structure):
    _fields_ = [("cr2", C.c_double), ("rr0", C.c_double), ("rb0", C.c_double), ("B", C.c_double),
                ("sig", C.c_double), ("vir", C.c_double), ("vib", C.c_double), ("wet", C.c_double),
                ("idir", C.c_int32), ("iipore", C.c_int32), ("eps", C.c_double), ("ee", C.c_double),
                ("lsize", C.c_double), ("bfr", C.c_double * 3), ("bfb", C.c_double * 3),
                ("dt", C.c_double), ("use_custom_S", C.c_int32), ("Sr", C.c_double * 19),
                ("Sb", C.c_double * 19), ("emulate_int32_nsum", C.c_int32)]


class LBMDiag(C.Structure):
    _fields_ = [("permr", C.c_double), ("permb", C.c_double), ("Qr", C.c_double * 3),
                ("Qb", C.c_double * 3), ("um", C.c_double * 3), ("umr", C.c_double * 3),
                ("umb", C.c_double * 3), ("tfr", C.c_double * 3), ("tfb", C.c_double * 3),
                ("nr", C.c_double), ("nb", C.c_double), ("tmr", C.c_double), ("tmb", C.c_double),
                ("tm", C.c_double), ("Cpr", C.c_double), ("Cpb", C.c_double), ("Re", C.c_double),
                ("isneg", C.c_int32), ("negdens_r", C.c_int32), ("negdens_b", C.c_int32),
                ("pad_", C.c_int32), ("outm", C.c_double * 12)]


class LSMParams(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("pad_axis", C.c_int32),
                ("jump_comp", C.c_int32), ("nthread", C.c_int32), ("alf1", C.c_double),
                ("bet1", C.c_double), ("alf2", C.c_double), ("bet2", C.c_double),
                ("rs", C.c_double), ("dt", C.c_double), ("gstrain", C.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(ORACLE_DIR, "liboracle.so")
        if not os.path.exists(path):
            subprocess.check_call(["make", "-s", "-C", ORACLE_DIR, "liboracle.so"])
        L = C.CDLL(path)
        vp = C.c_void_p
        L.or_lbm_create.restype = vp
        L.or_lbm_create.argtypes = [C.POINTER(LBMParams), C.c_int32, C.c_int32, C.c_int32, _P(np.int32)]
        L.or_lbm_destroy.argtypes = [vp]
        L.or_lbm_count.restype = C.c_int64
        L.or_lbm_count.argtypes = [vp, C.c_int32]
        L.or_lbm_get_tables.argtypes = [vp, vp, vp, vp, vp]
        L.or_lbm_init_random.argtypes = [vp, C.c_double]
        L.or_lbm_set_state.argtypes = [vp, _P(np.float64), _P(np.float64)]
        L.or_lbm_get_state.argtypes = [vp, _P(np.float64), _P(np.float64)]
        L.or_lbm_get_fields.argtypes = [vp, vp, vp, vp]
        L.or_lbm_step.argtypes = [vp, C.c_int64]
        L.or_lbm_check_state.argtypes = [vp, C.c_int64, C.c_int64, C.POINTER(LBMDiag)]
        L.or_lbm_get_S.argtypes = [vp, _P(np.float64), _P(np.float64)]
        L.or_lbm_save_fields.argtypes = [vp, vp, vp, vp, vp]
        L.or_collision.argtypes = [C.c_double] * 3 + [_P(np.float64), _P(np.float64), C.c_double,
                                                     _P(np.float64), _P(np.float64), _P(np.float64),
                                                     C.c_double]
        L.or_collmatr_trt.argtypes = [C.c_double, _P(np.float64)]
        L.or_lsm_create.restype = vp
        L.or_lsm_create.argtypes = [C.POINTER(LSMParams), _P(np.int32), _P(np.int32), _P(np.int32),
                                    _P(np.float64), _P(np.float64)]
        L.or_lsm_destroy.argtypes = [vp]
        L.or_num_threads.restype = C.c_int
        L.or_num_threads.argtypes = []
        L.or_set_threads.argtypes = [C.c_int]
        L.or_lsm_count.restype = C.c_int64
        L.or_lsm_count.argtypes = [vp, C.c_int32]
        L.or_lsm_get_nodes.argtypes = [vp, vp, vp]
        L.or_lsm_layers.argtypes = [vp, vp, vp, vp]
        L.or_lsm_get_tables.argtypes = [vp] + [vp] * 6
        L.or_lsm_get_state.argtypes = [vp, vp, vp, vp]
        L.or_lsm_set_state.argtypes = [vp, vp, vp, vp]
        L.or_lsm_forces.argtypes = [vp, _P(np.float64), _P(np.float64)]
        L.or_lsm_verlet.argtypes = [vp, C.c_int64]
        L.or_lsm_stress.restype = C.c_int64
        L.or_lsm_stress.argtypes = [vp, _P(np.float64), _P(np.float64), _P(np.float64)]
        L.or_lsm_anf.argtypes = [_P(np.float64), _P(np.int32)]
        L.or_lbm_fluid_force.argtypes = [vp, C.c_double, _P(np.float64)]
        L.or_lsm_coupled_step.argtypes = [vp, _P(np.float64), C.c_double]
        L.or_lsm_restart.argtypes = [vp, vp, vp, vp]
        L.or_lsm_calc_ab.argtypes = [vp, _P(np.int32), C.c_int32, C.c_int32, C.c_double, _P(np.float64)]
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def lbm_params(inp: pio.LBMInput, S=None, emulate_nsum=False) -> LBMParams:
    p = LBMParams()
    for n in ["cr2", "rr0", "rb0", "B", "sig", "vir", "vib", "wet", "eps", "ee", "lsize", "dt"]:
        setattr(p, n, float(getattr(inp, n)))
    p.idir, p.iipore = int(inp.idir), int(inp.iipore)
    p.bfr[:] = [float(x) for x in inp.bfr]
    p.bfb[:] = [float(x) for x in inp.bfb]
    if S is not None:
        p.use_custom_S = 1
        p.Sr[:] = list(S[0])
        p.Sb[:] = list(S[1])
    p.emulate_int32_nsum = 1 if emulate_nsum else 0
    return p


class OracleLBM:
    """The reference LBM restated in C; state arrays are (nl, 19) = Fortran (19, nl)."""

    def __init__(self, inp: pio.LBMInput, ind: np.ndarray, S=None, emulate_nsum=False):
        self.L = lib()
        self.inp = inp
        nz, ny, nx = ind.shape
        self._ind = np.ascontiguousarray(ind, dtype=np.int32)
        self.p = lbm_params(inp, S, emulate_nsum)
        self.h = self.L.or_lbm_create(C.byref(self.p), nx, ny, nz, self._ind)
        self.nl = self.L.or_lbm_count(self.h, 0)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.or_lbm_destroy(self.h)
            self.h = None

    def counts(self):
        return {k: self.L.or_lbm_count(self.h, i) for i, k in enumerate(["nl", "nps", "npb2", "npb4", "ns"])}

    def tables(self):
        c = self.counts()
        icor = np.zeros((c["nl"], 3), np.int32)
        ips = np.zeros((c["nps"], 3), np.int32)
        ipb2 = np.zeros((c["npb2"], 4), np.int32)
        ipb4 = np.zeros((c["npb4"], 6), np.int32)
        self.L.or_lbm_get_tables(self.h, _ptr(icor), _ptr(ips), _ptr(ipb2), _ptr(ipb4))
        return icor, ips, ipb2, ipb4

    def init_random(self, sat):
        self.L.or_lbm_init_random(self.h, float(sat))

    def set_state(self, fr, fb):
        self.L.or_lbm_set_state(self.h, np.ascontiguousarray(fr, np.float64),
                                np.ascontiguousarray(fb, np.float64))

    def state(self):
        fr = np.zeros((self.nl, 19))
        fb = np.zeros((self.nl, 19))
        self.L.or_lbm_get_state(self.h, fr, fb)
        return fr, fb

    def fields(self):
        rr, rb, cf = (np.zeros(self.nl) for _ in range(3))
        self.L.or_lbm_get_fields(self.h, _ptr(rr), _ptr(rb), _ptr(cf))
        return rr, rb, cf

    def save_fields(self):
        """save_results fields on the voxel grid (nz, ny, nx[, 3])."""
        nz, ny, nx = self._ind.shape
        u = np.zeros((nz, ny, nx, 3))
        col, pre = np.zeros((nz, ny, nx)), np.zeros((nz, ny, nx))
        ph = np.zeros((nz, ny, nx), np.int32)
        self.L.or_lbm_save_fields(self.h, _ptr(u), _ptr(col), _ptr(pre), _ptr(ph))
        return {"velocity": u, "color": col, "pressure": pre, "phase": ph}

    def step(self, n=1):
        self.L.or_lbm_step(self.h, int(n))

    def check_state(self, it1=1, itw=1) -> LBMDiag:
        d = LBMDiag()
        self.L.or_lbm_check_state(self.h, int(it1), int(itw), C.byref(d))
        return d

    def S(self):
        a, b = np.zeros(19), np.zeros(19)
        self.L.or_lbm_get_S(self.h, a, b)
        return a, b


def equilibrium_state(inp: pio.LBMInput, blue_nodes: np.ndarray):
    """Rest equilibrium per node, all-red or all-blue, as initdistr builds it
    (lsm_lbm_static_alpha_beta_multipore.f:2834-2840): fr = tr*c*rr0, fb = tb*(1-c)*rb0."""
    dt = inp.dt
    cb2 = inp.rr0 * inp.cr2 / inp.rb0
    tr = np.array([1.0 - 2.0 * inp.cr2 * dt * dt] + [inp.cr2 * dt * dt * (1.0 / 6.0)] * 6
                  + [inp.cr2 * dt * dt * (1.0 / 12.0)] * 12)
    tb = np.array([1.0 - 2.0 * cb2 * dt * dt] + [cb2 * dt * dt * (1.0 / 6.0)] * 6
                  + [cb2 * dt * dt * (1.0 / 12.0)] * 12)
    cfield = np.where(blue_nodes, 0.0, 1.0)
    fr = tr[None, :] * cfield[:, None] * inp.rr0
    fb = tb[None, :] * (1.0 - cfield[:, None]) * inp.rb0
    return fr, fb


def fluid_order_mask(ind: np.ndarray, mask: np.ndarray, iipore=1) -> np.ndarray:
    """Restrict a voxel mask [k,j,i] to fluid nodes in the reference's node order
    (x outer, y, z fastest; lsm_lbm_static_alpha_beta_multipore.f:1060-1072)."""
    t_ind = np.transpose(ind, (2, 1, 0))   # [i, j, k]
    t_m = np.transpose(mask, (2, 1, 0))
    return t_m[t_ind == iipore]


class OracleLSM:
    def __init__(self, ind, case="xx", alf=(1.0, 0.1), bet=None, alf2=None, bet2=None,
                 rs=1.0, dt=0.01, gstrain=0.01, nthread=1, cube=None):
        self.L = lib()
        nz, ny, nx = ind.shape
        # "periodic": L1's solid lattice (no padded axis, L1:501-1017)
        pad, comp = (-1, 0) if case == "periodic" else CASES[case]
        p = LSMParams()
        p.nx, p.ny, p.nz, p.pad_axis, p.jump_comp, p.nthread = nx, ny, nz, pad, comp, nthread
        a1, b1 = alf if bet is None else (alf, bet)
        p.alf1, p.bet1 = a1, b1
        p.alf2 = a1 if alf2 is None else alf2
        p.bet2 = b1 if bet2 is None else bet2
        p.rs, p.dt, p.gstrain = rs, dt, gstrain
        self.p = p
        self._ind = np.ascontiguousarray(ind, np.int32)
        cb, ca, cs, cab = cube if cube is not None else geo.cube_tables()
        self._cube = [np.ascontiguousarray(x) for x in (cb, ca, cs, cab)]
        self.h = self.L.or_lsm_create(C.byref(p), self._ind, *self._cube)
        self.ns = self.L.or_lsm_count(self.h, 0)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.or_lsm_destroy(self.h)
            self.h = None

    def counts(self):
        return {k: self.L.or_lsm_count(self.h, i)
                for i, k in enumerate(["ns", "nbo", "nna", "Gx", "Gy", "Gz", "npair"])}

    def layers(self):
        """(layb, lib, lia): node, bond and angle offsets of the z-layer groups."""
        n = 2 * self.p.nthread + 1
        out = [np.zeros(n, np.int64) for _ in range(3)]
        self.L.or_lsm_layers(self.h, *(_ptr(x) for x in out))
        return tuple(out)

    def nodes(self):
        c = self.counts()
        icors = np.zeros((c["ns"], 3), np.int32)
        mijks = np.zeros((c["Gz"], c["Gy"], c["Gx"]), np.int32)
        self.L.or_lsm_get_nodes(self.h, _ptr(icors), _ptr(mijks))
        return icors, mijks

    def tables(self):
        c = self.counts()
        bon = np.zeros((c["nbo"], 3), np.int32)
        s1, s2 = np.zeros(c["nbo"]), np.zeros(c["nbo"])
        ang = np.zeros((c["nna"], 6), np.int32)
        a1, a2 = np.zeros(c["nna"]), np.zeros(c["nna"])
        self.L.or_lsm_get_tables(self.h, *(_ptr(x) for x in (bon, s1, s2, ang, a1, a2)))
        return bon, s1, s2, ang, a1, a2

    def state(self):
        r, v, a = (np.zeros((self.ns, 3)) for _ in range(3))
        self.L.or_lsm_get_state(self.h, _ptr(r), _ptr(v), _ptr(a))
        return r, v, a

    def set_state(self, r=None, v=None, a=None):
        conv = [None if x is None else np.ascontiguousarray(x, np.float64) for x in (r, v, a)]
        self.L.or_lsm_set_state(self.h, *(_ptr(x) for x in conv))

    def restart(self, r, v, a):
        """icon = 1: state from iteration2, glued pairs reset (XX:1366-1375)."""
        self._rs = [np.ascontiguousarray(x, np.float64) for x in (r, v, a)]
        self.L.or_lsm_restart(self.h, *(_ptr(x) for x in self._rs))

    def forces(self, r):
        f = np.zeros((self.ns, 3))
        self.L.or_lsm_forces(self.h, np.ascontiguousarray(r, np.float64), f)
        return f

    def verlet(self, n):
        self.L.or_lsm_verlet(self.h, int(n))

    def stress(self):
        s, e, m = np.zeros(9), np.zeros(9), np.zeros(9)
        ncub = self.L.or_lsm_stress(self.h, s, e, m)
        return s.reshape(3, 3, order="F"), e.reshape(3, 3, order="F"), m.reshape(3, 3, order="F"), ncub


# load case -> (padded axis, jump component); xx and yz are the shipped
# XX_P.f / YZ_P.f programs, the others generalise them cyclically.
CASES = {"xx": (0, 0), "yy": (1, 1), "zz": (2, 2), "yz": (1, 2), "zx": (2, 0), "xy": (0, 1)}


# ----------------------------------------------------------- reference runs --
def ref_available(name: str) -> bool:
    return os.path.exists(os.path.join(REF_DIR, name))


def _prep_dir(ind: np.ndarray) -> str:
    d = tempfile.mkdtemp(prefix="psx_ref_")
    pio.write_ind(os.path.join(d, "ind"), ind)
    pio.write_velocities(os.path.join(d, "velocities19"), pio.D3Q19)
    pio.write_velocities(os.path.join(d, "velocities9"), pio.BOND9)
    pio.write_cube_tables(d, *geo.cube_tables())
    return d


def _run(exe: str, d: str, timeout: float) -> None:
    cmd = f"ulimit -s unlimited; exec {os.path.join(REF_DIR, exe)}"
    subprocess.run(["bash", "-c", cmd], cwd=d, check=True, timeout=timeout,
                   stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)


def run_ref_lbm(inp: pio.LBMInput, ind: np.ndarray, restart=None, timeout=600.0, fields=False):
    """Run the flang-compiled L1 program; returns (fr, fb, output.m rows, it1)
    and, with ``fields``, the last save_results files as a dict (velocity
    (nv, 3), color (nv,), pressure (nv,), voxels i fastest).

    ``restart`` = (fr, fb) injects an initial state through iteration2 (icon=1);
    the program then runs 2 pre-steps plus (itm - 1) * itw steps.
    """
    d = _prep_dir(ind)
    try:
        if restart is not None:
            inp = _copy(inp, icon=1)
            pio.write_lbm_restart(os.path.join(d, "iteration2"), 0, restart[0], restart[1])
        pio.write_lbm_input(os.path.join(d, "INPUT_PARAMETERS"), inp)
        nl = int((ind == inp.iipore).sum())
        _run("l1_ref", d, timeout)
        it1, fr, fb = pio.read_lbm_restart(os.path.join(d, "iteration2"), nl)
        outm = pio.read_output_m(os.path.join(d, "output.m"))
        if not fields:
            return fr, fb, outm, it1
        return fr, fb, outm, it1, read_field_files(d, it1, ind.shape)
    finally:
        shutil.rmtree(d, ignore_errors=True)


def run_ref_coupled(inp: pio.LBMInput, ind: np.ndarray, timeout=600.0):
    """Run L1 (icon as in inp); returns (fr, fb, output.m rows, it1, solid
    files of the last save (read_lsm_files), last alpha_beta block)."""
    d = _prep_dir(ind)
    try:
        pio.write_lbm_input(os.path.join(d, "INPUT_PARAMETERS"), inp)
        nl = int((ind == inp.iipore).sum())
        _run("l1_ref", d, timeout)
        it1, fr, fb = pio.read_lbm_restart(os.path.join(d, "iteration2"), nl)
        outm = pio.read_output_m(os.path.join(d, "output.m"))
        return (fr, fb, outm, it1, read_lsm_files(d, it1, ind.shape),
                read_alpha_beta(os.path.join(d, "alpha_beta"), int(inp.npores)))
    finally:
        shutil.rmtree(d, ignore_errors=True)


def read_field_files(d: str, it1: int, shape) -> dict:
    """velocity_field, color_field, pressure<it1> and the last ind_red_blue
    block as written by save_results; ``shape`` = (ncz, ncy, ncx)."""
    with open(os.path.join(d, "velocity_field")) as fh:
        vel = pio.parse_fortran_numbers(fh.read()).reshape(-1, 3)
    with open(os.path.join(d, "color_field")) as fh:
        col = pio.parse_fortran_numbers(fh.read())
    with open(os.path.join(d, f"pressure{it1}")) as fh:
        pre = pio.parse_fortran_numbers(fh.read())
    ncz, ncy, ncx = shape
    it, ph = pio.read_ind_red_blue(os.path.join(d, "ind_red_blue"), ncx, ncy, ncz)[-1]
    return {"velocity": vel.reshape(ncz, ncy, ncx, 3), "color": col.reshape(shape),
            "pressure": pre.reshape(shape), "phase": ph, "iter": it}


def read_lsm_block_files(d: str) -> dict:
    """The field files XX_P.f writes after each block (XX:1546-1592): coord,
    displ, vel, force as (rows, 3) over the padded grid (x fastest) and the
    FTO/VTO/RTO sums of 'write'."""
    out = {}
    for key, name in (("coord", "LSM_coord_field"), ("displ", "LSM_displ_field"), ("vel", "LSM_velocity_field"),
                      ("force", "LSM_force_field")):
        with open(os.path.join(d, name)) as fh:
            out[key] = pio.parse_fortran_numbers(fh.read()).reshape(-1, 3)
    with open(os.path.join(d, "write")) as fh:
        lines = fh.read().splitlines()
    for key in ("FTO", "VTO", "RTO"):
        line = next(l for l in lines if l.strip().startswith(key))
        out[key] = pio.parse_fortran_numbers(line.split("=", 1)[1])
    return out


def run_ref_lsm(case: str, ind: np.ndarray, inp: pio.LSMInput, restart=None, timeout=600.0, files=False,
                settings=False):
    """Run XX_P.f ('xx') or YZ_P.f ('yz'); returns (r, v, a, stress_tensor text)
    and, with ``files``, the last block's field files (read_lsm_block_files);
    with ``settings``, also the text of lsm_settings."""
    exe = {"xx": "xx_ref", "yz": "yz_ref"}[case]
    d = _prep_dir(ind)
    try:
        if restart is not None:
            inp = _copy(inp, icon=1)
            pio.write_lsm_restart(os.path.join(d, "iteration2"), *restart)
        pio.write_lsm_input(os.path.join(d, "INPUT_PARAMETERS"), inp)
        _run(exe, d, timeout)
        with open(os.path.join(d, "stress_tensor")) as fh:
            st = fh.read()
        with open(os.path.join(d, "icors")) as fh:
            ns = sum(1 for line in fh if line.strip())
        r, v, a = pio.read_lsm_restart(os.path.join(d, "iteration2"), ns)
        out = (r, v, a, st)
        if files:
            out += (read_lsm_block_files(d),)
        if settings:
            with open(os.path.join(d, "lsm_settings")) as fh:
                out += (fh.read(),)
        return out
    finally:
        shutil.rmtree(d, ignore_errors=True)


def settings_tables(text: str) -> dict:
    """nbo, nna and the layer arrays from the text of an lsm_settings file."""
    out = {}
    for line in text.splitlines():
        key, sep, vals = line.partition("=")
        key = key.strip()
        if sep and key in ("nbo", "nna", "layers", "layb", "lib", "lia"):
            out[key] = [int(x) for x in vals.split()]
    return out


def _copy(obj, **kw):
    import dataclasses
    return dataclasses.replace(obj, **kw)


def parse_stress_tensor(text: str) -> dict:
    """Last Exxxx/Eyyxx/Eyyzz/Eyzyz values of a stress_tensor file."""
    out = {}
    for line in text.splitlines():
        s = line.strip()
        for key in ("Exxxx", "Eyyxx", "Eyyzz", "Eyzyz"):
            if s.startswith(key):
                out[key] = pio.parse_fortran_numbers(s.split("=", 1)[1])[0]
    return out


def coupled_springs(Em: float, nu: float, dl: float):
    """alf, bet of the coupled model (L1:497-498)."""
    alf = dl * Em / (5.0 - 10.0 * nu)
    bet = dl * dl * dl * Em * (4 * nu - 1) / (20.0 * (2 * nu * nu + nu - 1))
    return alf, bet


class OracleCoupled:
    """L1's one-way coupled model: the LBM oracle, and the solid lattice of
    the same ind (corners of solid voxels, periodic in x, y and z) moved by
    the fluid force of every main-loop step (L1:3849-3888)."""

    def __init__(self, inp: pio.LBMInput, ind: np.ndarray, S=None, cube=None):
        self.lbm = OracleLBM(inp, ind, S)
        self.L = self.lbm.L
        self.inp = inp
        nz, ny, nx = ind.shape
        alf, bet = coupled_springs(inp.Em, inp.nu, inp.dl)
        p = LSMParams()
        p.nx, p.ny, p.nz, p.pad_axis, p.jump_comp, p.nthread = nx, ny, nz, -1, 0, 1
        p.alf1 = p.alf2 = alf
        p.bet1 = p.bet2 = bet
        p.rs, p.dt, p.gstrain = float(inp.rs), float(inp.dt), 0.0
        self.p = p
        self._raw = np.ascontiguousarray(ind, np.int32)
        self._ind = np.ascontiguousarray(np.where(ind == 0, 0, 1), np.int32)
        cb, ca, cs, cab = cube if cube is not None else geo.cube_tables()
        self._cube = [np.ascontiguousarray(x) for x in (cb, ca, cs, cab)]
        self.h = self.L.or_lsm_create(C.byref(p), self._ind, *self._cube)
        self.ns = self.L.or_lsm_count(self.h, 0)
        assert self.ns == self.lbm.counts()["ns"]
        self.af = np.zeros((self.ns, 3))

    def __del__(self):
        if getattr(self, "h", None):
            self.L.or_lsm_destroy(self.h)
            self.h = None

    def step(self, n=1):
        """n main-loop steps: LBM step, fluid force, lsm_step."""
        rsi = 1.0 / float(self.inp.rs)
        for _ in range(int(n)):
            self.lbm.step(1)
            self.L.or_lbm_fluid_force(self.lbm.h, rsi, self.af)
            self.L.or_lsm_coupled_step(self.h, self.af, 0.25)

    def state(self):
        """r, v, a, af as (ns, 3) in the L1 solid numbering."""
        r, v, a = (np.zeros((self.ns, 3)) for _ in range(3))
        self.L.or_lsm_get_state(self.h, _ptr(r), _ptr(v), _ptr(a))
        return r, v, a, self.af.copy()

    def set_state(self, r, v, a):
        self._st = [np.ascontiguousarray(x, np.float64) for x in (r, v, a)]
        self.L.or_lsm_set_state(self.h, *(_ptr(x) for x in self._st))

    def mijks(self):
        nz, ny, nx = self._ind.shape
        icors = np.zeros((self.ns, 3), np.int32)
        mijks = np.zeros((nz, ny, nx), np.int32)
        self.L.or_lsm_get_nodes(self.h, _ptr(icors), _ptr(mijks))
        return mijks

    def calc_ab(self, pres=None):
        """calc_ab: dict of asigmap, asigmam, astrain (3,3) Fortran order
        (row = force component), epsi, alphai, ncub, betai (npores,)."""
        inp = self.inp
        if pres is None:
            pres = (1.0 - 0.99) * inp.rr0 * inp.cr2  # L1:3987
        out = np.zeros(38 + int(inp.npores))
        self.L.or_lsm_calc_ab(self.h, self._raw, int(inp.iipore), int(inp.npores), float(pres), out)
        t = lambda a: a.reshape(3, 3).T  # column-major -> [row][col]
        return {"asigmap": t(out[0:9]), "asigmam": t(out[9:18]), "astrain": t(out[18:27]), "epsi": out[27],
                "alphai": t(out[28:37]), "ncub": int(out[37]), "betai": out[38:], "pres": pres}


def read_lsm_files(d: str, it1: int, shape) -> dict:
    """The solid-lattice files of save_results (L1:2176-2207) as arrays over
    the voxel grid (nz, ny, nx, 3): disp (LSM_disf<it1>), vel, force
    (adat/rsi), fluid_force (afdat/rsi); zero rows where no solid node."""
    out = {}
    for key, name in (("disp", f"LSM_disf{it1}"), ("vel", "LSM_velocity_field"), ("force", "LSM_force_field"),
                      ("fluid_force", "LSM_fluid_force")):
        with open(os.path.join(d, name)) as fh:
            out[key] = pio.parse_fortran_numbers(fh.read()).reshape(tuple(shape) + (3,))
    return out


def read_alpha_beta(path: str, npores: int) -> dict:
    """Last block appended to alpha_beta by calc_ab (L1:4447-4466)."""
    with open(path) as fh:
        lines = fh.read().splitlines()
    start = max(i for i, l in enumerate(lines) if l.strip().startswith("Elastic medium properties"))
    block = lines[start:]

    def after(label, count):
        i = next(k for k, l in enumerate(block) if l.strip().startswith(label))
        rest = block[i].split("=", 1)[1]
        vals = list(pio.parse_fortran_numbers(rest)) if rest.strip() else []
        k = i + 1
        while len(vals) < count:
            vals += list(pio.parse_fortran_numbers(block[k]))
            k += 1
        return np.array(vals[:count])

    t = lambda a: a.reshape(3, 3).T
    return {"asigmap": t(after("asigmap", 9)), "asigmam": t(after("asigmam", 9)),
            "astrain": t(after("astrain", 9)), "epsi": after("epsi", 1)[0], "alphai": t(after("alphai", 9)),
            "betai": after("betai", npores), "ncub": int(after("Solid part volume", 1)[0]),
            "pres": after("pres", 1)[0]}
