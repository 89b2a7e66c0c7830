"""Lattice spring model on the GPU (host-side interface).

Mirrors the reference programs ``latticespringmodel``
(``LSM2S/Cxxxx/XX_P.f``, ``LSM2S/Cyzyz/YZ_P.f``): voxel codes (0 = solid 1,
3 = solid 2, else pore), spring constants from (Em, nu) as XX:411-414, the
damped-Verlet solver of XX:1377-1545 and calc_stress (XX:1628-1997); plus an
equilibrium solver, the six-case stiffness tensor, batched sub-volumes and
the adjoint of the moduli.

Displacements live on the periodic node grid (node (x,y,z) = lower corner of
voxel (x,y,z)), shaped (3, nz, ny, nx).  A load case imposes the jump
J = 2*gstrain*e_d across the period of axis p; ``reference_node_map``
converts to and from the reference's padded node numbering.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np

from . import _lib
from .io import LSMInput

__all__ = ["LSM", "spring_constants", "moduli_from_C", "reference_node_map", "batch_moduli",
           "CASES"]

# load case -> (padded axis p, jump component d); xx = XX_P.f, yz = YZ_P.f
CASES = {"xx": (0, 0), "yy": (1, 1), "zz": (2, 2), "yz": (1, 2), "zx": (2, 0), "xy": (0, 1)}
_CASE_ID = {k: i for i, k in enumerate(_lib.LOAD_CASES)}


def spring_constants(Em: float, nu: float, dl: float = 1.0):
    """alpha, beta from Young's modulus and Poisson's ratio (XX_P.f:411-414)."""
    alf = dl * Em / (5.0 - 10.0 * nu)
    bet = dl * dl * dl * Em * (4 * nu - 1) / (20.0 * (2 * nu * nu + nu - 1))
    return alf, bet


def moduli_from_C(C6: np.ndarray):
    """Voigt bulk and shear moduli of a 6x6 stiffness matrix."""
    a = C6[0, 0] + C6[1, 1] + C6[2, 2]
    b = C6[0, 1] + C6[0, 2] + C6[1, 2] + C6[1, 0] + C6[2, 0] + C6[2, 1]
    c = C6[3, 3] + C6[4, 4] + C6[5, 5]
    return (a + b) / 9.0, (a - 0.5 * b + 3.0 * c) / 15.0


def _case(c) -> int:
    if isinstance(c, str):
        return _CASE_ID[c]
    return int(c)


def reference_node_map(ind: np.ndarray, case: str):
    """Node numbering of XX_P.f/YZ_P.f (XX:474-578) for load case ``case``.

    Returns (grid, twin, G): for every reference node (icors order, z outer,
    x fastest on the grid padded by 2 along the load axis) its flat index on
    the periodic node grid, whether it is the far twin (displacement = grid
    value + J), and the padded grid dims (Gx, Gy, Gz).
    """
    pa = CASES[case][0]
    solid = (ind == 0) | (ind == 3)             # [z, y, x]
    nz, ny, nx = ind.shape
    n = [nx, ny, nz]
    G = [n[a] + (2 if a == pa else 0) for a in range(3)]
    # node (i,j,k) on the padded grid (0-based) is solid if any of its up to
    # 8 adjacent voxels is solid; along the padded axis voxel index i-1 and i
    # must lie in 0..n-1, along periodic axes they wrap.
    node = np.zeros((G[2], G[1], G[0]), bool)
    for c in range(8):
        off = [c & 1, (c >> 1) & 1, c >> 2]
        idx = []
        valid = np.ones((G[2], G[1], G[0]), bool)
        for a in range(3):
            g = np.arange(G[a]) - off[a]
            if a == pa:
                ok = (g >= 0) & (g < n[a])
                g = np.where(ok, g, 0)
            else:
                ok = np.ones(G[a], bool)
                g = np.mod(g, n[a])
            shape = [1, 1, 1]
            shape[2 - a] = G[a]
            idx.append(g.reshape(shape))
            valid &= ok.reshape(shape)
        node |= solid[idx[2], idx[1], idx[0]] & valid
    kz, ky, kx = np.nonzero(node)              # C order = z outer, x fastest
    coords = [kx, ky, kz]
    gc = [np.mod(coords[a], n[a]) for a in range(3)]
    twin = coords[pa] == n[pa]
    grid = gc[0] + nx * (gc[1] + ny * gc[2])
    return grid.astype(np.int64), twin, tuple(G)


class LSM:
    """One LSM volume on the GPU.  ``ind`` is indexed [k, j, i]."""

    def __init__(self, ind: np.ndarray, alf1: float, bet1: float, alf2: Optional[float] = None,
                 bet2: Optional[float] = None, rs: float = 1.0, dt: float = 0.01, gstrain: float = 0.01,
                 precision: str = "f64", stiff: Optional[np.ndarray] = None, cube=None):
        L = _lib.lib()
        self.shape = ind.shape
        nz, ny, nx = ind.shape
        p = _lib.LSMParams()
        p.alf1, p.bet1 = float(alf1), float(bet1)
        p.alf2 = float(alf1 if alf2 is None else alf2)
        p.bet2 = float(bet1 if bet2 is None else bet2)
        p.rs, p.dt, p.gstrain = float(rs), float(dt), float(gstrain)
        p.precision = {"f64": _lib.PREC_F64, "f32": _lib.PREC_F32}[precision]
        self.params = p
        self.precision = precision
        self._ind = _lib.i32(ind)
        self._stiff = None if stiff is None else _lib.f64(stiff).reshape(-1)
        cb = [None] * 4 if cube is None else [_lib.i32(cube[0]), _lib.i32(cube[1]), _lib.f64(cube[2]),
                                              _lib.f64(cube[3])]
        h = C.c_void_p()
        _lib.check(L.psx_lsm_create(C.byref(p), nx, ny, nz, _lib.ptr(self._ind), _lib.ptr(self._stiff),
                                    *(_lib.ptr(x) for x in cb), C.byref(h)))
        self._h, self._L = h, L
        c = np.zeros(4, np.int64)
        _lib.check(L.psx_lsm_counts(h, _lib.ptr(c)))
        self.nsolid_nodes, self.ncubes, self.N = int(c[0]), int(c[1]), int(c[2])

    def set_springs(self, alf1: float, bet1: float, alf2: Optional[float] = None, bet2: Optional[float] = None):
        """New spring constants for the same geometry; the handle, its node
        grid and multigrid hierarchy are kept and refreshed in place."""
        a2 = alf1 if alf2 is None else alf2
        b2 = bet1 if bet2 is None else bet2
        _lib.check(self._L.psx_lsm_set_springs(self._h, float(alf1), float(bet1), float(a2), float(b2)))
        self.params.alf1, self.params.bet1, self.params.alf2, self.params.bet2 = alf1, bet1, a2, b2

    @classmethod
    def from_input(cls, inp: LSMInput, ind: np.ndarray, **kw):
        a1, b1 = spring_constants(inp.Em1, inp.nu1, inp.dl)
        a2, b2 = spring_constants(inp.Em2, inp.nu2, inp.dl)
        return cls(ind, a1, b1, a2, b2, rs=inp.rs, dt=inp.dt, gstrain=inp.gstrain, **kw)

    def close(self):
        if getattr(self, "_h", None):
            self._L.psx_lsm_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _vec(self, u):
        u = _lib.f64(u)
        if u.size != 3 * self.N:
            raise ValueError(f"displacement must have 3*{self.N} values")
        return u.reshape(3, -1)

    def set_u(self, case, u):
        _lib.check(self._L.psx_lsm_set_u(self._h, _case(case), _lib.ptr(self._vec(u))))

    def get_u(self, case) -> np.ndarray:
        u = np.zeros((3,) + tuple(self.shape))
        _lib.check(self._L.psx_lsm_get_u(self._h, _case(case), _lib.ptr(u)))
        return u

    def forces(self, case, u) -> np.ndarray:
        """F = -K(u + jump) on the node grid (the force loops of XX:1404-1500)."""
        f = np.zeros((3,) + tuple(self.shape))
        _lib.check(self._L.psx_lsm_forces(self._h, _case(case), _lib.ptr(self._vec(u)), _lib.ptr(f)))
        return f

    def verlet(self, case, nsteps: int, restart: bool = True):
        """nsteps of the reference's damped velocity-Verlet loop."""
        _lib.check(self._L.psx_lsm_verlet(self._h, _case(case), int(nsteps), 1 if restart else 0))

    def reference_nodes(self, case, with_icors: bool = False):
        """(grid index, far-twin flag[, icors (ns, 3) 1-based on the padded
        grid]) of every node of the reference's node list for ``case``
        (XX:474-578)."""
        n = C.c_int64(0)
        _lib.check(self._L.psx_lsm_reference_nodes(self._h, _case(case), C.byref(n), None, None, None))
        grid = np.zeros(n.value, np.int64)
        twin = np.zeros(n.value, np.int8)
        icors = np.zeros((n.value, 3), np.int32)
        _lib.check(self._L.psx_lsm_reference_nodes(self._h, _case(case), C.byref(n), _lib.ptr(grid), _lib.ptr(twin),
                                                   _lib.ptr(icors)))
        return (grid, twin.astype(bool), icors) if with_icors else (grid, twin.astype(bool))

    def get_verlet(self, case):
        """rdat, vdat, adat (ns, 3) of the Verlet state, reference node order
        (what XX writes to iteration2)."""
        ns = len(self.reference_nodes(case)[0])
        r, v, a = (np.zeros((ns, 3)) for _ in range(3))
        _lib.check(self._L.psx_lsm_get_verlet(self._h, _case(case), *(_lib.ptr(x) for x in (r, v, a))))
        return r, v, a

    def set_verlet(self, case, r, v, a):
        """Restart the Verlet loop from an iteration2 state (icon = 1); the
        glued pairs restart at -+gstrain as in XX:1366-1375."""
        arrs = [np.ascontiguousarray(x, np.float64) for x in (r, v, a)]
        _lib.check(self._L.psx_lsm_set_verlet(self._h, _case(case), *(_lib.ptr(x) for x in arrs)))

    def solve(self, case, rtol: float = 1e-10, maxit: int = 20000):
        it, rel = C.c_int32(0), C.c_double(0)
        _lib.check(self._L.psx_lsm_solve(self._h, _case(case), float(rtol), int(maxit), C.byref(it), C.byref(rel)))
        return it.value, rel.value

    def calc_stress(self, case):
        """(asigmap (3,3) summed over cubes, moduli (3,3) = asigmap/(V*2g/L))."""
        s, m = np.zeros(9), np.zeros(9)
        _lib.check(self._L.psx_lsm_stress(self._h, _case(case), _lib.ptr(s), _lib.ptr(m)))
        return s.reshape(3, 3, order="F"), m.reshape(3, 3, order="F")

    def moduli(self, rtol: float = 1e-10, maxit: int = 20000):
        """6x6 Voigt stiffness (rows xx,yy,zz,yz,zx,xy; columns load cases), K, G."""
        Cm = np.zeros(36)
        K, G = C.c_double(0), C.c_double(0)
        _lib.check(self._L.psx_lsm_moduli(self._h, float(rtol), int(maxit), _lib.ptr(Cm), C.byref(K), C.byref(G)))
        return Cm.reshape(6, 6, order="F"), K.value, G.value

    def adjoint(self, wK: float = 1.0, wG: float = 0.0, rtol: float = 1e-10, maxit: int = 20000,
                want_stiff: bool = True):
        """Gradient of wK*K + wG*G: d/d(alf1, bet1, alf2, bet2) and per-voxel
        d/d(stiffness multiplier) (shape of ind)."""
        if not getattr(self, "_h", None):
            raise ValueError("LSM handle is closed")
        dab = np.zeros(4)
        ds = np.zeros(self.shape) if want_stiff else None
        _lib.check(self._L.psx_lsm_adjoint(self._h, float(wK), float(wG), float(rtol), int(maxit),
                                           _lib.ptr(dab), _lib.ptr(ds)))
        return dab, ds

    def profile(self, n: int = 20):
        """Mean ms of one FP64 and one FP32 operator application."""
        ms = np.zeros(2)
        _lib.check(self._L.psx_lsm_profile(self._h, int(n), _lib.ptr(ms)))
        return tuple(ms)


def table_counts(ind: np.ndarray, case, nthread: int = 1) -> dict:
    """Sizes of the reference's bond and angle tables for ``case`` and its
    z-layer partition (the lsm_settings entries nbo, nna, layers, layb, lib,
    lia of XX:1314-1348), from the voxel codes ``ind`` [z, y, x] alone."""
    L = _lib.lib()
    nz, ny, nx = ind.shape
    nth2 = 2 * int(nthread)
    counts = np.zeros(3, np.int64)
    layers = np.zeros(nth2, np.int32)
    layb, lib_, lia = (np.zeros(nth2 + 1, np.int64) for _ in range(3))
    _lib.check(L.psx_lsm_table_counts(nx, ny, nz, _lib.ptr(_lib.i32(ind).reshape(-1)), _case(case), int(nthread),
                                      _lib.ptr(counts), _lib.ptr(layers), _lib.ptr(layb), _lib.ptr(lib_),
                                      _lib.ptr(lia)))
    return {"nbo": int(counts[0]), "nna": int(counts[1]), "ns": int(counts[2]), "layers": layers, "layb": layb,
            "lib": lib_, "lia": lia}


def batch_moduli(volumes: Sequence[np.ndarray], alf1: float, bet1: float, alf2: float, bet2: float,
                 gstrain: float = 0.01, precision: str = "f32", rtol: float = 1e-10, maxit: int = 20000):
    """Moduli of many sub-volumes, all six load cases solved together.

    Returns (C (V,6,6), KG (V,2), elapsed ms, iterations)."""
    L = _lib.lib()
    p = _lib.LSMParams()
    p.alf1, p.bet1, p.alf2, p.bet2 = float(alf1), float(bet1), float(alf2), float(bet2)
    p.rs, p.dt, p.gstrain = 1.0, 0.01, float(gstrain)
    p.precision = {"f64": _lib.PREC_F64, "f32": _lib.PREC_F32}[precision]
    dims = np.array([[v.shape[2], v.shape[1], v.shape[0]] for v in volumes], np.int32)
    cat = np.concatenate([_lib.i32(v).reshape(-1) for v in volumes])
    Cm = np.zeros(36 * len(volumes))
    KG = np.zeros(2 * len(volumes))
    ms = np.zeros(2)
    _lib.check(L.psx_lsm_batch_moduli(C.byref(p), len(volumes), _lib.ptr(dims), _lib.ptr(cat), float(rtol),
                                      int(maxit), _lib.ptr(Cm), _lib.ptr(KG), _lib.ptr(ms)))
    Cv = Cm.reshape(len(volumes), 6, 6).transpose(0, 2, 1)
    return Cv, KG.reshape(-1, 2), ms[0], int(ms[1])
