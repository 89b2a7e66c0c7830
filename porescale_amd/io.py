"""Readers and writers for the reference's file formats (the drop-in API surface).

Formats follow the Fortran programs' list-directed and fixed-format I/O:

* ``INPUT_PARAMETERS`` of the two-phase LBM program
  (``LBM-LSM/lsm_lbm_static_alpha_beta_multipore.f:435-469``) and of the LSM
  programs (``LSM2S/Cxxxx/XX_P.f:364-382``): one positional value per line,
  the rest of each line is a comment, trailing lines are ignored.
* ``ind``: one record per (j, k), k outer, j inner, each record a blank and
  ``ncx`` single digits (``lsm_lbm_static_alpha_beta_multipore.f:532-538``).
* ``iteration2`` restart files: LBM ``it1, fr(19,nl), fb(19,nl)``
  (``lsm_lbm_static_alpha_beta_multipore.f:3951-3961, 2845-2850``); LSM
  ``r, v, a (3,ns)`` (``XX_P.f:1599-1609, 1288-1294``).
* ``output.m`` rows (``lsm_lbm_static_alpha_beta_multipore.f:3274-3282``).
* ``velocities19``/``velocities9`` (``format(19i3)``/``(9i3)``) and the
  ``cube_*`` element tables read by ``calc_stress`` (``XX_P.f:1700-1714``).
"""
from __future__ import annotations

import os
import re
from dataclasses import dataclass, field, fields
from typing import List, Sequence

import numpy as np

__all__ = [
    "LBMInput", "LSMInput", "read_lbm_input", "write_lbm_input", "read_lsm_input",
    "write_lsm_input", "read_ind", "write_ind", "read_lbm_restart", "write_lbm_restart",
    "read_lsm_restart", "write_lsm_restart", "read_output_m", "read_velocities",
    "write_velocities", "read_cube_tables", "write_cube_tables", "parse_fortran_numbers",
    "D3Q19", "BOND9",
]

# D3Q19 order of the reference's velocities19 file, and the 9 bond vectors of
# velocities9.  Population indices in every state array follow this order.
D3Q19 = np.array([
    [0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0],
    [0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1],
    [0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1],
], dtype=np.int32)
BOND9 = np.array([
    [1, 0, 0, 1, 1, 1, 1, 0, 0],
    [0, 1, 0, 1, -1, 0, 0, 1, 1],
    [0, 0, 1, 0, 0, 1, -1, 1, -1],
], dtype=np.int32)

_NUM = re.compile(r"^[\s]*([-+0-9.eEdD]+)")


def _fnum(tok: str) -> float:
    return float(tok.replace("D", "E").replace("d", "e"))


def parse_fortran_numbers(text: str) -> np.ndarray:
    """Parse list-directed Fortran output (handles D exponents and r*c repeats)."""
    out: List[float] = []
    for tok in text.replace(",", " ").split():
        if "*" in tok:
            rep, val = tok.split("*", 1)
            out.extend([_fnum(val)] * int(rep))
        else:
            out.append(_fnum(tok))
    return np.asarray(out, dtype=np.float64)


def _values(path: str, count: int) -> List[str]:
    """First whitespace token of the first ``count`` lines that carry one."""
    vals: List[str] = []
    with open(path) as fh:
        for line in fh:
            if len(vals) == count:
                break
            s = line.strip()
            if not s:
                continue
            vals.append(s)
    if len(vals) < count:
        raise ValueError(f"{path}: expected {count} parameter lines, found {len(vals)}")
    return vals


def _first(s: str) -> str:
    return s.split()[0]


@dataclass
class LBMInput:
    """INPUT_PARAMETERS of the two-phase LBM program (30 positional lines)."""
    cr2: float = 1.0 / 3.0
    rr0: float = 1.0
    rb0: float = 1.0
    B: float = 0.7
    sig: float = 0.01
    vir: float = 1.0 / 6.0
    vib: float = 1.0 / 6.0
    wet: float = 0.0
    idir: int = 3
    itm: int = 1
    itw: int = 10
    its: int = 1
    eps: float = -1e-7
    mtype: int = 0
    icon: int = 0
    sat: float = 0.0
    iipore: int = 1
    npores: int = 1
    ee: float = 0.0
    lsize: float = 1.0
    bfr: Sequence[float] = field(default_factory=lambda: [0.0, 0.0, 1e-5])
    bfb: Sequence[float] = field(default_factory=lambda: [0.0, 0.0, 1e-5])
    ncx: int = 16
    ncy: int = 16
    ncz: int = 16
    Em: float = 94.5
    nu: float = 0.0742
    rs: float = 2.65
    dt: float = 1.0
    dl: float = 1.0


def read_lbm_input(path: str) -> LBMInput:
    v = _values(path, 30)
    p = LBMInput()
    scal = ["cr2", "rr0", "rb0", "B", "sig", "vir", "vib", "wet", "idir", "itm", "itw", "its",
            "eps", "mtype", "icon", "sat", "iipore", "npores", "ee", "lsize"]
    for name, s in zip(scal, v[:20]):
        cur = getattr(p, name)
        setattr(p, name, int(_first(s)) if isinstance(cur, int) else _fnum(_first(s)))
    p.bfr = [_fnum(t) for t in v[20].split()[:3]]
    p.bfb = [_fnum(t) for t in v[21].split()[:3]]
    p.ncx, p.ncy, p.ncz = (int(_first(s)) for s in v[22:25])
    p.Em, p.nu, p.rs, p.dt, p.dl = (_fnum(_first(s)) for s in v[25:30])
    return p


def _fd(x: float) -> str:
    return repr(float(x)).replace("e", "D") if "e" in repr(float(x)) else repr(float(x)) + "D0"


def write_lbm_input(path: str, p: LBMInput) -> None:
    names = ["cr2", "rr0", "rb0", "B", "sig", "vir", "vib", "wet", "idir", "itm", "itw", "its",
             "eps", "mtype", "icon", "sat", "iipore", "npores", "ee", "lsize"]
    lines = []
    for n in names:
        val = getattr(p, n)
        lines.append(f"{str(val) if isinstance(val, int) else _fd(val):<24s}! {n}")
    lines.append(" ".join(_fd(x) for x in p.bfr) + "   ! bfr")
    lines.append(" ".join(_fd(x) for x in p.bfb) + "   ! bfb")
    for n in ["ncx", "ncy", "ncz"]:
        lines.append(f"{getattr(p, n):<24d}! {n}")
    for n in ["Em", "nu", "rs", "dt", "dl"]:
        lines.append(f"{_fd(getattr(p, n)):<24s}! {n}")
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


@dataclass
class LSMInput:
    """INPUT_PARAMETERS of the two-solid LSM programs (15 positional lines)."""
    ncx: int = 16
    ncy: int = 16
    ncz: int = 16
    Em1: float = 1.0
    nu1: float = 0.25
    Em2: float = 1.0
    nu2: float = 0.25
    rs: float = 1.0
    dt: float = 0.01
    dl: float = 1.0
    itm: int = 1
    its: int = 100
    gstrain: float = 0.01
    nthread: int = 1
    icon: int = 0


def read_lsm_input(path: str) -> LSMInput:
    v = _values(path, 15)
    p = LSMInput()
    for f_, s in zip(fields(LSMInput), v):
        setattr(p, f_.name, int(_first(s)) if f_.type in (int, "int") else _fnum(_first(s)))
    return p


def write_lsm_input(path: str, p: LSMInput) -> None:
    lines = []
    for f_ in fields(LSMInput):
        val = getattr(p, f_.name)
        lines.append(f"{str(val) if isinstance(val, int) else _fd(val):<24s}! {f_.name}")
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def read_ind(path: str, ncx: int, ncy: int, ncz: int) -> np.ndarray:
    """Return codes as an int32 array indexed [k, j, i] (C order = Fortran ind(i,j,k))."""
    out = np.empty((ncz, ncy, ncx), dtype=np.int32)
    with open(path) as fh:
        for k in range(ncz):
            for j in range(ncy):
                line = fh.readline()
                if not line:
                    raise ValueError(f"{path}: truncated at record j={j + 1} k={k + 1}")
                digits = line.rstrip("\r\n")[1:1 + ncx]
                if len(digits) != ncx:
                    raise ValueError(f"{path}: record j={j + 1} k={k + 1} has {len(digits)} digits")
                out[k, j, :] = np.frombuffer(digits.encode(), dtype=np.uint8) - 48
    return out


def write_ind(path: str, ind: np.ndarray) -> None:
    ind = np.asarray(ind)
    if ind.min() < 0 or ind.max() > 9:
        raise ValueError("ind codes must be single digits")
    nz, ny, nx = ind.shape
    rows = (ind.reshape(nz * ny, nx).astype(np.uint8) + 48)
    with open(path, "w") as fh:
        for r in rows:
            fh.write(" " + r.tobytes().decode() + "\n")


def read_lbm_restart(path: str, nl: int):
    with open(path) as fh:
        it1 = int(fh.readline().split()[0])
        vals = parse_fortran_numbers(fh.read())
    if vals.size != 38 * nl:
        raise ValueError(f"{path}: expected {38 * nl} values, got {vals.size}")
    fr = vals[:19 * nl].reshape(nl, 19)
    fb = vals[19 * nl:].reshape(nl, 19)
    return it1, fr, fb


def _write_vals(fh, arr: np.ndarray, per_line: int = 4) -> None:
    flat = np.asarray(arr, dtype=np.float64).ravel()
    for s in range(0, flat.size, per_line):
        fh.write(" " + " ".join(repr(float(x)) for x in flat[s:s + per_line]) + "\n")


def write_lbm_restart(path: str, it1: int, fr: np.ndarray, fb: np.ndarray) -> None:
    """fr, fb shaped (nl, 19), i.e. Fortran (19, nl) column-major."""
    with open(path, "w") as fh:
        fh.write(f" {int(it1)}\n")
        _write_vals(fh, fr)
        _write_vals(fh, fb)


def read_lsm_restart(path: str, ns: int):
    with open(path) as fh:
        vals = parse_fortran_numbers(fh.read())
    if vals.size != 9 * ns:
        raise ValueError(f"{path}: expected {9 * ns} values, got {vals.size}")
    return (vals[:3 * ns].reshape(ns, 3), vals[3 * ns:6 * ns].reshape(ns, 3),
            vals[6 * ns:].reshape(ns, 3))


def write_lsm_restart(path: str, r: np.ndarray, v: np.ndarray, a: np.ndarray) -> None:
    with open(path, "w") as fh:
        _write_vals(fh, r)
        _write_vals(fh, v)
        _write_vals(fh, a)


def read_output_m(path: str) -> np.ndarray:
    rows = []
    with open(path) as fh:
        for line in fh:
            s = line.strip()
            if not s or s.startswith("%") or s.startswith("out") or s.startswith("]"):
                continue
            rows.append(parse_fortran_numbers(s.rstrip(";")))
    return np.array(rows)


def read_ind_red_blue(path: str, ncx: int, ncy: int, ncz: int):
    """Blocks of the phase file save_results appends to
    (lsm_lbm_static_alpha_beta_multipore.f:2143-2155): a list of
    (iteration, codes (ncz, ncy, ncx) int32), in file order."""
    blocks = []
    with open(path) as fh:
        lines = fh.read().splitlines()
    k = 0
    while k < len(lines):
        s = lines[k].strip()
        if s.startswith("ITER="):
            it = int(s[5:].rstrip(";"))
            if lines[k + 1].strip() != "a=[":
                raise ValueError(f"{path}: line {k + 2}: expected 'a=['")
            rows = lines[k + 2:k + 2 + ncy * ncz]
            if len(rows) != ncy * ncz:
                raise ValueError(f"{path}: block ITER={it} is truncated")
            a = np.array([[int(r[2 * i:2 * i + 2]) for i in range(ncx)] for r in rows], dtype=np.int32)
            blocks.append((it, a.reshape(ncz, ncy, ncx)))
            k += 2 + ncy * ncz
        else:
            k += 1
    return blocks


def write_save_fields(directory: str, it1: int, itw: int, fields: dict) -> None:
    """The fluid files of save_results (lsm_lbm_static_alpha_beta_multipore.f:
    2067-2155) from the arrays ``LBM.save_fields`` returns: velocity_field
    (3 values a line), color_field, pressure<it1> (one value a line, voxels x
    fastest) and one block appended to ind_red_blue."""
    u = np.asarray(fields["velocity"]).reshape(-1, 3)
    with open(os.path.join(directory, "velocity_field"), "w") as fh:
        for row in u:
            fh.write(" " + " ".join(repr(float(x)) for x in row) + "\n")
    for name, key in (("color_field", "color"), (f"pressure{int(it1)}", "pressure")):
        with open(os.path.join(directory, name), "w") as fh:
            for x in np.asarray(fields[key]).reshape(-1):
                fh.write(f" {float(x)!r}\n")
    ph = np.asarray(fields["phase"])
    ncz, ncy, ncx = ph.shape
    with open(os.path.join(directory, "ind_red_blue"), "a") as fh:
        fh.write(f"\n\n\n     ITER={int(it1) * int(itw):7d};\n a=[\n")
        for kk in range(ncz):
            for jj in range(ncy):
                fh.write("".join(f"{int(x):2d}" for x in ph[kk, jj]) + "\n")
        fh.write(" ];\n clm\n")


def read_velocities(path: str, n: int) -> np.ndarray:
    with open(path) as fh:
        rows = [[int(line[3 * c:3 * c + 3]) for c in range(n)] for line in fh.readlines()[:3]]
    return np.array(rows, dtype=np.int32)


def write_velocities(path: str, table: np.ndarray) -> None:
    with open(path, "w") as fh:
        for row in np.asarray(table):
            fh.write("".join(f"{int(x):3d}" for x in row) + "\n")


def read_cube_tables(directory: str):
    """Return (cube_bon (24,3), cube_ang (72,6), cube_sbond (24,), cube_abond (72,))."""
    def nums(name):
        with open(os.path.join(directory, name)) as fh:
            return parse_fortran_numbers(fh.read())
    bon = nums("cube_bon").astype(np.int32).reshape(24, 3)
    ang = nums("cube_ang").astype(np.int32).reshape(72, 6)
    sb = nums("cube_sbond").reshape(24)
    ab = nums("cube_abond").reshape(72)
    return bon, ang, sb, ab


def write_cube_tables(directory: str, bon, ang, sb, ab) -> None:
    with open(os.path.join(directory, "cube_bon"), "w") as fh:
        for r in np.asarray(bon):
            fh.write("".join(f"{int(x):13d}" for x in r) + "\n")
    with open(os.path.join(directory, "cube_ang"), "w") as fh:
        for r in np.asarray(ang):
            fh.write("".join(f"{int(x):13d}" for x in r) + "\n")
    with open(os.path.join(directory, "cube_sbond"), "w") as fh:
        for x in np.asarray(sb):
            fh.write(f"   {float(x):.16f}     \n")
    with open(os.path.join(directory, "cube_abond"), "w") as fh:
        for x in np.asarray(ab):
            fh.write(f"   {float(x):.16f}     \n")
