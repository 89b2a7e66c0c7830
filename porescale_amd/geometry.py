"""Seeded synthetic inputs: periodic sphere packs, phase blobs, lattice tables.

Everything here is deterministic on any machine: the random stream is a
counter-based splitmix64 hash written in numpy (no library RNG), so the same
``seed`` always yields the same ``ind`` file.
"""
from __future__ import annotations

import numpy as np

from .io import BOND9, D3Q19

__all__ = ["uniform", "sphere_pack", "two_component", "phase_blobs", "cube_tables",
           "angle_types"]

_M = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix64(x: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = (x + np.uint64(0x9E3779B97F4A7C15)) & _M
        z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M
        z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M
        return z ^ (z >> np.uint64(31))


def uniform(seed: int, stream: int, n: int, start: int = 0) -> np.ndarray:
    """``n`` doubles in [0,1) from counter ``start..start+n-1`` of ``(seed, stream)``."""
    with np.errstate(over="ignore"):
        base = np.uint64((seed * 0x100000001B3 + stream * 0x9E3779B1) & 0xFFFFFFFFFFFFFFFF)
        ctr = np.arange(start, start + n, dtype=np.uint64) + base
    return (_splitmix64(ctr) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def _stamp(solid: np.ndarray, centres: np.ndarray, radii: np.ndarray) -> None:
    """Mark voxels whose centre lies within radius of a sphere centre (periodic)."""
    nz, ny, nx = solid.shape
    for r in np.unique(radii):
        sel = centres[radii == r]
        ir = int(np.ceil(r))
        g = np.arange(-ir, ir + 1)
        dz, dy, dx = np.meshgrid(g, g, g, indexing="ij")
        # voxel (i,j,k) has centre (i+0.5, ...); centres are continuous
        for s in range(0, len(sel), 512):
            c = sel[s:s + 512]
            base = np.floor(c).astype(np.int64)
            vx = base[:, None, 0] + dx.ravel()[None, :]
            vy = base[:, None, 1] + dy.ravel()[None, :]
            vz = base[:, None, 2] + dz.ravel()[None, :]
            d2 = ((vx + 0.5 - c[:, None, 0]) ** 2 + (vy + 0.5 - c[:, None, 1]) ** 2
                  + (vz + 0.5 - c[:, None, 2]) ** 2)
            m = d2 <= r * r
            solid[vz[m] % nz, vy[m] % ny, vx[m] % nx] = True


def _ball_index(shape, c, r: float) -> np.ndarray:
    """Flat indices of the voxels one ball marks (same rule as ``_stamp``)."""
    nz, ny, nx = shape
    ir = int(np.ceil(r))
    g = np.arange(-ir, ir + 1)
    dz, dy, dx = np.meshgrid(g, g, g, indexing="ij")
    base = np.floor(c).astype(np.int64)
    vx = base[0] + dx.ravel()
    vy = base[1] + dy.ravel()
    vz = base[2] + dz.ravel()
    d2 = (vx + 0.5 - c[0]) ** 2 + (vy + 0.5 - c[1]) ** 2 + (vz + 0.5 - c[2]) ** 2
    m = d2 <= r * r
    return np.unique(((vz[m] % nz) * ny + vy[m] % ny) * nx + vx[m] % nx)


def _grow_balls(mask: np.ndarray, region: np.ndarray, target: float, seed: int, stream: int,
                radius: float) -> None:
    """Add seeded balls (centres from counter stream ``stream``) to ``mask``,
    one at a time, while fewer than ``target`` voxels of ``region`` are marked."""
    nz, ny, nx = mask.shape
    fm, fr = mask.reshape(-1), region.reshape(-1)
    count = int(np.count_nonzero(mask & region))
    ctr, chunk = 0, 1024
    while count < target:
        u = uniform(seed, stream, 3 * chunk, start=3 * ctr)
        for b in range(chunk):
            if count >= target:
                break
            c = np.array([u[3 * b] * nx, u[3 * b + 1] * ny, u[3 * b + 2] * nz])
            idx = _ball_index(mask.shape, c, radius)
            new = idx[~fm[idx]]
            fm[new] = True
            count += int(np.count_nonzero(fr[new]))
        ctr += chunk


def sphere_pack(shape, porosity: float, radius=6.0, seed: int = 0, invert: bool = False,
                batch: int = 64) -> np.ndarray:
    """Periodic pack of overlapping solid spheres with uniform centres.

    Spheres are added in batches of ``batch`` (radius fixed, or drawn uniformly
    from ``radius=(rmin, rmax)``) until the pore fraction drops to ``porosity``;
    the last batch is trimmed sphere by sphere so the target is met closely.
    Returns ``ind`` codes indexed [k, j, i]: 1 = pore, 0 = solid.  With
    ``invert`` the spheres are pore and the rest solid (same realisation).
    """
    nz, ny, nx = shape
    solid = np.zeros(shape, dtype=bool)
    ctr = 0
    while True:
        u = uniform(seed, 1, 4 * batch, start=4 * ctr)
        ctr += batch
        c = np.stack([u[0::4] * nx, u[1::4] * ny, u[2::4] * nz], axis=1)
        if np.isscalar(radius):
            r = np.full(batch, float(radius))
        else:
            r = np.round(radius[0] + u[3::4] * (radius[1] - radius[0]))
        trial = solid.copy()
        _stamp(trial, c, r)
        if 1.0 - trial.mean() > porosity:
            solid = trial
            continue
        for s in range(batch):  # trim the final batch
            _stamp(solid, c[s:s + 1], r[s:s + 1])
            if 1.0 - solid.mean() <= porosity:
                break
        break
    pore = ~solid
    if invert:
        pore = ~pore
    return np.where(pore, 1, 0).astype(np.int32)


def two_component(ind: np.ndarray, frac2: float, seed: int = 0, radius=4.0) -> np.ndarray:
    """Split the solid (code 0) into component 1 (code 0) and component 2 (code 3).

    Component 2 is grown as random balls (seeded) inside the solid until a
    fraction ``frac2`` of the solid voxels carries code 3.
    """
    out = ind.copy()
    solid = ind == 0
    nsolid = int(solid.sum())
    if nsolid == 0 or frac2 <= 0:
        return out
    comp2 = np.zeros(ind.shape, dtype=bool)
    _grow_balls(comp2, solid, frac2 * nsolid, seed, 7, float(radius))
    out[comp2 & solid] = 3
    return out


def phase_blobs(ind: np.ndarray, saturation: float, seed: int = 0, radius: float = 3.0,
                iipore: int = 1) -> np.ndarray:
    """Boolean 'blue' mask over the voxel grid: random balls of non-wetting fluid
    in the pore space, added until ``saturation`` of the pore voxels is blue."""
    pore = ind == iipore
    npore = int(pore.sum())
    blue = np.zeros(ind.shape, dtype=bool)
    if npore == 0 or saturation <= 0:
        return blue
    _grow_balls(blue, pore, saturation * npore, seed, 3, float(radius))
    return blue & pore


# ---------------------------------------------------------------- tables --
_RING8 = [[2, 8, 4, 9, 3, 11, 5, 10], [2, 12, 6, 13, 3, 15, 7, 14], [4, 16, 6, 17, 5, 19, 7, 18]]
_RING4 = [[12, 16, 13, 17], [14, 18, 15, 19], [8, 12, 10, 14], [9, 13, 11, 15],
          [8, 16, 9, 18], [10, 17, 11, 19]]


def angle_types() -> np.ndarray:
    """The 48 angle types (arm j, arm k), 1-based D3Q19 indices, in the
    reference's order: 45-degree rings in the xy, xz, yz planes, then the six
    60-degree rings of diagonals sharing a +-1 component."""
    t = []
    for ring in _RING8:
        t += [(ring[m], ring[(m + 1) % 8]) for m in range(8)]
    for ring in _RING4:
        t += [(ring[m], ring[(m + 1) % 4]) for m in range(4)]
    return np.array(t, dtype=np.int32)


_CORNERS = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0],
                     [0, 0, 1], [1, 0, 1], [0, 1, 1], [1, 1, 1]], dtype=np.int32)


def _corner(v) -> int:
    for c, cv in enumerate(_CORNERS):
        if tuple(cv) == tuple(v):
            return c + 1
    return 0


def cube_tables():
    """Element tables of one voxel cube, same content and order as the
    reference's cube_bon/cube_ang/cube_sbond/cube_abond files.

    Bonds: from each corner (1..8) along each of the 9 bond vectors that stays
    in the cube; weight 1/4 for an edge (shared by 4 cubes), 1/2 for a face
    diagonal (shared by 2 cubes).  Angles: at each corner, each of the 48
    angle types whose two arms stay in the cube; weight 1/2 for a 45-degree
    angle (lies in a face) and 1 for a 60-degree angle (inside one cube).
    """
    bon, sb = [], []
    for p1 in range(8):
        for b in range(9):
            p2 = _corner(_CORNERS[p1] + BOND9[:, b])
            if p2:
                bon.append((p1 + 1, p2, b + 1))
                sb.append(0.25 if b < 3 else 0.5)
    ang, ab = [], []
    types = angle_types()
    for p1 in range(8):
        for t, (qj, qk) in enumerate(types):
            pj = _corner(_CORNERS[p1] + D3Q19[:, qj - 1])
            pk = _corner(_CORNERS[p1] + D3Q19[:, qk - 1])
            if pj and pk:
                ang.append((p1 + 1, pj, pk, qj, qk, t + 1))
                ab.append(1.0 if (qj > 7 and qk > 7) else 0.5)
    return (np.array(bon, dtype=np.int32), np.array(ang, dtype=np.int32),
            np.array(sb, dtype=np.float64), np.array(ab, dtype=np.float64))
