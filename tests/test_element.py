"""The LSM element-matrix formulation, checked on the CPU against the oracle's
bond/angle force loops (the reference algorithm) through the reference's node
numbering.  This pins the exact matrices generated into lsm_element.h, the
periodic-with-jump representation of the glued boundary, and
reference_node_map, before any GPU run."""
import os
import sys

import numpy as np
import pytest

import oracle_bind as ob
from porescale_amd import geometry as geo
from porescale_amd.lsm import CASES, reference_node_map

sys.path.insert(0, os.path.join(ob.ROOT, "porescale_amd", "csrc"))
from gen_lsm_element import build  # noqa: E402

CORNER = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0], [0, 0, 1], [1, 0, 1], [0, 1, 1], [1, 1, 1]])


@pytest.fixture(scope="module")
def mats():
    KA, KB, SA, SB = build()
    return tuple(np.array(M, dtype=float) for M in (KA, KB, SA, SB))


def cube_fields(ind, alf, bet):
    comp = np.where(ind == 0, 1, np.where(ind == 3, 2, 0))
    ka = np.where(comp == 1, alf[0], np.where(comp == 2, alf[1], 0.0))
    kb = np.where(comp == 1, bet[0], np.where(comp == 2, bet[1], 0.0))
    return ka, kb


def gather_cubes(u, J, vox):
    """u (3,nz,ny,nx) -> (nvox, 24) corner displacements with the jump."""
    _, nz, ny, nx = u.shape
    kz, ky, kx = vox
    out = np.zeros((len(kx), 24))
    for p, (a, b, c) in enumerate(CORNER):
        x, y, z = kx + a, ky + b, kz + c
        wx, wy, wz = x // nx, y // ny, z // nz
        for e in range(3):
            out[:, 3 * p + e] = u[e, z % nz, y % ny, x % nx] + wx * J[0][e] + wy * J[1][e] + wz * J[2][e]
    return out, (kx, ky, kz)


def element_forces(u, ind, J, KA, KB, alf, bet):
    ka, kb = cube_fields(ind, alf, bet)
    vox = np.nonzero((ka != 0) | (kb != 0))
    uc, (kx, ky, kz) = gather_cubes(u, J, vox)
    fc = -(ka[vox][:, None] * (uc @ KA.T) + kb[vox][:, None] * (uc @ KB.T))
    f = np.zeros_like(u)
    nz, ny, nx = ind.shape
    for p, (a, b, c) in enumerate(CORNER):
        for e in range(3):
            np.add.at(f[e], ((kz + c) % nz, (ky + b) % ny, (kx + a) % nx), fc[:, 3 * p + e])
    return f


def jump(case, g):
    p, d = CASES[case]
    J = np.zeros((3, 3))
    J[p][d] = 2 * g
    return J


@pytest.mark.parametrize("case", ["xx", "yz", "zz", "xy"])
def test_element_operator_equals_reference_force_loops(case, mats):
    KA, KB, _, _ = mats
    ind = geo.two_component(geo.sphere_pack((7, 8, 9), 0.45, 2.0, seed=21), 0.4, seed=22, radius=1.5)
    alf, bet = (1.0, 0.37), (0.1, 0.021)
    g = 0.01
    o = ob.OracleLSM(ind, case, alf=alf[0], bet=bet[0], alf2=alf[1], bet2=bet[1], gstrain=g)
    grid, twin, G = reference_node_map(ind, case)
    icors, _ = o.nodes()
    assert len(grid) == o.ns
    nz, ny, nx = ind.shape
    # the map reproduces the reference's node coordinates
    pa = CASES[case][0]
    gx = grid % nx
    gy = (grid // nx) % ny
    gz = grid // (nx * ny)
    coords = np.stack([gx, gy, gz], 1)
    coords[:, pa] += np.where(twin, ind.shape[2 - pa], 0)
    assert np.array_equal(coords + 1, icors)
    rng = np.random.default_rng(3)
    u = rng.standard_normal((3, nz, ny, nx)) * 1e-3
    J = jump(case, g)
    uref = u.reshape(3, -1)[:, grid].T + np.where(twin[:, None], J[pa][None, :], 0.0)
    fref = o.forces(uref)
    # glue: both twins act on the one periodic node
    fgrid = np.zeros((3, nx * ny * nz))
    for e in range(3):
        np.add.at(fgrid[e], grid, fref[:, e])
    fel = element_forces(u, ind, J, KA, KB, alf, bet).reshape(3, -1)
    scale = np.abs(fgrid).max()
    np.testing.assert_allclose(fel, fgrid, atol=1e-12 * scale)


@pytest.mark.parametrize("case", ["xx", "yz"])
def test_element_stress_equals_calc_stress(case, mats):
    KA, KB, SA, SB = mats
    ind = geo.two_component(geo.sphere_pack((6, 7, 8), 0.4, 2.0, seed=31), 0.3, seed=32, radius=1.5)
    alf, bet = (1.0, 0.5), (0.1, 0.05)
    g = 0.01
    o = ob.OracleLSM(ind, case, alf=alf[0], bet=bet[0], alf2=alf[1], bet2=bet[1], gstrain=g)
    grid, twin, _ = reference_node_map(ind, case)
    nz, ny, nx = ind.shape
    rng = np.random.default_rng(4)
    u = rng.standard_normal((3, nz, ny, nx)) * 1e-3
    J = jump(case, g)
    pa = CASES[case][0]
    uref = u.reshape(3, -1)[:, grid].T + np.where(twin[:, None], J[pa][None, :], 0.0)
    o.set_state(r=uref)
    asig, _, _, _ = o.stress()
    ka, kb = cube_fields(ind, alf, bet)
    vox = np.nonzero((ka != 0) | (kb != 0))
    uc, _ = gather_cubes(u, J, vox)
    sig = (ka[vox][:, None] * (uc @ SA.T) + kb[vox][:, None] * (uc @ SB.T)).sum(0)
    np.testing.assert_allclose(sig.reshape(3, 3, order="F"), asig, atol=1e-12 * np.abs(asig).max())


def test_element_matrices_structure(mats):
    KA, KB, SA, SB = mats
    assert np.array_equal(KA, KA.T) and np.array_equal(KB, KB.T)
    assert np.linalg.eigvalsh(KA).min() > -1e-12 and np.linalg.eigvalsh(KB).min() > -1e-12
    for d in range(3):  # translations carry no energy and no stress
        t = np.zeros(24)
        t[d::3] = 1
        assert np.abs(KA @ t).max() == 0 and np.abs(KB @ t).max() == 0
        assert np.abs(SA @ t).max() == 0 and np.abs(SB @ t).max() == 0
    # homogeneous uniaxial strain: energy density 3a+8b (C11), stress the same
    u = np.zeros(24)
    for p, x in enumerate(CORNER):
        u[3 * p] = x[0]
    assert u @ KA @ u == pytest.approx(3.0) and u @ KB @ u == pytest.approx(8.0)
    assert (SA @ u)[0] == pytest.approx(3.0) and (SB @ u)[0] == pytest.approx(8.0)


def _node_stencil_from_header():
    """The generated node_stencil routine of lsm_element.h as a function of
    (ka[8], kb[8]) -> (27, 3, 3) blocks, read from the header text."""
    import re
    text = open(os.path.join(ob.ROOT, "porescale_amd", "csrc", "lsm_element.h")).read()
    body = text[text.index("void node_stencil("):]
    body = body[:body.index("\n}\n")]
    lines = re.findall(r"out\[(\d+) \* stride\] = (.*);", body)
    assert len(lines) == 243 and sorted(int(k) for k, _ in lines) == list(range(243))
    terms = []
    for k, expr in lines:
        tl = []
        for coef, arr, p in re.findall(r"T\(([-0-9.e]+)\) \* (ka|kb)\[(\d)\]", expr):
            tl.append((float(coef), arr, int(p)))
        assert expr == "T(0)" or len(tl) == expr.count("*")
        terms.append((int(k), tl))

    def stencil(ka8, kb8):
        out = np.zeros(243)
        for k, tl in terms:
            out[k] = sum(c * (ka8[p] if arr == "ka" else kb8[p]) for c, arr, p in tl)
        return out.reshape(27, 3, 3)
    return stencil


def test_generated_node_stencil_equals_element_assembly(mats):
    """Row blocks K(n, n + offset) of the assembled operator (used to build
    the multigrid hierarchy) against a direct sum of the element matrices of
    the 8 cubes around the node, with random per-cube scales."""
    KA, KB = mats[0], mats[1]
    stencil = _node_stencil_from_header()
    rng = np.random.default_rng(4)
    for _ in range(3):
        ka8, kb8 = rng.random(8), rng.random(8)
        ref = np.zeros((27, 3, 3))
        for p in range(8):          # node is corner p of the cube at node - CORNER[p]
            for q in range(8):      # neighbour corner q of that cube
                off = CORNER[q] - CORNER[p]
                s = (off[0] + 1) + 3 * (off[1] + 1) + 9 * (off[2] + 1)
                ref[s] += ka8[p] * KA[3 * p:3 * p + 3, 3 * q:3 * q + 3] + kb8[p] * KB[3 * p:3 * p + 3, 3 * q:3 * q + 3]
        np.testing.assert_allclose(stencil(ka8, kb8), ref, rtol=0, atol=1e-14)
    # the rows of a homogeneous solid sum to zero (rigid translations)
    st = stencil(np.ones(8), np.full(8, 0.1))
    np.testing.assert_allclose(st.sum(axis=0), 0.0, atol=1e-13)


def _periodic_interp(nf):
    """1D trilinear interpolation of the multigrid: coarse node I on fine 2I,
    odd fine nodes halfway between their two (periodic) even neighbours."""
    P = np.zeros((nf, (nf + 1) // 2))
    for j in range(nf):
        if j % 2 == 0:
            P[j, j // 2] = 1.0
        else:
            P[j, (j - 1) // 2] += 0.5
            P[j, (0 if j + 1 == nf else j + 1) // 2] += 0.5
    return P


def _assemble(n, elems):
    """Dense matrix of a periodic grid (x fastest) from per-element 24x24 blocks."""
    nx, ny, nz = n
    N = nx * ny * nz
    A = np.zeros((3 * N, 3 * N))
    for (x, y, z), K in elems:
        nodes = [((x + c[0]) % nx) + nx * (((y + c[1]) % ny) + ny * ((z + c[2]) % nz)) for c in CORNER]
        for f in range(8):
            for h in range(8):
                A[3 * nodes[f]:3 * nodes[f] + 3, 3 * nodes[h]:3 * nodes[h] + 3] += K[3 * f:3 * f + 3, 3 * h:3 * h + 3]
    return A


@pytest.mark.parametrize("n", [(6, 6, 6), (7, 6, 8), (7, 9, 5)])
def test_first_galerkin_product_from_projected_element_matrices(n, mats):
    """The first coarse operator P^T A P (trilinear P, periodic, odd or even
    axes) equals the sum over fine voxels of a_v P_p^T KA P_p + b_v P_p^T KB
    P_p placed on coarse element floor(v/2), with the per-axis patterns the
    multigrid set-up tabulates (even voxel [[1,0],[1/2,1/2]], odd
    [[1/2,1/2],[0,1]], the wrapping voxel of an odd axis the identity).  A
    wrong pattern or a transposed corner order moves entries by O(1)."""
    KA, KB = mats[0], mats[1]
    rng = np.random.default_rng(3)
    nx, ny, nz = n
    N = nx * ny * nz
    ka = rng.uniform(0.2, 1.5, N)
    kb = rng.uniform(0.01, 0.2, N)
    ka[rng.random(N) < 0.3] = 0.0
    kb[ka == 0.0] = 0.0
    vid = lambda x, y, z: x + nx * (y + ny * z)  # noqa: E731
    fine = [((x, y, z), ka[vid(x, y, z)] * KA + kb[vid(x, y, z)] * KB)
            for z in range(nz) for y in range(ny) for x in range(nx)]
    P = np.kron(np.kron(_periodic_interp(nz), _periodic_interp(ny)), _periodic_interp(nx))
    P3 = np.kron(P, np.eye(3))
    want = P3.T @ _assemble(n, fine) @ P3
    pat = [np.array([[1.0, 0.0], [0.5, 0.5]]), np.array([[0.5, 0.5], [0.0, 1.0]]), np.eye(2)]
    nc = [(a + 1) // 2 for a in n]
    coarse = []
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                f = (x, y, z)
                E = tuple(c // 2 for c in f)
                p = [2 if (n[d] % 2 == 1 and f[d] == n[d] - 1) else f[d] % 2 for d in range(3)]
                Pe = np.array([[np.prod([pat[p[d]][(a >> d) & 1, (b >> d) & 1] for d in range(3)])
                                for b in range(8)] for a in range(8)])
                Pe3 = np.kron(Pe, np.eye(3))
                v = vid(x, y, z)
                coarse.append((E, Pe3.T @ (ka[v] * KA + kb[v] * KB) @ Pe3))
    got = _assemble(tuple(nc), coarse)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12 * np.abs(want).max())
