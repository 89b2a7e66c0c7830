"""LSM HIP kernels against the oracle and closed forms (GPU)."""
import os

import numpy as np
import pytest

import oracle_bind as ob
from golden.make_golden import LSM_INP, lsm_converged_geometry, lsm_geometry
from porescale_amd import geometry as geo
from porescale_amd.lsm import CASES, LSM, batch_moduli, moduli_from_C, reference_node_map, spring_constants

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _say(msg):
    print(msg, flush=True)


def _to_ref(u, ind, case, g):
    grid, twin, _ = reference_node_map(ind, case)
    p, d = CASES[case]
    J = np.zeros(3)
    J[d] = 2 * g
    return u.reshape(3, -1)[:, grid].T + np.where(twin[:, None], J[None, :], 0.0)


@pytest.mark.parametrize("case", ["xx", "yz", "zx"])
def test_forces_match_reference_force_loops(case):
    ind = geo.two_component(geo.sphere_pack((9, 10, 11), 0.45, 2.5, seed=1), 0.4, seed=2, radius=1.5)
    alf, bet = (1.0, 0.4), (0.1, 0.03)
    g = 0.01
    _say(f"case forces {case}")
    o = ob.OracleLSM(ind, case, alf=alf[0], bet=bet[0], alf2=alf[1], bet2=bet[1], gstrain=g)
    rng = np.random.default_rng(1)
    u = rng.standard_normal((3,) + ind.shape) * 1e-3
    with LSM(ind, alf[0], bet[0], alf[1], bet[1], gstrain=g) as s:
        f = s.forces(case, u).reshape(3, -1)
    fref = o.forces(_to_ref(u, ind, case, g))
    grid, _, _ = reference_node_map(ind, case)
    fg = np.zeros_like(f)
    for e in range(3):
        np.add.at(fg[e], grid, fref[:, e])
    np.testing.assert_allclose(f, fg, atol=1e-12 * np.abs(fg).max())


@pytest.mark.parametrize("case", ["xx", "yz"])
def test_verlet_trajectory_matches_golden_reference(case):
    ind = lsm_geometry()
    g = np.load(os.path.join(GOLD, f"lsm_{case}.npz"))
    a1, b1 = spring_constants(LSM_INP["Em1"], LSM_INP["nu1"])
    a2, b2 = spring_constants(LSM_INP["Em2"], LSM_INP["nu2"])
    _say(f"case verlet {case}")
    with LSM(ind, a1, b1, a2, b2, rs=1.0, dt=0.01, gstrain=0.01) as s:
        s.verlet(case, int(g["steps"]))
        u = s.get_u(case)
        _, m = s.calc_stress(case)
    r = _to_ref(u, ind, case, 0.01)
    np.testing.assert_allclose(r, g["r"], rtol=0, atol=1e-13 * np.abs(g["r"]).max())
    if case == "xx":
        assert m[0, 0] == pytest.approx(float(g["Exxxx"]), rel=1e-11)
        assert m[1, 1] == pytest.approx(float(g["Eyyxx"]), rel=1e-10)
    else:
        assert m[1, 2] == pytest.approx(float(g["Eyzyz"]), rel=1e-11)


@pytest.mark.parametrize("case", ["xx", "yz"])
def test_verlet_restart_in_reference_node_order(case):
    """The state in the reference's node order (iteration2): after 100 steps
    it is the reference's own iteration2; restarting from it (icon = 1, glued
    pairs reset) and running 50 more steps gives the reference's restart run."""
    ind = lsm_geometry()
    g = np.load(os.path.join(GOLD, f"lsm_{case}.npz"))
    g1 = np.load(os.path.join(GOLD, f"lsm_{case}_restart.npz"))
    a1, b1 = spring_constants(LSM_INP["Em1"], LSM_INP["nu1"])
    a2, b2 = spring_constants(LSM_INP["Em2"], LSM_INP["nu2"])
    _say(f"case verlet restart {case}")
    with LSM(ind, a1, b1, a2, b2, rs=1.0, dt=0.01, gstrain=0.01) as s:
        grid, twin = s.reference_nodes(case)
        assert len(grid) == g["r"].shape[0]
        s.verlet(case, int(g["steps"]))
        r, v, a = s.get_verlet(case)
        for x, key in ((r, "r"), (v, "v"), (a, "a")):
            np.testing.assert_allclose(x, g[key], rtol=0, atol=1e-13 * np.abs(g[key]).max())
        s.set_verlet(case, g["r"], g["v"], g["a"])
        s.verlet(case, int(g1["steps"]), restart=False)
        r, v, a = s.get_verlet(case)
    for x, key in ((r, "r"), (v, "v"), (a, "a")):
        np.testing.assert_allclose(x, g1[key], rtol=0, atol=1e-12 * np.abs(g1[key]).max())


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_all_solid_closed_form(prec):
    a, b = 1.0, 0.1
    ind = np.zeros((8, 8, 8), np.int32)
    _say(f"case all-solid {prec}")
    with LSM(ind, a, b, precision=prec) as s:
        C6, K, G = s.moduli(rtol=1e-12, maxit=5000)
    c11, c12, c44 = 3 * a + 8 * b, a - 4 * b, a + 6 * b
    expect = np.zeros((6, 6))
    expect[:3, :3] = c12
    np.fill_diagonal(expect[:3, :3], c11)
    expect[3, 3] = expect[4, 4] = expect[5, 5] = c44
    np.testing.assert_allclose(C6, expect, atol=1e-9)
    assert K == pytest.approx(5 * a / 3, rel=1e-10) and G == pytest.approx(c44, rel=1e-10)


def test_solve_equals_converged_verlet_and_fp32_refinement():
    ind = geo.sphere_pack((10, 10, 10), 0.3, 3.0, seed=3)
    _say("case solve vs verlet")
    with LSM(ind, 1.0, 0.1, dt=0.05) as s:
        it, rel = s.solve("xx", 1e-11, 20000)
        assert rel <= 1e-11
        _, m_cg = s.calc_stress("xx")
        s.verlet("xx", 40000)
        _, m_v = s.calc_stress("xx")
    with LSM(ind, 1.0, 0.1, precision="f32") as s32:
        it32, rel32 = s32.solve("xx", 1e-11, 20000)
        _, m32 = s32.calc_stress("xx")
    assert rel32 <= 1e-11
    assert m_v[0, 0] == pytest.approx(m_cg[0, 0], rel=1e-6)
    assert m32[0, 0] == pytest.approx(m_cg[0, 0], rel=1e-9)


def test_moduli_symmetry_and_batch_equals_single():
    vols = [geo.two_component(geo.sphere_pack(sh, 0.35, 2.5, seed=s), 0.3, seed=s + 1, radius=1.5)
            for s, sh in ((1, (10, 12, 9)), (2, (8, 8, 8)), (3, (12, 10, 11)))]
    pars = (1.0, 0.1, 0.3, 0.02)
    singles = []
    for v in vols:
        with LSM(v, *pars) as s:
            C6, K, G = s.moduli(rtol=1e-10)
            singles.append((C6, K, G))
            assert (K, G) == pytest.approx(moduli_from_C(C6), rel=1e-12)
            np.testing.assert_allclose(C6, C6.T, atol=1e-7 * np.abs(C6).max())
    _say("case batch of 3")
    for prec in ("f64", "f32"):
        Cb, KG, ms, its = batch_moduli(vols, *pars, precision=prec, rtol=1e-10)
        for (C6, K, G), cb, kg in zip(singles, Cb, KG):
            np.testing.assert_allclose(cb, C6, atol=1e-8 * np.abs(C6).max())
            assert kg[0] == pytest.approx(K, rel=1e-8) and kg[1] == pytest.approx(G, rel=1e-8)


def test_all_rhs_coarse_step_batch_equals_single(monkeypatch):
    """Eight 64^3 volumes in one FP32 multigrid batch: the first coarse level
    holds 8 x 32^3 nodes = 1024 tiles, so its Chebyshev steps apply each
    stencil block to all six right-hand sides in one thread; each volume alone
    (128 tiles there) takes one grid row per right-hand side.  Both converge
    to the same FP64-refined moduli."""
    vols = [geo.two_component(geo.sphere_pack((64, 64, 64), 0.35, 5.0, seed=60 + i), 0.5, seed=70 + i,
                              radius=4.0) for i in range(8)]
    pars = (1.0, 0.1, 0.3, 0.02)
    singles = []
    for v in vols:
        with LSM(v, *pars, precision="f32") as s:
            singles.append(s.moduli(rtol=1e-10))
    runs = {}
    for label, env in (("all right-hand sides per thread", {}),
                       ("one grid row per right-hand side", {"PSX_MG_RALL_TILES": "0"})):
        for k, val in env.items():
            monkeypatch.setenv(k, val)
        _say(f"case batch of eight 64^3 volumes, level 1 {label}")
        runs[label] = batch_moduli(vols, *pars, precision="f32", rtol=1e-10)
        for k in env:
            monkeypatch.delenv(k)
        Cb, KG, ms, its = runs[label]
        for (C6, K, G), cb, kg in zip(singles, Cb, KG):
            np.testing.assert_allclose(cb, C6, atol=1e-8 * np.abs(C6).max())
            assert kg[0] == pytest.approx(K, rel=1e-8) and kg[1] == pytest.approx(G, rel=1e-8)
    # the two kernels apply the same operator (up to FP32 rounding), so the
    # preconditioner and the iteration count stay the same: a wrong block
    # would still converge, only more slowly
    its = [r[3] for r in runs.values()]
    assert max(its) - min(its) <= 2, its


def test_multigrid_setup_variants_same_preconditioner(monkeypatch):
    """The first Galerkin product from the projected element matrices (the
    default) and from the assembled fine stencil (PSX_MG_GALERKIN_ELEM=0),
    and the FP16 smoothing stencils (default) against the FP32 ones
    (PSX_MG_HALF=0), on a batch with odd and even axes (the wrapping coarse
    element of an odd axis holds one fine voxel): the same moduli and, the
    preconditioners agreeing to rounding, the same iteration counts within
    one."""
    vols = [geo.two_component(geo.sphere_pack(sh, 0.35, 3.0, seed=80 + i), 0.4, seed=90 + i, radius=2.0)
            for i, sh in enumerate([(20, 20, 20), (23, 18, 21), (17, 25, 19)])]
    pars = (1.0, 0.1, 0.3, 0.02)
    runs = {}
    for label, env in (("default", {}), ("assembled-stencil Galerkin product", {"PSX_MG_GALERKIN_ELEM": "0"}),
                       ("FP32 smoothing stencils", {"PSX_MG_HALF": "0"})):
        for k, val in env.items():
            monkeypatch.setenv(k, val)
        _say(f"case multigrid set-up: {label}")
        runs[label] = batch_moduli(vols, *pars, precision="f32", rtol=1e-10)
        for k in env:
            monkeypatch.delenv(k)
    C0, KG0, _, it0 = runs["default"]
    for label, (C1, KG1, _, it1) in runs.items():
        np.testing.assert_allclose(C1, C0, rtol=0, atol=1e-8 * np.abs(C0).max())
        np.testing.assert_allclose(KG1, KG0, rtol=1e-8)
        assert abs(it1 - it0) <= 1, (label, it1, it0)


@pytest.mark.parametrize("levels", ["1", "4"])
def test_multigrid_coarsest_degree_one_same_moduli(monkeypatch, levels):
    """A coarsest level of Chebyshev degree 1 takes no smoothing step, so its
    x must be zeroed where the pass starts (every other level's first step
    stores x = d).  One level: the CG update zeroes the finest x; four: the
    restriction zeroes the coarsest."""
    ind = geo.two_component(geo.sphere_pack((24, 20, 22), 0.35, 3.0, seed=3), 0.3, seed=4, radius=2.0)
    pars = (1.0, 0.1, 0.3, 0.02)
    with LSM(ind, *pars, precision="f32") as s:
        C6, K, G = s.moduli(rtol=1e-10)
    monkeypatch.setenv("PSX_MG_CDEG", "1")
    monkeypatch.setenv("PSX_MG_LEVELS", levels)
    monkeypatch.setenv("PSX_MG_CMAX", "3")  # coarsen as deep as the levels allow
    _say(f"case coarsest degree 1, {levels} level(s)")
    with LSM(ind, *pars, precision="f32") as s:
        C1, K1, G1 = s.moduli(rtol=1e-10, maxit=200000)
    np.testing.assert_allclose(C1, C6, atol=1e-8 * np.abs(C6).max())
    assert K1 == pytest.approx(K, rel=1e-8) and G1 == pytest.approx(G, rel=1e-8)


def test_adjoint_closed_form_homogeneous():
    a, b = 1.0, 0.1
    ind = np.zeros((6, 6, 6), np.int32)
    with LSM(ind, a, b) as s:
        s.moduli(rtol=1e-12)
        dK, dsK = s.adjoint(1.0, 0.0, rtol=1e-12)
        dG, dsG = s.adjoint(0.0, 1.0, rtol=1e-12)
    # K = 5a/3, G = a + 6b for one component (code 0)
    assert dK[0] == pytest.approx(5.0 / 3.0, rel=1e-8) and abs(dK[1]) < 1e-9
    assert dG[0] == pytest.approx(1.0, rel=1e-8) and dG[1] == pytest.approx(6.0, rel=1e-8)
    # uniform stiffness field: dK/ds_v = K/N
    np.testing.assert_allclose(dsK, (5 * a / 3) / ind.size, rtol=1e-7)


def test_adjoint_matches_finite_differences():
    ind = geo.two_component(geo.sphere_pack((8, 9, 10), 0.35, 2.5, seed=8), 0.4, seed=9, radius=1.5)
    ab = np.array([1.0, 0.1, 0.3, 0.02])
    rng = np.random.default_rng(5)
    stiff = np.exp(0.3 * rng.standard_normal(ind.shape))
    _say("case adjoint fd")

    def KG(ab_, st_):
        with LSM(ind, *ab_, stiff=st_) as s:
            _, K, G = s.moduli(rtol=1e-12, maxit=20000)
        return np.array([K, G])

    with LSM(ind, *ab, stiff=stiff) as s:
        s.moduli(rtol=1e-12, maxit=20000)
        gK, gsK = s.adjoint(1.0, 0.0, rtol=1e-12)
        gG, _ = s.adjoint(0.0, 1.0, rtol=1e-12)
    h = 1e-6
    for k in range(4):
        e = np.zeros(4)
        e[k] = h * ab[k]
        fd = (KG(ab + e, stiff) - KG(ab - e, stiff)) / (2 * e[k])
        assert gK[k] == pytest.approx(fd[0], rel=1e-5, abs=1e-8)
        assert gG[k] == pytest.approx(fd[1], rel=1e-5, abs=1e-8)
    solid = np.argwhere(ind != 1)
    for idx in solid[[0, len(solid) // 2, -1]]:
        e = np.zeros(ind.shape)
        e[tuple(idx)] = h
        fd = (KG(ab, stiff + e)[0] - KG(ab, stiff - e)[0]) / (2 * h)
        assert gsK[tuple(idx)] == pytest.approx(fd, rel=1e-4, abs=1e-9)


def test_matvec_profile():
    ind = geo.sphere_pack((32, 32, 32), 0.35, 6.0, seed=0)
    with LSM(ind, 1.0, 0.1) as s:
        t64, t32 = s.profile(5)
    assert t64 > 0 and t32 > 0


def test_cyclic_axis_permutation_maps_all_six_cases():
    """The reference ships only the xx and yz programs.  yy, zz, zx and xy are
    pinned to them by symmetry: relabelling the axes cyclically (x->y->z->x)
    maps load case xx->yy->zz, yz->zx->xy and the stress rows the same way,
    so C'[s(I), s(J)] = C[I, J] exactly (to the solver tolerance)."""
    ind = geo.two_component(geo.sphere_pack((7, 9, 8), 0.35, 2.0, seed=21), 0.4, seed=22, radius=1.5)
    rot = np.transpose(ind, (1, 2, 0))  # new (z', y', x') = old (y, x, z)
    _say("case cyclic permutation")
    with LSM(ind, 1.0, 0.1, 0.3, 0.02) as s:
        C, K, G = s.moduli(rtol=1e-12, maxit=20000)
    with LSM(rot, 1.0, 0.1, 0.3, 0.02) as s:
        Cr, Kr, Gr = s.moduli(rtol=1e-12, maxit=20000)
    sig = np.array([1, 2, 0, 4, 5, 3])
    mapped = np.empty_like(C)
    mapped[np.ix_(sig, sig)] = C
    np.testing.assert_allclose(Cr, mapped, rtol=0, atol=1e-9 * np.abs(C).max())
    assert Kr == pytest.approx(K, rel=1e-10) and Gr == pytest.approx(G, rel=1e-10)
    # a non-trivial geometry: the six cases really differ
    assert abs(C[0, 0] - C[1, 1]) > 1e-4 * C[0, 0]


def test_multigrid_preconditioner_cuts_iterations(monkeypatch):
    """The FP32 inner solves use the multigrid V-cycle; block Jacobi
    (PSX_MG=0) must need several times more iterations for the same answer."""
    ind = geo.two_component(geo.sphere_pack((24, 20, 22), 0.35, 3.0, seed=3), 0.3, seed=4, radius=2.0)
    out = {}
    for mg in ("0", "1"):
        monkeypatch.setenv("PSX_MG", mg)
        _say(f"case multigrid PSX_MG={mg}")
        with LSM(ind, 1.0, 0.1, 0.3, 0.02, precision="f32") as s:
            it, rel = s.solve("yz", 1e-10, 50000)
            _, m = s.calc_stress("yz")
        assert rel <= 1e-10
        out[mg] = (it, m[1, 2])
    assert out["1"][0] * 3 < out["0"][0]
    assert out["1"][1] == pytest.approx(out["0"][1], rel=1e-9)


def test_edge_volumes_single_and_batched():
    """Edge volumes, alone and together in one batch with an ordinary pack:
    an all-pore volume (no solid node: every load case is the zero system
    and the moduli are zero, as the reference's empty node loops give), the
    smallest volume (2 voxels a side, all solid: the closed form), and a
    solid island that touches no face (a floating cluster: its right-hand
    side is zero, so the singular block does not disturb the solve)."""
    a, b = 1.0, 0.1
    c11, c12, c44 = 3 * a + 8 * b, a - 4 * b, a + 6 * b
    closed = np.zeros((6, 6))
    closed[:3, :3] = c12
    np.fill_diagonal(closed[:3, :3], c11)
    closed[3, 3] = closed[4, 4] = closed[5, 5] = c44
    pore = np.ones((6, 7, 5), np.int32)
    tiny = np.zeros((2, 2, 2), np.int32)
    island = geo.sphere_pack((10, 9, 11), 0.4, 2.5, seed=41)
    island[4:6, 4:6, 4:6] = 0
    island[3, 3:7, 3:7] = 1
    island[6, 3:7, 3:7] = 1
    island[3:7, 3, 3:7] = 1
    island[3:7, 6, 3:7] = 1
    island[3:7, 3:7, 3] = 1
    island[3:7, 3:7, 6] = 1
    vols = [pore, tiny, island]
    singles = []
    for v, name in zip(vols, ("all-pore", "2x2x2 solid", "island")):
        for prec in ("f64", "f32"):
            _say(f"case edge volume {name} {prec}")
            with LSM(v, a, b, precision=prec) as s:
                singles.append(s.moduli(rtol=1e-10, maxit=20000))
    (Cp, Kp, Gp), (Cp32, _, _) = singles[0], singles[1]
    assert np.all(Cp == 0.0) and Kp == 0.0 and Gp == 0.0 and np.all(Cp32 == 0.0)
    for C6, _, _ in singles[2:4]:
        np.testing.assert_allclose(C6, closed, atol=1e-9)
    np.testing.assert_allclose(singles[5][0], singles[4][0], atol=1e-8 * np.abs(singles[4][0]).max())
    for prec in ("f64", "f32"):
        _say(f"case edge volumes in one batch {prec}")
        Cb, KG, ms, its = batch_moduli(vols, a, b, a, b, precision=prec, rtol=1e-10, maxit=20000)
        assert np.all(Cb[0] == 0.0) and np.all(KG[0] == 0.0)
        np.testing.assert_allclose(Cb[1], closed, atol=1e-9)
        np.testing.assert_allclose(Cb[2], singles[4][0], atol=1e-8 * np.abs(singles[4][0]).max())


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_moduli_equal_reference_run_to_equilibrium(precision):
    """The reference's own solver run to equilibrium (XX_P.f and YZ_P.f, 2 x
    20000 damped-Verlet steps on a 10^3 two-component pack; golden
    lsm_converged.npz): the six-case moduli of the GPU solvers (FP64 PCG, and
    FP32 multigrid PCG with FP64 refinement) and of the batched solve of that
    volume beside a second one equal its Exxxx, Eyyxx, Eyyzz and Eyzyz to
    1e-9 relative."""
    g = np.load(os.path.join(GOLD, "lsm_converged.npz"))
    # the reference itself is converged: its last two blocks agree
    for k in ("xx_Exxxx", "xx_Eyyxx", "xx_Eyyzz", "yz_Eyzyz"):
        assert abs(g[k][-1] - g[k][-2]) <= 1e-12 * abs(g[k][-1])
    ind = lsm_converged_geometry()
    a1, b1 = spring_constants(LSM_INP["Em1"], LSM_INP["nu1"])
    a2, b2 = spring_constants(LSM_INP["Em2"], LSM_INP["nu2"])
    want = {(0, 0): g["xx_Exxxx"][-1], (1, 0): g["xx_Eyyxx"][-1], (2, 0): g["xx_Eyyzz"][-1],
            (3, 3): g["yz_Eyzyz"][-1]}
    print(f"case converged reference moduli 10^3 {precision}", flush=True)
    with LSM(ind, a1, b1, a2, b2, precision=precision, gstrain=LSM_INP["gstrain"]) as s:
        C6, K, G = s.moduli(1e-11, 200000)
    other = geo.two_component(geo.sphere_pack((12, 9, 11), 0.4, 2.5, seed=13), 0.5, seed=14, radius=2.0)
    Cb, KG, _, _ = batch_moduli([other, ind], a1, b1, a2, b2, precision=precision, rtol=1e-11)
    for (i, j), v in want.items():
        assert C6[i, j] == pytest.approx(v, rel=1e-9), (i, j)
        assert Cb[1][i, j] == pytest.approx(v, rel=1e-9), (i, j)
