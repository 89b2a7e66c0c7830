"""LBM HIP kernels against the oracle (GPU).

FP64 storage must be bit-identical to the oracle (which is bit-identical to
the flang-compiled reference): states after N steps, tables, and check_state
to reduction rounding.  FP32 storage is held to stated tolerances.
"""
import os

import numpy as np
import pytest

import oracle_bind as ob
from golden.make_golden import lbm_cases
from porescale_amd import geometry as geo
from porescale_amd import io as pio
from porescale_amd.lbm import LBM, mrt_rates, trt_rates

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _say(msg):
    print(msg, flush=True)


def test_fp64_icon0_bitexact_vs_golden():
    (ind, inp), _ = lbm_cases()
    g = np.load(os.path.join(GOLD, "lbm_icon0.npz"))
    _say("case lbm_icon0 12^3 fp64")
    with LBM(inp, ind) as lb:
        lb.initdistr(inp.sat)
        lb.step(int(g["steps"]))
        fr, fb = lb.get_state()
    assert np.array_equal(fr, g["fr"]) and np.array_equal(fb, g["fb"])


def test_fp64_blobs_bitexact_and_check_state_vs_golden():
    _, (ind, inp, blue) = lbm_cases()
    g = np.load(os.path.join(GOLD, "lbm_blobs.npz"))
    _say("case lbm_blobs 18x16x14 fp64 wetting")
    with LBM(inp, ind) as lb:
        lb.set_state(*ob.equilibrium_state(inp, ob.fluid_order_mask(ind, blue)))
        rows = lb.run(inp.itm, inp.itw)
        fr, fb = lb.get_state()
    assert np.array_equal(fr, g["fr"]) and np.array_equal(fb, g["fb"])
    np.testing.assert_allclose(rows, g["outm"], rtol=2e-7, atol=1e-14, equal_nan=True)


def test_tables_identical_to_oracle():
    ind = geo.sphere_pack((11, 9, 13), 0.4, 2.5, seed=4)
    nz, ny, nx = ind.shape
    inp = pio.LBMInput(ncx=nx, ncy=ny, ncz=nz)
    o = ob.OracleLBM(inp, ind)
    with LBM(inp, ind) as lb:
        got = lb.tables()
        assert (lb.nl, lb.nps, lb.npb2, lb.npb4, lb.ns) == tuple(o.counts().values())
    for a, b in zip(got, o.tables()):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("variant", ["trt_quarter", "mrt", "idir1_ratio"])
def test_fp64_collision_variants_bitexact(variant):
    ind = geo.sphere_pack((12, 10, 14), 0.5, 2.5, seed=7)
    nz, ny, nx = ind.shape
    inp = pio.LBMInput(B=0.7, sig=0.03, vir=1 / 6, vib=1 / 6, wet=0.4, idir=3, sat=0.45, ncx=nx, ncy=ny,
                       ncz=nz, dt=1.0, bfr=[0, 0, 1e-5], bfb=[0, 0, 1e-5])
    coll, S = "ref", None
    if variant == "trt_quarter":
        Sr = trt_rates(inp.vir, 0.25)
        coll, S = ("trt", 0.25), (Sr, Sr)
    elif variant == "mrt":
        inp.vib = 1.0 / 60.0
        Sr, Sb = mrt_rates(inp.vir), mrt_rates(inp.vib)
        coll, S = ("mrt", Sr, Sb), (Sr, Sb)
    else:
        inp.idir, inp.bfr, inp.bfb, inp.vib, inp.cr2 = 1, [1e-5, 2e-6, 0], [3e-5, 0, 1e-6], 0.05, 0.3
    _say(f"case variant {variant}")
    o = ob.OracleLBM(inp, ind, S=S)
    o.init_random(inp.sat)
    o.step(25)
    with LBM(inp, ind, collision=coll) as lb:
        np.testing.assert_array_equal(lb.rates()[0], o.S()[0])
        lb.initdistr(inp.sat)
        lb.step(25)
        fr, fb = lb.get_state()
        d = lb.check_state(1, 25)
    ofr, ofb = o.state()
    assert np.array_equal(fr, ofr) and np.array_equal(fb, ofb)
    od = o.check_state(1, 25)
    for k in ("permr", "permb", "tmr", "tmb", "nr", "nb", "Cpr", "Cpb"):
        assert getattr(d, k) == pytest.approx(getattr(od, k), rel=1e-11, abs=1e-300, nan_ok=True)
    np.testing.assert_allclose(d.Qr[:], od.Qr[:], rtol=1e-10, atol=1e-18)


def test_single_phase_mass_conservation_and_rest_state():
    ind = geo.sphere_pack((16, 16, 16), 0.5, 3.0, seed=9)
    inp = pio.LBMInput(ncx=16, ncy=16, ncz=16, bfr=[0, 0, 0], bfb=[0, 0, 0], sat=0.0, sig=0.0)
    with LBM(inp, ind, precision="f64") as lb:
        lb.initdistr(0.0)
        fr0, _ = lb.get_state()
        lb.step(50)
        fr, fb = lb.get_state()
    # at rest with no force the equilibrium is a fixed point
    np.testing.assert_allclose(fr, fr0, rtol=0, atol=1e-15)
    assert np.all(fb == 0.0)


@pytest.mark.parametrize("blobs", [False, True])
def test_fp32_storage_tracks_fp64(blobs):
    ind = geo.sphere_pack((24, 24, 24), 0.4, 4.0, seed=5)
    inp = pio.LBMInput(B=0.7, sig=0.01, vir=1 / 6, vib=1 / 6, wet=0.0, idir=3, ncx=24, ncy=24, ncz=24,
                       dt=1.0, bfr=[0, 0, 1e-5], bfb=[0, 0, 1e-5])
    state = ob.equilibrium_state(inp, ob.fluid_order_mask(ind, geo.phase_blobs(ind, 0.5 if blobs else 0.0, 1)))
    out = {}
    for prec in ("f64", "f32"):
        _say(f"case fp32 vs fp64 blobs={blobs} {prec}")
        with LBM(inp, ind, precision=prec) as lb:
            lb.set_state(*state)
            lb.step(300)
            out[prec] = (lb.check_state(1, 300), lb.get_state())
    d64, d32 = out["f64"][0], out["f32"][0]
    # exact zeros of the minority phase survive FP32 storage
    assert d32.tmb == 0.0 if not blobs else d32.tmb > 0
    assert abs(d32.tm - d64.tm) / d64.tm < 1e-6
    # a two-phase transient amplifies storage rounding through the interface
    # dynamics; single-phase flow is held to the storage precision
    assert d32.permr == pytest.approx(d64.permr, rel=2e-3 if blobs else 1e-5)
    assert abs(d32.nr - d64.nr) <= 0.002 * (d64.nr + d64.nb)
    if not blobs:
        np.testing.assert_allclose(out["f32"][1][0], out["f64"][1][0], rtol=0, atol=2e-8)


def test_profile_reports_kernel_times():
    ind = geo.sphere_pack((32, 32, 32), 0.35, 6.0, seed=0)
    inp = pio.LBMInput(ncx=32, ncy=32, ncz=32)
    with LBM(inp, ind, precision="f32") as lb:
        lb.initdistr(0.5)
        tot, tc, td = lb.profile(10)
    assert tot > 0 and tc > 0 and td > 0 and tc + td <= tot * 1.01


@pytest.mark.parametrize("prec32", ["f32", "f32x"])
def test_fp32_state_round_trip_and_first_step(prec32):
    """FP32 storage holds post-collision populations ahead of the pull; set/get
    must invert each other to storage precision (each population's colour
    fraction to its word's 25 significand bits, or 33 with the extension
    byte), and one step must match the FP64 path to storage precision."""
    ind = geo.sphere_pack((14, 12, 10), 0.45, 2.5, seed=13)
    nz, ny, nx = ind.shape
    inp = pio.LBMInput(B=0.7, sig=0.02, vir=0.1, vib=0.2, wet=0.3, ncx=nx, ncy=ny, ncz=nz, dt=1.0,
                       bfr=[0, 1e-5, 0], bfb=[0, 1e-5, 0])
    state = ob.equilibrium_state(inp, ob.fluid_order_mask(ind, geo.phase_blobs(ind, 0.5, 2, radius=2.0)))
    rng = np.random.default_rng(0)
    state = (state[0] * (1 + 1e-3 * rng.standard_normal(state[0].shape)),
             state[1] * (1 + 1e-3 * rng.standard_normal(state[1].shape)))
    # colours mixed in every population, minority shares from 1/2 down to
    # 1e-12 of it, a fifth left single-coloured
    f = state[0] + state[1]
    r = 0.5 * 10.0 ** (-12.0 * rng.uniform(size=f.shape))
    r[rng.uniform(size=f.shape) < 0.2] = 0.0
    blue_minor = rng.uniform(size=f.shape) < 0.5
    state = (np.where(blue_minor, f - r * f, r * f), np.where(blue_minor, r * f, f - r * f))
    out = {}
    for prec in (prec32, "f64"):
        with LBM(inp, ind, precision=prec) as lb:
            lb.set_state(*state)
            back = lb.get_state()
            lb.step(1)
            out[prec] = (back, lb.get_state())
    for a, b in zip(out[prec32][0], state):
        np.testing.assert_allclose(a, b, rtol=0, atol=2e-8)
    # the minority colour's fraction of every population, whatever T's rounding
    fr, fb = state
    br, bb = out[prec32][0]
    minor_blue = np.abs(fb) < np.abs(fr)
    frac = np.where(minor_blue, fb, fr) / (fr + fb)
    frac_back = np.where(minor_blue, bb, br) / (br + bb)
    nz = frac != 0.0
    assert np.all(frac_back[~nz] == 0.0)  # single-colour populations stay exact
    bits = 33 if prec32 == "f32x" else 25
    assert np.max(np.abs(frac_back[nz] - frac[nz]) / np.abs(frac[nz])) <= 2.0 ** -bits * (1 + 1e-6)
    for a, b in zip(out[prec32][1], out["f64"][1]):
        np.testing.assert_allclose(a, b, rtol=0, atol=5e-8)


@pytest.mark.parametrize("prec32", ["f32", "f32x"])
def test_fp32_compressed_pull_table_is_exact(monkeypatch, prec32):
    """Every pull-table form addresses exactly the populations the collide
    kernel's element table does (precomputed element indices, voxel and
    coordinates; the default): with the element table off, the
    direction-major int32 table, the node-major one, and the int16 tables
    (direct offsets, flag 1; rank differences within x-planes, flag 2, the
    density kernel's form); with it on, each density-kernel form, the int16
    ones direction-major and node-major (the default).  States
    bit-identical."""
    ind = geo.sphere_pack((20, 14, 16), 0.4, 3.0, seed=17)
    nz, ny, nx = ind.shape
    inp = pio.LBMInput(B=0.7, sig=0.02, vir=0.1, vib=0.2, wet=0.3, ncx=nx, ncy=ny, ncz=nz, dt=1.0,
                       bfr=[1e-5, 0, 0], bfb=[1e-5, 0, 0])
    state = ob.equilibrium_state(inp, ob.fluid_order_mask(ind, geo.phase_blobs(ind, 0.5, 5, radius=2.5)))
    out = {}
    monkeypatch.setenv("PSX_LBM_FUSED", "0")  # the int16 table serves the two-kernel step
    forms = [(e, f, a, "1") for e in ("0", "1") for f, a in (("0", "0"), ("0", "1"), ("1", "1"), ("2", "1"))]
    forms += [("1", f, "1", "0") for f in ("1", "2")]
    for etab, flag, aos, dtab in forms:
        monkeypatch.setenv("PSX_LBM_ETAB", etab)
        monkeypatch.setenv("PSX_LBM_IDX16", flag)
        monkeypatch.setenv("PSX_LBM_SRCAOS", aos)
        monkeypatch.setenv("PSX_LBM_DTAB", dtab)
        _say(f"case {prec32} element table={etab} pull table int16={flag} node-major={aos} density table={dtab}")
        with LBM(inp, ind, precision=prec32) as lb:
            lb.set_state(*state)
            lb.step(40)
            out[etab + flag + aos + dtab] = lb.get_state()
    ref = out["1001"]
    for key, (fr, fb) in out.items():
        assert np.array_equal(ref[0], fr) and np.array_equal(ref[1], fb), key


@pytest.mark.parametrize("prec32", ["f32", "f32x"])
def test_fp32_wide_pull_offsets_are_exact(monkeypatch, prec32):
    """The collide kernel's 64-bit pull offsets (taken automatically once the
    population buffer reaches 2^31 floats, about 56 M fluid nodes; forced
    here with PSX_LBM_WIDE=1) give the same bits as the 32-bit ones, with the
    int32 and the compressed pull tables."""
    ind = geo.sphere_pack((18, 14, 16), 0.4, 3.0, seed=29)
    nz, ny, nx = ind.shape
    inp = pio.LBMInput(B=0.7, sig=0.02, vir=0.1, vib=0.2, wet=0.3, ncx=nx, ncy=ny, ncz=nz, dt=1.0,
                       bfr=[1e-5, 2e-6, 0], bfb=[0, 1e-5, 1e-6])
    state = ob.equilibrium_state(inp, ob.fluid_order_mask(ind, geo.phase_blobs(ind, 0.5, 4, radius=2.5)))
    monkeypatch.setenv("PSX_LBM_FUSED", "0")
    for idx16 in ("0", "1"):
        monkeypatch.setenv("PSX_LBM_IDX16", idx16)
        out = {}
        for wide in ("0", "1"):
            monkeypatch.setenv("PSX_LBM_WIDE", wide)
            _say(f"case {prec32} wide={wide} int16={idx16}")
            with LBM(inp, ind, precision=prec32) as lb:
                lb.set_state(*state)
                lb.step(25)
                out[wide] = lb.get_state()
        assert np.array_equal(out["0"][0], out["1"][0]) and np.array_equal(out["0"][1], out["1"][1])


@pytest.mark.parametrize("brick", ["8,8,16", "4,4,4", "3,5,7", "16,16,16"])
def test_fp32_fused_brick_step_is_exact(monkeypatch, brick):
    """The fused step (colour field of each brick and its halo in LDS, then
    collide) gives the same bits as the collide + density kernel pair, for
    bricks that do and do not divide the grid (and one larger than it)."""
    ind = geo.sphere_pack((14, 12, 18), 0.4, 3.0, seed=23)
    nz, ny, nx = ind.shape
    inp = pio.LBMInput(B=0.7, sig=0.02, vir=0.1, vib=0.2, wet=0.3, ncx=nx, ncy=ny, ncz=nz, dt=1.0,
                       bfr=[0, 1e-5, 2e-6], bfb=[1e-6, 1e-5, 0])
    state = ob.equilibrium_state(inp, ob.fluid_order_mask(ind, geo.phase_blobs(ind, 0.5, 6, radius=2.5)))
    out = {}
    monkeypatch.setenv("PSX_LBM_BRICK", brick)
    for flag in ("0", "1"):
        monkeypatch.setenv("PSX_LBM_FUSED", flag)
        _say(f"case fp32 fused={flag} brick={brick}")
        with LBM(inp, ind, precision="f32") as lb:
            lb.set_state(*state)
            lb.step(30)
            out[flag] = (lb.get_state(), lb.check_state(1, 30).outm, lb.save_fields())
    (a, da, fa), (b, db, fb_) = out["0"], out["1"]
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    np.testing.assert_allclose(da, db, rtol=1e-12, atol=1e-18, equal_nan=True)
    for k in ("velocity", "color", "pressure", "phase"):
        assert np.array_equal(fa[k], fb_[k]), k


def test_fp64_multipore_selects_iipore_bitexact():
    """Two pore families (codes 1 and 2) with iipore = 2: only code-2 voxels
    are fluid (L1:1046-1072).  The other family enters the colour stencil as
    `wet`, the defined form of L1:2673 whose reference reads cf(0) there."""
    ind = geo.sphere_pack((12, 14, 10), 0.55, 2.5, seed=21)
    fam = geo.phase_blobs(ind, 0.6, seed=4, radius=2.0)
    ind2 = np.where(fam, 2, ind).astype(np.int32)
    nz, ny, nx = ind.shape
    inp = pio.LBMInput(B=0.7, sig=0.02, vir=0.1, vib=0.2, wet=0.4, iipore=2, npores=2, sat=0.4, ncx=nx, ncy=ny,
                       ncz=nz, dt=1.0, bfr=[0, 0, 1e-5], bfb=[0, 0, 1e-5])
    _say("case multipore iipore=2")
    o = ob.OracleLBM(inp, ind2)
    o.init_random(inp.sat)
    o.step(20)
    with LBM(inp, ind2) as lb:
        assert lb.nl == int((ind2 == 2).sum()) == o.nl
        lb.initdistr(inp.sat)
        lb.step(20)
        fr, fb = lb.get_state()
    ofr, ofb = o.state()
    assert np.array_equal(fr, ofr) and np.array_equal(fb, ofb)


def test_fields_fp64_bitexact_vs_reference_files():
    """psx_lbm_fields from the blob run's final state against the reference's
    own save_results files (velocity_field, color_field, pressure3,
    ind_red_blue)."""
    _, (ind, inp, blue) = lbm_cases()
    g = np.load(os.path.join(GOLD, "lbm_blobs.npz"))
    f = np.load(os.path.join(GOLD, "lbm_blobs_fields.npz"))
    _say("case save_results fields 18x16x14 fp64")
    with LBM(inp, ind) as lb:
        lb.set_state(g["fr"], g["fb"])
        got = lb.save_fields()
    for k in ("velocity", "color", "pressure", "phase"):
        assert np.array_equal(got[k], f[k]), k


def test_run_writes_save_results_files(tmp_path):
    """LBM.run(save_dir=...) writes the reference's files each its-th
    iteration; the last ones equal the reference's."""
    _, (ind, inp, blue) = lbm_cases()
    f = np.load(os.path.join(GOLD, "lbm_blobs_fields.npz"))
    _say("case run() with save_results 18x16x14 fp64")
    with LBM(inp, ind) as lb:
        lb.set_state(*ob.equilibrium_state(inp, ob.fluid_order_mask(ind, blue)))
        lb.run(inp.itm, inp.itw, save_dir=str(tmp_path), its=1)
    back = ob.read_field_files(str(tmp_path), inp.itm, ind.shape)
    for k in ("velocity", "color", "pressure", "phase"):
        assert np.array_equal(back[k], f[k]), k
    assert sorted(p.name for p in tmp_path.glob("pressure*")) == [f"pressure{i}" for i in range(1, inp.itm + 1)]


def test_fields_after_gpu_steps_match_oracle():
    """Fields after 25 GPU steps: FP64 bit-identical to the oracle's
    save_results; FP32 storage from the same start within storage precision."""
    ind = geo.sphere_pack((20, 18, 16), 0.45, 3.0, seed=31)
    nz, ny, nx = ind.shape
    inp = pio.LBMInput(B=0.7, sig=0.02, vir=0.1, vib=0.2, wet=-0.4, idir=1, ncx=nx, ncy=ny, ncz=nz, dt=1.0,
                       bfr=[2e-5, 0, 1e-6], bfb=[1e-5, 0, 0])
    state = ob.equilibrium_state(inp, ob.fluid_order_mask(ind, geo.phase_blobs(ind, 0.45, 8, radius=2.5)))
    o = ob.OracleLBM(inp, ind)
    o.set_state(*state)
    o.step(25)
    want = o.save_fields()
    _say("case fields after 25 steps 20x18x16 fp64")
    with LBM(inp, ind) as lb:
        lb.set_state(*state)
        lb.step(25)
        got = lb.save_fields()
    for k in ("velocity", "color", "pressure", "phase"):
        assert np.array_equal(got[k], want[k]), k
    _say("case fields from the oracle state 20x18x16 fp32")
    with LBM(inp, ind, precision="f32") as lb:
        lb.set_state(*o.state())
        g32 = lb.save_fields()
    # FP32 storage rounds each population to ~6e-8 relative (|f| < 0.4)
    np.testing.assert_allclose(g32["velocity"], want["velocity"], rtol=0, atol=1e-7)
    np.testing.assert_allclose(g32["color"], want["color"], rtol=0, atol=1e-6)
    np.testing.assert_allclose(g32["pressure"], want["pressure"], rtol=0, atol=1e-7)
    cf = want["color"]
    near = (np.abs(cf - 0.9) < 1e-5) | (np.abs(cf + 0.9) < 1e-5) | (np.abs(cf) < 1e-5)
    assert np.array_equal(g32["phase"][~near], want["phase"][~near])


def _diag_arrays(d):
    return np.array([d.permr, d.permb, *d.Qr, *d.Qb, d.nr, d.nb, d.tmr, d.tmb, d.tm, *d.outm])


@pytest.mark.parametrize("shape, kind", [((5, 6, 7), "solid"), ((9, 8, 10), "fluid"), ((3, 3, 3), "pack")])
def test_edge_domains_match_oracle(shape, kind):
    """Edge geometries, FP64 bit-identical to the oracle: an all-solid domain
    (no fluid node: every kernel has an empty grid), a domain with no solid
    (every link streams, through the periodic wrap on all three axes, no
    bounce-back) and the smallest grid the stencil allows (3 voxels a side,
    where a node's -1 and +1 neighbours along an axis are one another's)."""
    nz, ny, nx = shape
    if kind == "solid":
        ind = np.zeros(shape, np.int32)
    elif kind == "fluid":
        ind = np.ones(shape, np.int32)
    else:
        ind = np.ones(shape, np.int32)
        ind[1, 1, 1] = 0
        ind[0, 2, 1] = 0
    inp = pio.LBMInput(B=0.7, sig=0.02, vir=0.1, vib=0.2, wet=0.3, ncx=nx, ncy=ny, ncz=nz, dt=1.0,
                       bfr=[1e-5, 2e-6, 0], bfb=[0, 1e-5, 3e-6])
    blobs = geo.phase_blobs(ind, 0.5, 3, radius=1.5) if kind != "solid" else np.zeros(shape, bool)
    state = ob.equilibrium_state(inp, ob.fluid_order_mask(ind, blobs))
    o = ob.OracleLBM(inp, ind)
    o.set_state(*state)
    o.step(12)
    want, dwant, fwant = o.state(), _diag_arrays(o.check_state(1, 12)), o.save_fields()
    _say(f"case edge domain {kind} {shape} fp64, nl={o.nl}")
    with LBM(inp, ind) as lb:
        lb.set_state(*state)
        lb.step(12)
        got, dgot, fgot = lb.get_state(), _diag_arrays(lb.check_state(1, 12)), lb.save_fields()
    assert got[0].shape == want[0].shape == (o.nl, 19)
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
    np.testing.assert_allclose(dgot, dwant, rtol=1e-12, atol=1e-18, equal_nan=True)
    for k in ("velocity", "color", "pressure", "phase"):
        assert np.array_equal(fgot[k], fwant[k]), k
    _say(f"case edge domain {kind} {shape} fp32")
    with LBM(inp, ind, precision="f32") as lb:
        lb.set_state(*state)
        lb.step(12)
        g32 = lb.get_state()
    assert g32[0].shape == (o.nl, 19)
    np.testing.assert_allclose(g32[0], want[0], rtol=0, atol=5e-8)
    np.testing.assert_allclose(g32[1], want[1], rtol=0, atol=5e-8)
