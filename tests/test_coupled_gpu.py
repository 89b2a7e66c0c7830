"""The coupled LBM -> LSM model of L1 on the GPU (fluid force of the
bounce-back links, lsm_step, calc_ab) against the reference's own files and
the oracle."""
import os

import numpy as np
import pytest

import oracle_bind as ob
from golden.make_golden import coupled_case
from porescale_amd.lbm import LBM, Coupled

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _say(msg):
    print(msg, flush=True)


def _run(precision="f64"):
    ind, inp = coupled_case()
    lb = LBM(inp, ind, precision=precision)
    lb.initdistr(inp.sat)
    lb.step(2)
    cp = Coupled(lb)
    cp.step(inp.itm * inp.itw)
    return ind, inp, lb, cp


def _grid(g, key, mijks):
    sol = mijks > 0
    out = np.zeros((int(mijks.max()), 3))
    out[mijks[sol] - 1] = g[key][sol]
    return out


def test_coupled_fp64_vs_reference_files():
    """2 pre-steps + 12 coupled steps: the fluid force is the reference's to
    the bit (same links, same summation order); the solid state, whose
    spring forces are summed per node here and per bond there, to 1e-12 of
    its scale."""
    _say("case coupled 14x10x12 fp64 vs L1 files")
    ind, inp, lb, cp = _run()
    g = np.load(os.path.join(GOLD, "lbm_coupled.npz"))
    fr, fb = lb.get_state()
    assert np.array_equal(fr, g["fr"]) and np.array_equal(fb, g["fb"])
    r, v, a, af = cp.get_state()
    mijks = ob.OracleCoupled(inp, ind).mijks()
    assert cp.ns == int(mijks.max())
    rsi = 1.0 / inp.rs
    assert np.array_equal(af / rsi, _grid(g, "lsm_fluid_force", mijks))
    for arr, key in ((r, "lsm_disp"), (v, "lsm_vel"), (a / rsi, "lsm_force")):
        want = _grid(g, key, mijks)
        np.testing.assert_allclose(arr, want, rtol=0, atol=1e-12 * np.abs(want).max())
    ab = cp.calc_ab()
    for k in ("alphai", "betai"):
        want = g[f"ab_{k}"]
        np.testing.assert_allclose(ab[k], want, rtol=0, atol=1e-10 * np.abs(want).max())
    assert ab["ncub"] == int(g["ab_ncub"]) and ab["epsi"] == g["ab_epsi"]
    cp.close()
    lb.close()


def test_coupled_calc_ab_equals_oracle_on_same_state():
    """calc_ab on the GPU and in the oracle from the same displacement: the
    per-cube terms are evaluated in the reference's order, the sums over
    cubes and faces in a tree, so the totals agree to summation rounding;
    counts exactly."""
    _say("case coupled calc_ab vs oracle")
    ind, inp, lb, cp = _run()
    r, v, a, af = cp.get_state()
    o = ob.OracleCoupled(inp, ind)
    o.set_state(r, v, a)
    want = o.calc_ab()
    got = cp.calc_ab()
    for k in ("asigmap", "asigmam", "astrain", "alphai", "betai"):
        w = np.asarray(want[k])
        np.testing.assert_allclose(np.asarray(got[k]), w, rtol=0, atol=1e-13 * np.abs(w).max())
    assert got["epsi"] == want["epsi"] and got["ncub"] == want["ncub"]
    assert np.array_equal(got["asigmam"], -got["asigmap"])
    cp.close()
    lb.close()


def test_coupled_state_round_trip_and_rest():
    """set/get state invert each other; a solid at rest under a fluid at rest
    with no force feels only the pressure term and its spring forces stay
    balanced by it (total fluid force zero on a periodic slab)."""
    ind = np.ones((8, 6, 10), np.int32)
    ind[:, :, 3:6] = 0
    from porescale_amd import io as pio
    nz, ny, nx = ind.shape
    inp = pio.LBMInput(ncx=nx, ncy=ny, ncz=nz, bfr=[0, 0, 0], bfb=[0, 0, 0], Em=0.2, nu=0.2, rs=2.0, sat=0.0)
    _say("case coupled slab at rest")
    with LBM(inp, ind) as lb:
        lb.initdistr(0.0)
        with Coupled(lb) as cp:
            rng = np.random.default_rng(1)
            st = [rng.standard_normal((cp.ns, 3)) for _ in range(3)]
            cp.set_state(*st)
            back = cp.get_state()
            for x, y in zip(back[:3], st):
                assert np.array_equal(x, y)
            cp.set_state(*(np.zeros((cp.ns, 3)) for _ in range(3)))
            cp.step(3)
            r, v, a, af = cp.get_state()
    o = ob.OracleCoupled(inp, ind)
    o.lbm.init_random(0.0)
    o.step(3)
    orr, ov, oa, oaf = o.state()
    assert np.array_equal(af, oaf)
    assert np.abs(af.sum(0)).max() < 1e-14
    np.testing.assert_allclose(r, orr, rtol=0, atol=1e-13 * max(np.abs(orr).max(), 1e-300))


def test_coupled_fp32_tracks_fp64():
    """FP32 populations: the fluid force within storage precision of the
    FP64 run, the solid displacement close to it."""
    _say("case coupled fp32 vs fp64")
    _, _, lb64, cp64 = _run("f64")
    _, _, lb32, cp32 = _run("f32")
    r64, _, _, af64 = cp64.get_state()
    r32, _, _, af32 = cp32.get_state()
    np.testing.assert_allclose(af32, af64, rtol=0, atol=1e-4 * np.abs(af64).max())
    np.testing.assert_allclose(r32, r64, rtol=0, atol=1e-4 * np.abs(r64).max())
    for x in (cp64, cp32, lb64, lb32):
        x.close()
