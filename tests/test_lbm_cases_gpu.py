"""LBM parity on the benchmark's own workloads and on closed forms (GPU).

* The TRT two-phase sphere pack at its full size (64^3, 2000 steps + 2
  pre-steps, magic parameter 1/4): FP64 storage bit-identical to the oracle,
  check_state to reduction rounding.
* The MRT two-phase pack with 1:10 viscosities and a 45 degree contact angle
  (the headline workload's physics at 48^3): FP64 bit-identical to the
  oracle over a short window; FP32 storage against FP64 over 2000 steps.
* Closed forms that hold independently of the reference: one step at
  tau = 1 with every rate 1 (TRT with magic parameter 1/4, or MRT with S = 1)
  returns f_eq + Fp streamed; the body-force slit between half-way
  bounce-back walls with the reference's TRT (magic parameter 3/16,
  collmatrTRT, lsm_lbm_static_alpha_beta_multipore.f:2858-2893) converges
  to the discrete Poiseuille profile.
"""
import numpy as np
import pytest

import oracle_bind as ob
from closed_forms import bgk_case, bgk_one_step, slit_case
from porescale_amd import cases
from porescale_amd.lbm import LBM

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _diag(d):
    return {k: getattr(d, k) for k in ("permr", "permb", "tmr", "tmb", "tm", "nr", "nb")}


def test_trt_pack_full_size_bitexact_vs_oracle():
    """64^3 pack, radius 6, porosity 0.35; TRT tau = 1, magic parameter 1/4;
    blob start S_b = 0.5; 2 pre-steps then 20 x 100 steps with check_state
    after each 100.  The oracle runs the same schedule with the same rates."""
    ind, inp, coll, S, blue = cases.twophase_trt_pack(64)
    state = cases.initial_state(inp, ind, blue)
    _say("case twophase TRT pack 64^3, 2002 steps fp64 (oracle)")
    o = ob.OracleLBM(inp, ind, S=S)
    o.set_state(*state)
    o.step(2)
    orows = []
    for it1 in range(1, 21):
        o.step(100)
        orows.append(list(o.check_state(it1, 100).outm))
    ofr, ofb = o.state()
    od = _diag(o.check_state(20, 100))
    _say(f"case twophase TRT pack 64^3 fp64 (GPU), nl={o.nl}")
    with LBM(inp, ind, collision=coll) as lb:
        np.testing.assert_array_equal(lb.rates()[0], S[0])
        lb.set_state(*state)
        rows = lb.run(20, 100)
        fr, fb = lb.get_state()
        d = _diag(lb.check_state(20, 100))
    assert np.array_equal(fr, ofr) and np.array_equal(fb, ofb)
    np.testing.assert_allclose(rows, np.array(orows), rtol=2e-7, atol=1e-14, equal_nan=True)
    for k in d:
        assert d[k] == pytest.approx(od[k], rel=1e-11, abs=1e-300), k
    # the blob start keeps the blue saturation near 0.5 and the red phase
    # flowing with the force (the blue flux of this transient is not
    # single-signed: blobs are pushed back by capillary forces)
    assert 0.4 < d["nb"] / (d["nr"] + d["nb"]) < 0.6 and d["permr"] > 0


def test_mrt_pack_fp64_bitexact_vs_oracle():
    """The headline workload's physics (MRT, viscosity ratio 1:10, wet
    0.7071, blob start) at 48^3: FP64 storage bit-identical to the oracle
    after 40 steps, check_state to reduction rounding."""
    ind, inp, coll, S, blue = cases.twophase_mrt_pack(48, radii=(2, 5), seed=3)
    state = cases.initial_state(inp, ind, blue)
    o = ob.OracleLBM(inp, ind, S=S)
    o.set_state(*state)
    o.step(40)
    _say(f"case twophase MRT pack 48^3 fp64, nl={o.nl}")
    with LBM(inp, ind, collision=coll) as lb:
        Sr, Sb = lb.rates()
        np.testing.assert_array_equal(Sr, S[0])
        np.testing.assert_array_equal(Sb, S[1])
        lb.set_state(*state)
        lb.step(40)
        fr, fb = lb.get_state()
        d = _diag(lb.check_state(1, 40))
    ofr, ofb = o.state()
    assert np.array_equal(fr, ofr) and np.array_equal(fb, ofb)
    od = _diag(o.check_state(1, 40))
    for k in d:
        assert d[k] == pytest.approx(od[k], rel=1e-11, abs=1e-300), k


@pytest.mark.parametrize("prec32", ["f32", "f32x"])
def test_mrt_pack_fp32_storage_tracks_fp64(prec32):
    """FP32 storage (total deviation + minority-colour fraction to 25 bits,
    or 33 with f32x) against FP64 storage on the 48^3 MRT workload over 2000
    steps, and the control: FP64 from the same start with every population
    scaled by 1 + 1e-7 u (u uniform in [-1, 1]), i.e. how far a perturbation
    of FP32 storage size moves this transient by itself.  Bounds: total mass to
    1e-10, each colour's mass, the blue saturation and the blue permeability
    to 1e-5 relative; the red permeability no further from FP64 than 1e-5 or
    three times the control's own spread, whichever is larger."""
    ind, inp, coll, S, blue = cases.twophase_mrt_pack(48, radii=(2, 5), seed=3)
    fr0, fb0 = cases.initial_state(inp, ind, blue)
    rng = np.random.default_rng(0)
    pert = (fr0 * (1.0 + 1e-7 * rng.uniform(-1.0, 1.0, fr0.shape)),
            fb0 * (1.0 + 1e-7 * rng.uniform(-1.0, 1.0, fb0.shape)))
    out = {}
    for name, prec, st in (("f64", "f64", (fr0, fb0)), ("ctl", "f64", pert), ("f32", prec32, (fr0, fb0))):
        _say(f"case twophase MRT pack 48^3 2000 steps {name}")
        with LBM(inp, ind, collision=coll, precision=prec) as lb:
            lb.set_state(*st)
            lb.step(2000)
            d = lb.check_state(1, 2000)
            out[name] = {"permr": d.permr, "permb": d.permb, "sat": d.outm[11], "tmr": d.tmr, "tmb": d.tmb,
                         "tm": d.tm}
    ref, f32, ctl = out["f64"], out["f32"], out["ctl"]
    rel = {k: abs(f32[k] - ref[k]) / abs(ref[k]) for k in ref}
    rel_ctl = {k: abs(ctl[k] - ref[k]) / abs(ref[k]) for k in ref}
    _say(f"fp32 rel {rel}")
    _say(f"control rel {rel_ctl}")
    assert rel["tm"] < 1e-10
    for k in ("tmr", "tmb", "sat", "permb"):
        assert rel[k] <= 1e-5, (k, rel[k])
    assert rel["permr"] <= max(1e-5, 3.0 * rel_ctl["permr"])


@pytest.mark.parametrize("rates", ["trt_quarter_tau1", "mrt_all_one"])
@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_bgk_limit_one_step_is_feq_plus_force(rates, precision):
    """tau = 1 with magic parameter 1/4 makes every TRT rate 1, and MRT with
    S = 1 is the same collision: one step must return f_eq + Fp streamed
    (closed_forms.bgk_one_step), whatever the incoming populations."""
    ind, inp, fr, fb = bgk_case()
    if rates == "trt_quarter_tau1":
        coll = ("trt", 0.25)
    else:
        coll = ("mrt", np.ones(19), np.ones(19))
    want = bgk_one_step(inp, fr)
    _say(f"case BGK limit {rates} {precision}")
    with LBM(inp, ind, collision=coll, precision=precision) as lb:
        np.testing.assert_allclose(lb.rates()[0][[1, 2, 4, 6, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18]], 1.0,
                                   rtol=0, atol=1e-15)
        lb.set_state(fr, fb)
        lb.step(1)
        gr, gb = lb.get_state()
    assert np.all(gb == 0.0)
    np.testing.assert_allclose(gr, want, rtol=0, atol=1e-15 if precision == "f64" else 4e-8)


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_poiseuille_slit_discrete_closed_form(precision):
    """Single-phase body-force slit between half-way bounce-back walls with
    the reference's TRT (magic parameter 3/16): the steady permeability is
    that of the discrete parabola, k = (H/nx)(H^2/12 + 1/24)
    (closed_forms.slit_case); a wall displaced by d would move it by ~d/H."""
    ind, inp, k = slit_case()
    _say(f"case Poiseuille slit {precision}")
    with LBM(inp, ind, precision=precision) as lb:
        lb.initdistr(0.0)
        lb.step(6000)
        d = lb.check_state(1, 6000)
    assert d.permr == pytest.approx(k, rel=1e-9 if precision == "f64" else 1e-5)
