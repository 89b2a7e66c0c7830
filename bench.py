#!/usr/bin/env python3
"""Benchmark of the two porescale hot paths on one MI355X.

Headline (the JSON line's metric/value): the D3Q19 two-phase colour-gradient
LBM step with MRT collision at 256^3 (sphere pack, radii 4..10, porosity
0.30, viscosity ratio 1:10, wetting 0.7071, blob saturation 0.5), FP32
population storage with the colour fraction to 33 bits (f32x) and FP64
arithmetic: MLUPS over fluid nodes.  The step is one pull+collide launch and
one pull+density launch; the roofline entry is the collide kernel, timed
with HIP events on the library's own stream.

Also reported (objects in the same line):
  lbm_f32      the same LBM workload with the 25-bit colour fraction (f32):
               ms per step and MLUPS.
  lbm_f64      the same LBM workload with FP64 storage (states bit-identical
               to the reference): ms per step and MLUPS.
  trt_pack_64  the reference method as configured: 64^3 radius-6 pack, FP64,
               TRT tau = 1 with magic parameter 1/4, 2 + 2000 steps: MLUPS,
               k_r, k_b and the blue saturation.
  precision    f32 and f32x storage against FP64 on the headline workload
               over 20000 steps, with the FP64 control started from a 1e-7
               perturbation: permeabilities, saturation, colour masses at the
               end, averaged over the second half and at worst over 40
               samples.
  coupled      L1's coupled model on the same 256^3 workload: ms per main-loop
               step (LBM + momentum exchange + lsm_step) and calc_ab.
  lsm          64^3 single-solid sphere pack (porosity 0.35), all six load
               cases at rtol 1e-10 and 1e-5: plain FP64 PCG one case at a time
               versus the batched FP32 multigrid-PCG / FP64-refinement solver,
               with the moduli error; the operator's roofline at 64^3 and
               128^3; 64 uneven two-component sub-volumes in one batch against
               plain FP64 on every volume; adjoint gradients of (K, G) at
               128^3 with a lognormal stiffness field, and a 50-step
               Gauss-Newton calibration on one handle.
  cpu_baseline the oracle (C restatement of the reference, OpenMP over up to
               16 host threads) timed on this host over a bounded sample of
               the headline workload; beside it the reference programs
               themselves (flang builds of L1 and XX_P.f, one core) on the
               64^3 cases, and the LSM oracle.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-lsm] [--no-cpu] [--no-coupled]
                    [--no-precision] [--calib-steps N] [--lbm-storage f32x|f32]
    python bench.py --lsm-only         # the LSM rows alone (profiling)
    python bench.py --lsm-matvec       # the LSM operator alone at 64^3 and 128^3 (profiling)
    python bench.py --precision-steps N  # FP32 storage against FP64 over N steps
    python bench.py --calib-copy MIB   # plain float32 elementwise kernel (torch.neg)
                                       # over MIB MiB in and out, ten launches:
                                       # scale factor for FETCH_SIZE/WRITE_SIZE
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def lbm_workload(n=256):
    from porescale_amd import cases
    ind, inp, coll, S, blue = cases.twophase_mrt_pack(n)
    return ind, inp, blue


def initial_state(inp, ind, blue):
    from porescale_amd import cases
    return cases.initial_state(inp, ind, blue)


# algorithmic HBM bytes per fluid node and step of the collide kernel: the
# 19 pulled populations read and the node's own 19 written, each as T and
# colour word (4 + 4 B) plus the extension byte (f32x), or as fr and fb
# (8 + 8 B, f64); plus the colour field, 2 x 4 B (f32, f32x) or 2 x 8 B (f64)
ALG_BYTES = {"f32": 2 * 19 * 8 + 8.0, "f32x": 2 * 19 * 9 + 8.0, "f64": 2 * 19 * 16 + 16.0}


def run_lbm(steps, warmup, before=None, prec="f32x"):
    """The headline: ``warmup`` untimed steps, then ``steps`` steps with HIP
    events around every kernel.  ``before(ind, inp, blue)`` runs once the
    handle holds its start state, just ahead of the warm-up (the other
    storage legs: their GPU work keeps the clocks up while the headline's
    host set-up would otherwise leave the device idle); its result is
    returned."""
    from porescale_amd.lbm import LBM, mrt_rates
    ind, inp, blue = lbm_workload()
    fr0, fb0 = initial_state(inp, ind, blue)
    coll = ("mrt", mrt_rates(inp.vir), mrt_rates(inp.vib))
    out = {}
    with LBM(inp, ind, collision=coll, precision=prec) as lb:
        lb.set_state(fr0, fb0)
        if before is not None:
            out["before"] = before(ind, inp, blue)
        lb.step(warmup)
        lb.sync()
        t0 = time.perf_counter()
        tot, tc, td = lb.profile(steps)   # events around every kernel, same stream
        wall = time.perf_counter() - t0
        d = lb.check_state(1, steps)
        out.update(nl=lb.nl, tot_ms=tot, coll_ms=tc, dens_ms=td, wall_s=wall, permr=d.permr, permb=d.permb,
                   sat_b=d.outm[11])
    return out, ind, inp, blue


STORAGE = {"f64": "f64 (bit-identical to the reference)",
           "f32": "f32 (FP32 total deviation + minority-colour fraction to 25 bits)",
           "f32x": "f32x (FP32 total deviation + minority-colour fraction to 33 bits: a 32-bit word and a byte)"}


def run_lbm_storage(ind, inp, blue, prec, steps=50, warmup=5):
    """The same workload in another storage: FP64, whose states are
    bit-identical to the reference's (what exact parity costs), or FP32
    storage with the other colour-fraction width."""
    from porescale_amd.lbm import LBM, mrt_rates
    coll = ("mrt", mrt_rates(inp.vir), mrt_rates(inp.vib))
    with LBM(inp, ind, collision=coll, precision=prec) as lb:
        lb.set_state(*initial_state(inp, ind, blue))
        lb.step(warmup)
        lb.sync()
        tot, tc, td = lb.profile(steps)
        nl = lb.nl
    alg = ALG_BYTES[prec]
    return {"storage": STORAGE[prec], "steps": steps, "ms_per_step": tot / steps,
            "collide_ms": tc, "density_ms": td, "mlups": nl * steps / (tot * 1e-3) / 1e6,
            "algorithmic_bytes_per_node": alg, "collide_gbs": alg * nl / (tc * 1e-3) / 1e9}


def run_other_storage(headline):
    """The storage legs other than the headline's, for run_lbm's ``before``."""
    def before(ind, inp, blue):
        return {p: run_lbm_storage(ind, inp, blue, p) for p in ("f64", "f32", "f32x") if p != headline}
    return before


def run_coupled(ind, inp, blue, steps=20):
    """L1's coupled model on the same 256^3 workload: each main-loop step is
    the FP32 LBM step, the bounce-back momentum exchange and lsm_step of the
    solid lattice (FP64, periodic node grid); stable springs Em 0.2, nu 0.2,
    rs 2.  Also times calc_ab (host, the reference's loops) once."""
    import dataclasses
    import numpy as np
    from porescale_amd.lbm import LBM, Coupled, mrt_rates
    inp2 = dataclasses.replace(inp, Em=0.2, nu=0.2, rs=2.0)
    coll = ("mrt", mrt_rates(inp.vir), mrt_rates(inp.vib))
    with LBM(inp2, ind, collision=coll, precision="f32x") as lb:
        lb.set_state(*initial_state(inp2, ind, blue))
        lb.step(2)
        with Coupled(lb) as cp:
            cp.step(2)
            lb.sync()
            t0 = time.perf_counter()
            cp.step(steps)
            lb.sync()
            t_c = time.perf_counter() - t0
            t0 = time.perf_counter()
            lb.step(steps)
            lb.sync()
            t_l = time.perf_counter() - t0
            r, v, a, af = cp.get_state()
            t0 = time.perf_counter()
            ab = cp.calc_ab()
            t_ab = time.perf_counter() - t0
            ns = cp.ns
    return {"workload": "coupled lbm-d3q19 256^3 + solid lattice (Em 0.2, nu 0.2, rs 2)", "solid_nodes": ns,
            "steps": steps, "ms_per_step": t_c / steps * 1e3, "lbm_ms_per_step": t_l / steps * 1e3,
            "solid_ms_per_step": (t_c - t_l) / steps * 1e3, "calc_ab_s": t_ab,
            "max_disp": float(np.abs(r).max()), "alphai_diag": np.diag(ab["alphai"]).tolist(),
            "betai": ab["betai"].tolist()}


def run_trt_pack(n=64, steps=2000):
    """The reference method as configured: FP64 storage (bit-identical to
    the reference), TRT with tau = 1 and magic parameter 1/4 on the 64^3
    radius-6 pack (porosity 0.35), A = 0.01, recolouring 0.7, body force
    1e-5 along z, blob start S_b = 0.5; 2 pre-steps then ``steps`` steps.
    Outputs of check_state (L1:3181-3182, 3280) after the last step."""
    from porescale_amd import cases
    from porescale_amd.lbm import LBM
    ind, inp, coll, S, blue = cases.twophase_trt_pack(n)
    out = {"workload": f"lbm-d3q19-twophase-trt-{n}^3 (tau 1, magic 1/4, radius-6 pack, porosity 0.35)",
           "steps": 2 + steps}
    for prec in ("f64", "f32", "f32x"):
        with LBM(inp, ind, collision=coll, precision=prec) as lb:
            lb.set_state(*cases.initial_state(inp, ind, blue))
            lb.step(2)
            lb.sync()
            tot, tc, td = lb.profile(steps)
            d = lb.check_state(1, steps)
            out[prec] = {"ms_per_step": tot / steps, "mlups": lb.nl * steps / (tot * 1e-3) / 1e6,
                         "permr": d.permr, "permb": d.permb, "sat_b": d.outm[11]}
            out["fluid_nodes"] = lb.nl
    out["storage"] = {p: STORAGE[p] for p in ("f64", "f32", "f32x")}
    return out


def run_precision(steps=20000, every=500, window_from=10000):
    """FP32 storage against FP64 on the headline workload over ``steps``
    steps: permeabilities, blue saturation and colour masses of four runs:
    FP64, FP32 (25-bit colour fraction) and FP32X (33-bit colour fraction)
    from the same start, and, as the control, FP64 from that start with
    every population scaled by 1 + 1e-7 u (u uniform in [-1, 1], seed 0):
    how far a perturbation of the size of FP32 storage rounding alone moves
    the observables of this transient.  Sampled every ``every`` steps;
    reported at the last step and as the mean over the samples from
    ``window_from`` on (the transient's permeabilities fluctuate)."""
    import numpy as np
    from porescale_amd.lbm import LBM, mrt_rates
    ind, inp, blue = lbm_workload()
    fr0, fb0 = initial_state(inp, ind, blue)
    coll = ("mrt", mrt_rates(inp.vir), mrt_rates(inp.vib))
    rng = np.random.default_rng(0)
    fr1 = fr0 * (1.0 + 1e-7 * rng.uniform(-1.0, 1.0, fr0.shape))
    fb1 = fb0 * (1.0 + 1e-7 * rng.uniform(-1.0, 1.0, fb0.shape))
    keys = ("permr", "permb", "sat_b", "mass_r", "mass_b")
    out = {"workload": "lbm-d3q19-twophase-mrt-256^3", "steps": steps, "sample_every": every,
           "window": [window_from, steps]}
    series = {}
    legs = ("f32", "f32x", "f64_perturbed_start")
    for name, prec, st in (("f64", "f64", (fr0, fb0)), ("f64_perturbed_start", "f64", (fr1, fb1)),
                           ("f32", "f32", (fr0, fb0)), ("f32x", "f32x", (fr0, fb0))):
        rows = []
        with LBM(inp, ind, collision=coll, precision=prec) as lb:
            lb.set_state(*st)
            done = 0
            while done < steps:
                lb.step(every)
                done += every
                d = lb.check_state(1, every)
                rows.append((done, d.permr, d.permb, d.outm[11], d.tmr, d.tmb))
                if done % 5000 == 0:
                    log(f"precision run {name}: {done}/{steps} steps")
        series[name] = np.array(rows)
        last = series[name][-1]
        out[name] = dict(zip(keys, (float(x) for x in last[1:])))
    w = series["f64"][:, 0] >= window_from
    for name in ("f64",) + legs:
        out[name]["window_mean"] = dict(zip(keys, (float(x) for x in series[name][w, 1:].mean(0))))
    for i, k in enumerate(keys):
        ref, ref_w = series["f64"][-1, i + 1], series["f64"][w, i + 1].mean()
        out[f"rel_{k}"] = {n: abs(series[n][-1, i + 1] - ref) / abs(ref) for n in legs}
        out[f"rel_{k}_window_mean"] = {n: abs(series[n][w, i + 1].mean() - ref_w) / abs(ref_w) for n in legs}
        out[f"max_rel_{k}_over_samples"] = {
            n: float(np.max(np.abs(series[n][:, i + 1] - series["f64"][:, i + 1]) / np.abs(series["f64"][:, i + 1])))
            for n in legs}
    return out


def cpu_baseline(ind, inp, blue, budget_s=12.0):
    """Oracle on the same 256^3 workload, its node and link loops on this
    process's cores (at most 16, the host share of one GPU): steps until the
    budget."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_bind as ob
    from porescale_amd.lbm import mrt_rates
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    ob.lib().or_set_threads(threads)
    threads = ob.lib().or_num_threads()
    S = (mrt_rates(inp.vir), mrt_rates(inp.vib))
    o = ob.OracleLBM(inp, ind, S=S)
    o.set_state(*initial_state(inp, ind, blue))
    n, t = 0, 0.0
    while t < budget_s and n < 50:
        t0 = time.perf_counter()
        o.step(1)
        t += time.perf_counter() - t0
        n += 1
    return {"value": o.nl * n / t / 1e6, "unit": "MLUPS", "cores": threads, "kind": "port",
            "sample": f"{n} steps of the same 256^3 two-phase MRT workload, oracle in C, {threads} OpenMP threads"}


def _time_ref(exe, d, timeout=600.0):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_bind as ob
    t0 = time.perf_counter()
    ob._run(exe, d, timeout)
    return time.perf_counter() - t0


def cpu_reference(lsm_steps_to_converge=None):
    """The reference programs themselves (oracle/_ref: L1 and XX_P.f
    compiled from the reference's sources with flang, serial) and the LSM
    oracle, each timed on this host over a bounded window.  Per-step times
    come from two runs that differ only in their step count, so set-up
    (tables, file I/O) cancels.
      L1: the 64^3 TRT pack (tau 1), icon 0 start at S_b = 0.5; one L1
          main-loop step is the LBM step plus lsm_step of the solid lattice
          (L1 always runs both).
      XX_P.f: the 64^3 solid pack, alpha 1, beta 0.1 (E, nu from the inverse
          of XX:411-414): damped-Verlet node-steps per second."""
    import dataclasses
    import shutil
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_bind as ob
    from porescale_amd import cases
    from porescale_amd import geometry as geo
    from porescale_amd import io as pio
    out = {}
    if ob.ref_available("l1_ref"):
        ind, inp, coll, S, blue = cases.twophase_trt_pack(64)
        times = {}
        for itw in (5, 25):
            d = ob._prep_dir(ind)
            try:
                pio.write_lbm_input(os.path.join(d, "INPUT_PARAMETERS"),
                                    dataclasses.replace(inp, itm=1, itw=itw, its=1000, sat=0.5, icon=0))
                times[itw] = _time_ref("l1_ref", d)
            finally:
                shutil.rmtree(d, ignore_errors=True)
        per = (times[25] - times[5]) / 20
        nl = int((ind == 1).sum())
        out["l1_ref"] = {"value": nl / per / 1e6, "unit": "MLUPS (fluid nodes, LBM + lsm_step per step)",
                         "s_per_step": per, "cores": 1, "kind": "reference",
                         "sample": "L1 compiled from the reference's sources (flang -O2, serial) on the 64^3 TRT "
                                   "pack: 25 minus 5 main-loop steps, each the LBM step and lsm_step"}
    ind = geo.sphere_pack((64, 64, 64), 0.35, 6.0, seed=0)
    nz, ny, nx = ind.shape
    r = 0.1  # beta / alpha; alpha = 1 (inverse of XX:411-414)
    nu = (1 - 4 * r) / (4 * (1 + r))
    E = 1.0 * (5 - 10 * nu)
    if ob.ref_available("xx_ref"):
        times = {}
        for its in (10, 60):
            d = ob._prep_dir(ind)
            try:
                pio.write_lsm_input(os.path.join(d, "INPUT_PARAMETERS"),
                                    pio.LSMInput(ncx=nx, ncy=ny, ncz=nz, Em1=E, nu1=nu, Em2=E, nu2=nu, rs=1.0,
                                                 dt=0.01, itm=1, its=its, gstrain=0.01, nthread=1))
                times[its] = _time_ref("xx_ref", d)
            finally:
                shutil.rmtree(d, ignore_errors=True)
        per = (times[60] - times[10]) / 50
        o = ob.OracleLSM(ind, "xx", alf=(1.0, 0.1))
        ns = o.ns
        ent = {"value": ns / per, "unit": "node-steps/s", "s_per_step": per, "nodes": ns, "cores": 1,
               "kind": "reference",
               "sample": "XX_P.f compiled from the reference's sources (flang -O2, serial) on the 64^3 solid "
                         "pack, alpha 1, beta 0.1: 60 minus 10 damped-Verlet steps"}
        if lsm_steps_to_converge:
            ent["verlet_steps_to_moduli_1e-5"] = lsm_steps_to_converge
            ent["solves_per_s_estimate"] = 1.0 / (lsm_steps_to_converge * per)
        out["xx_ref"] = ent
        t0 = time.perf_counter()
        o.verlet(20)
        per_o = (time.perf_counter() - t0) / 20
        out["lsm_oracle"] = {"value": ns / per_o, "unit": "node-steps/s", "s_per_step": per_o, "cores": 1,
                             "kind": "port", "sample": "the LSM oracle (C restatement of XX_P.f), same case, "
                                                       "20 damped-Verlet steps"}
    return out


def _plain_rate(ind, cases, rtol, pars=(1.0, 0.1), maxit=200000, stiff=None, end_to_end=False):
    """Plain FP64 GPU path: block-Jacobi PCG, one load case per solve.
    Returns the seconds of the solves alone (after a warm-up solve that
    allocates and builds the graph) and their iteration counts; with
    end_to_end also the seconds from handle creation to the last solve with
    no warm-up."""
    from porescale_amd.lsm import LSM
    if end_to_end:
        t0 = time.perf_counter()
        with LSM(ind, *pars, precision="f64", stiff=stiff) as s:
            for c in cases:
                s.solve(c, rtol, maxit)
        t_e2e = time.perf_counter() - t0
    with LSM(ind, *pars, precision="f64", stiff=stiff) as s:
        s.solve(cases[0], rtol, maxit)  # warm-up (allocation, graph build)
        t0 = time.perf_counter()
        its = [s.solve(c, rtol, maxit)[0] for c in cases]
        t = time.perf_counter() - t0
    return (t, its, t_e2e) if end_to_end else (t, its)


def lsm_roofline():
    """The LSM operator (k_apply, one right-hand side) against the FP32 and
    FP64 vector peaks.  Algorithmic flops per application are the
    reference's force loops over the same volume (XX:1406-1500): 26 per
    bond and 57 per angle, over its bond and angle tables (nbo, nna of load
    case xx); the kernel time is the mean of 20 launches with HIP events on
    the handle's stream."""
    from porescale_amd import geometry as geo
    from porescale_amd.lsm import LSM, table_counts
    out = {}
    for n, ind in ((64, geo.sphere_pack((64, 64, 64), 0.35, 6.0, seed=0)),
                   (128, geo.two_component(geo.sphere_pack((128, 128, 128), 0.35, 6.0, seed=5), 0.3, seed=6,
                                           radius=4.0))):
        tc = table_counts(ind, "xx")
        flops = 26.0 * tc["nbo"] + 57.0 * tc["nna"]
        with LSM(ind, 1.0, 0.1, 0.3, 0.02, precision="f32") as s:
            s.profile(5)
            t64, t32 = s.profile(20)
            nodes = s.nsolid_nodes
        out[f"{n}^3"] = {
            "bound": "valu", "kernel": "k_apply<float> (one right-hand side)", "unit": "TFLOP/s",
            "algorithmic_flops_per_launch": flops, "ref_nodes": tc["ns"], "bonds": tc["nbo"], "angles": tc["nna"],
            "flops_per_node": flops / tc["ns"], "solid_nodes": nodes,
            "fp32": {"mean_launch_ms": t32, "achieved": flops / (t32 * 1e-3) / 1e12, "peak": 157.3,
                     "frac": flops / (t32 * 1e-3) / 1e12 / 157.3},
            "fp64": {"mean_launch_ms": t64, "achieved": flops / (t64 * 1e-3) / 1e12, "peak": 78.6,
                     "frac": flops / (t64 * 1e-3) / 1e12 / 78.6}}
    return out


def verlet_convergence(ind, chunk=10000, max_steps=300000, tol=1e-5):
    """Damped-Verlet steps (the reference's solver, XX:1377-1545, on the
    GPU in FP64) until C11 of load case xx is within ``tol`` of its
    equilibrium value (PCG to 1e-12): the step count one reference solve
    needs."""
    from porescale_amd.lsm import LSM
    with LSM(ind, 1.0, 0.1, precision="f64") as s:
        s.solve("xx", 1e-12, 200000)
        c_eq = s.calc_stress("xx")[1][0, 0]
        done, err = 0, None
        t0 = time.perf_counter()
        while done < max_steps:
            s.verlet("xx", chunk, restart=(done == 0))
            done += chunk
            c = s.calc_stress("xx")[1][0, 0]
            err = abs(c - c_eq) / abs(c_eq)
            if err <= tol:
                break
        t = time.perf_counter() - t0
    return {"steps": done, "converged": bool(err is not None and err <= tol), "C11_rel_err": err,
            "C11_equilibrium": c_eq, "gpu_seconds": t, "tol": tol}


def run_lsm(calib_steps=50):
    """LSM rows: six-case moduli at 64^3 (plain FP64 vs batched FP32+MG, at
    rtol 1e-10 and 1e-5), uneven sub-volumes in one batch, adjoint gradients
    at 128^3 with a lognormal stiffness field and Gauss-Newton calibration."""
    import numpy as np
    from porescale_amd import geometry as geo
    from porescale_amd.autograd import calibrate
    from porescale_amd.lsm import LSM, batch_moduli
    res = {}
    ind = geo.sphere_pack((64, 64, 64), 0.35, 6.0, seed=0)
    cfg2 = {"cfg": "64^3 sphere pack (r=6, porosity 0.35, seed 0), alpha=1, beta=0.1, 6 load cases"}
    for rtol in (1e-10, 1e-5):
        t_plain, its = _plain_rate(ind, range(6), rtol)
        with LSM(ind, 1.0, 0.1, precision="f32") as s:
            s.moduli(rtol, 200000)
            t0 = time.perf_counter()
            C6, K, G = s.moduli(rtol, 200000)
            t_b = time.perf_counter() - t0
        with LSM(ind, 1.0, 0.1, precision="f64") as s:
            _, K64, G64 = s.moduli(1e-12, 200000)
        cfg2[f"rtol_{rtol:g}"] = {
            "plain_fp64_solves_per_s": 6 / t_plain, "plain_fp64_iters": its,
            "batched_mg_solves_per_s": 6 / t_b, "speedup": t_plain / t_b,
            "K_rel_err_vs_fp64": abs(K - K64) / K64, "G_rel_err_vs_fp64": abs(G - G64) / G64,
            "C11": C6[0, 0], "C44": C6[3, 3], "K": K, "G": G}
    with LSM(ind, 1.0, 0.1, precision="f32") as s:
        t64, t32 = s.profile(20)
        cfg2["matvec_ms"] = {"fp64": t64, "fp32": t32, "solid_nodes": s.nsolid_nodes}
    res["moduli_64"] = cfg2
    res["roofline"] = lsm_roofline()
    res["verlet_64"] = verlet_convergence(ind)
    # uneven sub-volumes, two components, one batch
    rng = np.random.default_rng(0)
    sizes = rng.integers(32, 97, size=64)
    vols = [geo.two_component(geo.sphere_pack((int(L), int(L), int(L)), 0.35, 5.0, seed=100 + i), 0.7,
                              seed=200 + i, radius=4.0) for i, L in enumerate(sizes)]
    pars = (1.0, 0.1, 0.3, 0.02)
    nv = len(vols)
    res["batch_64"] = {"volumes": nv, "sizes": [int(x) for x in sizes], "systems": 6 * nv,
                       "nodes": int(sum(int(x) ** 3 for x in sizes))}
    for rtol in (1e-10, 1e-5):
        t0 = time.perf_counter()
        Cb, KG, ms, it = batch_moduli(vols, *pars, precision="f32", rtol=rtol)
        t_b = time.perf_counter() - t0
        # second call of the same batch: the first pays one-time set-up
        # (module load, tile autotuning) that the plain leg below does not
        t0 = time.perf_counter()
        _, KGw, ms_w, it = batch_moduli(vols, *pars, precision="f32", rtol=rtol)
        t_bw = time.perf_counter() - t0
        t_ps = t_pe = 0.0
        for v in vols:
            t, _, te = _plain_rate(v, range(6), rtol, pars, end_to_end=True)
            t_ps += t
            t_pe += te
        # end to end: set-up (tables, allocation, multigrid hierarchy or
        # graph capture) and solves on both legs; solve only: the plain leg's
        # timed solves and the batch's own event time around right-hand
        # sides, solve and stresses.  The plain leg solves every volume of
        # the batch, one load case at a time
        res["batch_64"][f"rtol_{rtol:g}"] = {
            "seconds_first_call": t_b, "seconds": t_bw, "solve_seconds": ms_w / 1e3,
            "solves_per_s": 6 * nv / t_bw, "iterations": it,
            "end_to_end": {"plain_fp64_seconds": t_pe, "plain_fp64_solves_per_s": 6 * nv / t_pe,
                           "batched_mg_solves_per_s": 6 * nv / t_bw, "speedup": t_pe / t_bw},
            "solve_only": {"plain_fp64_seconds": t_ps, "plain_fp64_solves_per_s": 6 * nv / t_ps,
                           "batched_mg_solves_per_s": 6 * nv / (ms_w / 1e3), "speedup": t_ps / (ms_w / 1e3)}}
        if rtol == 1e-10:
            KG10 = KGw
        else:
            res["batch_64"]["rtol_1e-05"]["KG_max_rel_err_vs_rtol_1e-10"] = float(np.max(np.abs(KGw - KG10) / KG10))
    # adjoint at 128^3: two components, lognormal stiffness (sigma 0.3, seed 1)
    n5 = 128
    ind5 = geo.two_component(geo.sphere_pack((n5, n5, n5), 0.35, 6.0, seed=5), 0.3, seed=6, radius=4.0)
    stiff = np.exp(0.3 * np.random.default_rng(1).standard_normal(ind5.shape))
    adj = {"cfg": "128^3 two-component sphere pack, lognormal stiffness sigma 0.3 seed 1, rtol 1e-10"}
    for prec in ("f64", "f32"):
        with LSM(ind5, *pars, precision=prec, stiff=stiff) as s:
            t0 = time.perf_counter()
            _, K5, G5 = s.moduli(1e-10, 200000)
            t_f = time.perf_counter() - t0
            t0 = time.perf_counter()
            s.adjoint(1.0, 0.0, 1e-10, 200000)
            t_a = time.perf_counter() - t0
        adj[prec] = {"forward_s": t_f, "adjoint_s": t_a, "grads_per_s": 1.0 / (t_f + t_a), "K": K5, "G": G5}
    adj["speedup"] = adj["f32"]["grads_per_s"] / adj["f64"]["grads_per_s"]
    # the same gradient at the metric's tolerance 1e-5 (forward and adjoint)
    a5 = {}
    for prec in ("f64", "f32"):
        with LSM(ind5, *pars, precision=prec, stiff=stiff) as s:
            s.moduli(1e-5, 200000)  # warm-up: hierarchy and graphs
            t0 = time.perf_counter()
            _, K5, G5 = s.moduli(1e-5, 200000)
            dab, _ = s.adjoint(1.0, 0.0, 1e-5, 200000, want_stiff=False)
            t_g = time.perf_counter() - t0
        a5[prec] = {"seconds": t_g, "grads_per_s": 1.0 / t_g, "K": K5, "G": G5, "dK_dab": dab.tolist()}
    a5["speedup"] = a5["f32"]["grads_per_s"] / a5["f64"]["grads_per_s"]
    a5["K_rel_err_vs_fp64"] = abs(a5["f32"]["K"] - a5["f64"]["K"]) / a5["f64"]["K"]
    adj["rtol_1e-05"] = a5
    # plain FP64 (one load case per block-Jacobi PCG solve) on the forward
    # moduli.  The gradient of K weights only the three normal load cases, so
    # its adjoint has three nonzero systems of the same matrix; the batched
    # FP64 adjoint above solves them together and is no slower than solving
    # them one at a time, so plain forward + FP64 adjoint bounds a plain FP64
    # gradient from below
    t_plain, its5 = _plain_rate(ind5, range(6), 1e-10, pars, stiff=stiff)
    t_grad = t_plain + adj["f64"]["adjoint_s"]
    f32_grad = adj["f32"]["forward_s"] + adj["f32"]["adjoint_s"]
    adj["plain_fp64_forward"] = {"seconds": t_plain, "solves_per_s": 6 / t_plain, "iters": its5,
                                 "speedup_f32_forward": t_plain / adj["f32"]["forward_s"]}
    adj["plain_fp64_gradient_lower_bound"] = {"seconds": t_grad, "speedup_f32": t_grad / f32_grad}
    steps = calib_steps
    tgt = np.array([adj["f32"]["K"], adj["f32"]["G"]]) * np.array([1.05, 0.97])
    t0 = time.perf_counter()
    hist, losses = calibrate(ind5, np.array(pars), tgt, steps=steps, rtol=1e-10, precision="f32", stiff=stiff)
    t_c = time.perf_counter() - t0
    adj["calibration"] = {"steps": steps, "handle": "one LSM handle, springs replaced in place each step",
                          "seconds": t_c, "s_per_step": t_c / steps,
                          "loss_first": float(losses[0]), "loss_last": float(losses[-1]),
                          "target_KG": tgt.tolist(), "fitted_ab": hist[-1].tolist()}
    res["adjoint_128"] = adj
    return res


def pmc_traffic(kernel, profile=None):
    """HBM bytes per launch of ``kernel`` from the committed rocprofv3 --pmc
    summary (FETCH_SIZE and WRITE_SIZE, each scaled by a copy of known size)."""
    profile = profile or f"profiles/{kernel}_lbm256_mrt.md"
    try:
        with open(os.path.join(ROOT, profile)) as f:
            for line in f:
                t = line.split()
                if len(t) == 3 and t[0] == "TRAFFIC_BYTES_PER_LAUNCH" and t[1] == kernel:
                    return float(t[2]), profile
    except OSError:
        pass
    return None, None


def calib_copy(mib):
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(10):
        torch.neg(a, out=b)
    torch.cuda.synchronize()
    log(f"calib-copy: 10 launches of torch.neg over {mib} MiB float32 in, {mib} MiB out")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--lsm", dest="lsm", action="store_true", default=True)
    ap.add_argument("--no-lsm", dest="lsm", action="store_false")
    ap.add_argument("--no-cpu", dest="cpu", action="store_false", default=True)
    ap.add_argument("--no-coupled", dest="coupled", action="store_false", default=True)
    ap.add_argument("--no-precision", dest="precision", action="store_false", default=True)
    ap.add_argument("--calib-copy", type=int, default=0, metavar="MIB")
    ap.add_argument("--calib-steps", type=int, default=50, help="Gauss-Newton calibration steps in the LSM rows")
    ap.add_argument("--lbm-storage", choices=("f32x", "f32"), default="f32x",
                    help="storage of the headline LBM workload (profiling the other one)")
    ap.add_argument("--lsm-only", action="store_true", help="only the LSM rows (for profiling them)")
    ap.add_argument("--lsm-matvec", action="store_true", help="only the LSM operator roofline (profiling)")
    ap.add_argument("--precision-steps", type=int, default=0, metavar="N",
                    help="FP32 storage against FP64 on the headline workload over N steps (own JSON line)")
    args = ap.parse_args()
    ws = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or ws != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}, WORLD_SIZE {ws}): the LBM and LSM "
              "paths here are single-GPU builds; multi-GPU sharding is a design note, not measured.",
              file=sys.stderr)
        sys.exit(2)
    if args.calib_copy:
        calib_copy(args.calib_copy)
        return
    if args.lsm_matvec:
        print(json.dumps({"lsm_roofline": lsm_roofline()}), flush=True)
        return
    if args.lsm_only:
        print(json.dumps({"lsm": run_lsm(args.calib_steps)}), flush=True)
        return
    if args.precision_steps:
        print(json.dumps({"precision": run_precision(args.precision_steps)}), flush=True)
        return
    from porescale_amd import _lib
    dev, ncu, gb = _lib.device_info()
    log(f"device {dev}, {ncu} CUs, {gb:.0f} GB")
    prec = args.lbm_storage
    lbm, ind, inp, blue = run_lbm(args.steps, args.warmup, before=run_other_storage(prec), prec=prec)
    ms_step = lbm["tot_ms"] / args.steps
    mlups = lbm["nl"] * args.steps / (lbm["tot_ms"] * 1e-3) / 1e6
    kname = "k_collide_f32x" if prec == "f32x" else "k_collide_f32"
    traffic, traffic_src = pmc_traffic(kname)
    alg_b = ALG_BYTES[prec]
    coll_gbs = alg_b * lbm["nl"] / (lbm["coll_ms"] * 1e-3) / 1e9
    step_gbs = alg_b * lbm["nl"] / (ms_step * 1e-3) / 1e9
    log(f"LBM nl={lbm['nl']} {ms_step:.3f} ms/step ({lbm['coll_ms']:.3f} collide + {lbm['dens_ms']:.3f} density)"
        f" {mlups:.0f} MLUPS, step {step_gbs:.0f} GB/s algorithmic")
    out = {
        "metric": "LBM MLUPS (D3Q19 two-phase colour-gradient MRT, fluid nodes)",
        "value": mlups,
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "storage": STORAGE[prec],
        "data": "synthetic (seeded sphere pack and blob phase distribution)",
        "config": {"workload": "lbm-d3q19-twophase-mrt-256^3", "grid": [256, 256, 256],
                   "fluid_nodes": lbm["nl"], "porosity": 0.30, "sphere_radii": [4, 10],
                   "viscosity_ratio": "1:10", "wet": 0.7071, "saturation_blue": 0.5},
        "hbm_frac_step": step_gbs / HBM_PEAK_GBS,
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": coll_gbs, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": coll_gbs / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": traffic_src,
                     "algorithmic_bytes_per_node": alg_b, "nodes_per_launch": lbm["nl"],
                     "mean_launch_ms": lbm["coll_ms"]},
        "lbm_diag": {"permr": lbm["permr"], "permb": lbm["permb"], "sat_b": lbm["sat_b"]},
    }
    for p, leg in lbm["before"].items():
        out[f"lbm_{p}"] = leg
        log(f"LBM {p} storage {json.dumps(leg)}")
    out["trt_pack_64"] = run_trt_pack()
    log(f"TRT pack 64^3 {json.dumps(out['trt_pack_64'])}")
    if args.precision:
        out["precision"] = run_precision()
        log(f"precision {json.dumps(out['precision'])[:600]}")
    if args.coupled:
        out["coupled"] = run_coupled(ind, inp, blue)
        log(f"coupled {json.dumps(out['coupled'])}")
    if args.lsm:
        out["lsm"] = run_lsm(args.calib_steps)
        log(f"LSM {json.dumps(out['lsm'])[:400]}")
    if args.cpu:
        out["cpu_baseline"] = cpu_baseline(ind, inp, blue)
        conv = out.get("lsm", {}).get("verlet_64", {})
        out["cpu_baseline"]["reference"] = cpu_reference(conv.get("steps") if conv.get("converged") else None)
        log(f"CPU baseline {out['cpu_baseline']}")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
