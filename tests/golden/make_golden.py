"""Writes the golden fixtures of tests/golden/ by running the flang-compiled
reference programs (oracle/_ref, built by `make -C oracle ref` from
/root/reference) on small seeded inputs.

    python3 tests/golden/make_golden.py

Fixtures (compressed .npz, inputs and reference outputs only):
  lbm_icon0.npz   L1, random all-red/all-blue start (initdistr icon=0),
                  12^3 sphere pack, 2 pre-steps + 5 steps.
  lbm_blobs.npz   L1, blob start injected through iteration2 (icon=1),
                  18x16x14 pack, wetting 0.7071, anisotropic body forces,
                  unequal viscosities, 2 + 3*20 steps, output.m rows.
  lbm_blobs_fields.npz  the save_results files of the lbm_blobs run's last
                  iteration: velocity_field, color_field, pressure<it1> and
                  the last ind_red_blue block.
  lbm_coupled.npz L1 with its coupled solid lattice (12x10x14, Em 0.2, nu 0.2,
                  rs 2): 2 pre-steps + 2*6 coupled steps from initdistr; the
                  LSM_disf2/LSM_velocity_field/LSM_force_field/LSM_fluid_force
                  files and the last alpha_beta block (averaged tensors as
                  printed: sums / ncub).
  lsm_xx.npz      XX_P.f, two-component 10x12x14 pack, 100 Verlet steps,
                  nthread=3: r, v, a and stress_tensor moduli.
  lsm_yz.npz      YZ_P.f, same geometry, nthread=1.
  lsm_xx_fields.npz, lsm_yz_fields.npz  the field files of the last block
                  (LSM_coord_field, LSM_displ_field, LSM_velocity_field,
                  LSM_force_field, and the sums of 'write').
  lsm_xx_restart.npz, lsm_yz_restart.npz  icon=1 from the state above
                  (iteration2), 50 more steps: r, v, a and the moduli.
  lsm_settings_xx.txt, lsm_settings_yz.txt  the lsm_settings file of the
                  lsm_xx / lsm_yz runs as the reference wrote it (table sizes
                  nbo, nna and the z-layer partition included).
                  `make_golden.py settings` rewrites only these.
  ref_tables.npz  the reference's velocities19/velocities9 and cube_* files
                  as read from /root/reference (data tables of the API).
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle_bind as ob  # noqa: E402
from porescale_amd import geometry as geo  # noqa: E402
from porescale_amd import io as pio  # noqa: E402
from porescale_amd.lsm import spring_constants  # noqa: E402


def lbm_cases():
    n = 12
    ind0 = geo.sphere_pack((n, n, n), 0.5, 3.0, seed=1)
    inp0 = pio.LBMInput(B=0.7, sig=0.01, vir=1 / 6, vib=1 / 6, wet=0.0, idir=3, itm=1, itw=5, sat=0.5,
                        ncx=n, ncy=n, ncz=n, dt=1.0, bfr=[0, 0, 1e-5], bfb=[0, 0, 1e-5])
    ind1 = geo.sphere_pack((14, 16, 18), 0.45, 3.0, seed=2)
    nz, ny, nx = ind1.shape
    inp1 = pio.LBMInput(B=0.7, sig=0.02, vir=0.1, vib=0.3, wet=0.7071, idir=3, itm=3, itw=20, sat=0.4,
                        ncx=nx, ncy=ny, ncz=nz, dt=1.0, bfr=[0, 1e-5, 2e-5], bfb=[1e-6, 0, 2e-5])
    blue = geo.phase_blobs(ind1, 0.4, seed=3, radius=2.5)
    return (ind0, inp0), (ind1, inp1, blue)


def coupled_case():
    """L1 with its coupled solid: stable springs (Em 0.2, nu 0.2, rs 2)."""
    ind = geo.sphere_pack((12, 10, 14), 0.45, 2.5, seed=7)
    nz, ny, nx = ind.shape
    inp = pio.LBMInput(B=0.7, sig=0.02, vir=0.1, vib=0.2, wet=0.3, idir=3, itm=2, itw=6, sat=0.5,
                       ncx=nx, ncy=ny, ncz=nz, dt=1.0, bfr=[0, 1e-4, 2e-4], bfb=[1e-5, 0, 2e-4],
                       Em=0.2, nu=0.2, rs=2.0, dl=1.0)
    return ind, inp


def lsm_geometry():
    ind = geo.sphere_pack((10, 12, 14), 0.4, 3.0, seed=5)
    return geo.two_component(ind, 0.3, seed=6, radius=2.0)


LSM_INP = dict(Em1=18.639, nu1=0.35065, Em2=94.529, nu2=0.0742, rs=1.0, dt=0.01, itm=2, its=50,
               gstrain=0.01)


def settings19_golden():
    """settings_19 as the reference L1 writes it (L1:1442-1514) for the blob
    case (icon = 1) and the coupled case (icon = 0)."""
    import dataclasses
    import shutil
    import oracle_bind as ob
    (_, _), (ind1, inp1, blue) = lbm_cases()
    ind2, inp2 = coupled_case()
    for name, ind, inp, restart in (("blobs", ind1, inp1, blue), ("coupled", ind2, inp2, None)):
        d = ob._prep_dir(ind)
        try:
            if restart is not None:
                inp = dataclasses.replace(inp, icon=1)
                pio.write_lbm_restart(os.path.join(d, "iteration2"), 0,
                                      *ob.equilibrium_state(inp, ob.fluid_order_mask(ind, restart)))
            pio.write_lbm_input(os.path.join(d, "INPUT_PARAMETERS"), dataclasses.replace(inp, itm=1, itw=1))
            ob._run("l1_ref", d, 600.0)
            shutil.copy(os.path.join(d, "settings_19"), os.path.join(HERE, f"settings_19_{name}.txt"))
        finally:
            shutil.rmtree(d, ignore_errors=True)


def lsm_converged_geometry():
    """10^3 two-component pack (solid 0.66, 21 % code 3) for the converged
    reference moduli."""
    return geo.two_component(geo.sphere_pack((10, 10, 10), 0.35, 2.5, seed=11), 0.3, seed=12, radius=2.0)


def lsm_converged_golden(its=20000, itm=2):
    """The reference's own solver run to equilibrium: XX_P.f and YZ_P.f for
    itm * its damped-Verlet steps on lsm_converged_geometry() with the
    LSM_INP springs (the moduli of the last stress_tensor block agree with
    the block before to ~1e-14 at 20000 steps a block)."""
    import oracle_bind as ob
    ind = lsm_converged_geometry()
    nz, ny, nx = ind.shape
    inp = pio.LSMInput(ncx=nx, ncy=ny, ncz=nz, nthread=1, **dict(LSM_INP, itm=itm, its=its))
    out = {"steps": itm * its}
    for case in ("xx", "yz"):
        st = ob.run_ref_lsm(case, ind, inp, timeout=3000)[3]
        vals = {}
        for line in st.splitlines():
            for key in ("Exxxx", "Eyyxx", "Eyyzz", "Eyzyz"):
                if line.strip().startswith(key):
                    vals.setdefault(key, []).append(pio.parse_fortran_numbers(line.split("=", 1)[1])[0])
        for key, v in vals.items():
            out[f"{case}_{key}"] = np.array(v)
    np.savez_compressed(os.path.join(HERE, "lsm_converged.npz"), **out)


def lsm_settings_golden():
    import oracle_bind as ob
    from porescale_amd import io as pio
    ind = lsm_geometry()
    nz, ny, nx = ind.shape
    for case, nth in (("xx", 3), ("yz", 1)):
        inp = pio.LSMInput(ncx=nx, ncy=ny, ncz=nz, nthread=nth, **LSM_INP)
        text = ob.run_ref_lsm(case, ind, inp, settings=True)[-1]
        with open(os.path.join(HERE, f"lsm_settings_{case}.txt"), "w") as fh:
            fh.write(text)


def main():
    (ind0, inp0), (ind1, inp1, blue) = lbm_cases()
    fr, fb, outm, it1 = ob.run_ref_lbm(inp0, ind0)
    np.savez_compressed(os.path.join(HERE, "lbm_icon0.npz"), fr=fr, fb=fb, outm=outm, steps=7)
    fr0, fb0 = ob.equilibrium_state(inp1, ob.fluid_order_mask(ind1, blue))
    fr, fb, outm, it1, fl = ob.run_ref_lbm(inp1, ind1, restart=(fr0, fb0), fields=True)
    np.savez_compressed(os.path.join(HERE, "lbm_blobs.npz"), fr=fr, fb=fb, outm=outm, steps=62)
    np.savez_compressed(os.path.join(HERE, "lbm_blobs_fields.npz"), it1=it1, iter=fl["iter"],
                        velocity=fl["velocity"], color=fl["color"], pressure=fl["pressure"], phase=fl["phase"])
    ind, inp = coupled_case()
    fr, fb, outm, it1, lsmf, ab = ob.run_ref_coupled(inp, ind)
    np.savez_compressed(os.path.join(HERE, "lbm_coupled.npz"), fr=fr, fb=fb, it1=it1,
                        **{f"lsm_{k}": v for k, v in lsmf.items()},
                        **{f"ab_{k}": np.asarray(v) for k, v in ab.items()})
    ind = lsm_geometry()
    nz, ny, nx = ind.shape
    for case, nth in (("xx", 3), ("yz", 1)):
        inp = pio.LSMInput(ncx=nx, ncy=ny, ncz=nz, nthread=nth, **LSM_INP)
        r, v, a, st, fl = ob.run_ref_lsm(case, ind, inp, files=True)
        np.savez_compressed(os.path.join(HERE, f"lsm_{case}_fields.npz"), **fl)
        mods = ob.parse_stress_tensor(st)
        np.savez_compressed(os.path.join(HERE, f"lsm_{case}.npz"), r=r, v=v, a=a, steps=100, nthread=nth,
                            **{k: np.float64(v_) for k, v_ in mods.items()})
        # icon = 1 from that state: 1 block of 50 more steps
        inp1 = pio.LSMInput(ncx=nx, ncy=ny, ncz=nz, nthread=nth, **dict(LSM_INP, itm=1))
        r1, v1, a1, st1 = ob.run_ref_lsm(case, ind, inp1, restart=(r, v, a))
        np.savez_compressed(os.path.join(HERE, f"lsm_{case}_restart.npz"), r=r1, v=v1, a=a1, steps=50,
                            **{k: np.float64(v_) for k, v_ in ob.parse_stress_tensor(st1).items()})
    ref = "/root/reference/LSM2S/Cxxxx"
    b, an, sb, ab = pio.read_cube_tables(ref)
    np.savez_compressed(os.path.join(HERE, "ref_tables.npz"), velocities19=pio.read_velocities(
        os.path.join(ref, "velocities19"), 19), velocities9=pio.read_velocities(os.path.join(ref, "velocities9"), 9),
        cube_bon=b, cube_ang=an, cube_sbond=sb, cube_abond=ab)
    lsm_settings_golden()
    settings19_golden()
    lsm_converged_golden()
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    if sys.argv[1:] == ["settings"]:
        lsm_settings_golden()
    elif sys.argv[1:] == ["settings19"]:
        settings19_golden()
    elif sys.argv[1:] == ["converged"]:
        lsm_converged_golden()
    else:
        main()
