"""File formats of the reference (the API surface) and the seeded generators."""
import os

import numpy as np
import pytest

from porescale_amd import geometry as geo
from porescale_amd import io as pio

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
REF = "/root/reference"


def test_generated_tables_equal_reference_files():
    g = np.load(os.path.join(GOLD, "ref_tables.npz"))
    b, a, sb, ab = geo.cube_tables()
    assert np.array_equal(b, g["cube_bon"]) and np.array_equal(a, g["cube_ang"])
    assert np.array_equal(sb, g["cube_sbond"]) and np.array_equal(ab, g["cube_abond"])
    assert np.array_equal(pio.D3Q19, g["velocities19"]) and np.array_equal(pio.BOND9, g["velocities9"])


def test_cube_and_velocity_files_round_trip(tmp_path):
    pio.write_cube_tables(str(tmp_path), *geo.cube_tables())
    back = pio.read_cube_tables(str(tmp_path))
    for x, y in zip(back, geo.cube_tables()):
        assert np.array_equal(x, y)
    pio.write_velocities(str(tmp_path / "v19"), pio.D3Q19)
    assert np.array_equal(pio.read_velocities(str(tmp_path / "v19"), 19), pio.D3Q19)


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference inputs not mounted")
def test_reads_reference_input_files():
    p = pio.read_lbm_input(os.path.join(REF, "LBM-LSM", "INPUT_PARAMETERS"))
    assert (p.ncx, p.ncy, p.ncz, p.idir, p.itw) == (120, 120, 120, 1, 10000)
    assert p.cr2 == pytest.approx(1 / 3) and p.vir == 0.00625 and p.dt == 0.1
    q = pio.read_lsm_input(os.path.join(REF, "LSM2S", "Cxxxx", "INPUT_PARAMETERS"))
    assert (q.ncx, q.ncy, q.ncz, q.nthread, q.icon) == (150, 150, 54, 8, 0)
    assert q.Em2 == 94.529 and q.gstrain == 0.01
    ind = pio.read_ind(os.path.join(REF, "LSM2S", "Cxxxx", "ind"), 150, 150, 54)
    vals, counts = np.unique(ind, return_counts=True)
    assert dict(zip(vals.tolist(), counts.tolist())) == {0: 62957, 1: 280228, 3: 871815}


def test_input_parameters_round_trip(tmp_path):
    p = pio.LBMInput(sig=0.0123, bfr=[1e-5, 0, -2e-6], ncx=7, itw=33, wet=-0.25)
    pio.write_lbm_input(str(tmp_path / "IP"), p)
    assert pio.read_lbm_input(str(tmp_path / "IP")) == p
    q = pio.LSMInput(ncx=5, Em1=18.639, nu1=0.35065, gstrain=0.005, nthread=3)
    pio.write_lsm_input(str(tmp_path / "IPL"), q)
    assert pio.read_lsm_input(str(tmp_path / "IPL")) == q


def test_ind_and_restart_round_trip(tmp_path):
    ind = geo.two_component(geo.sphere_pack((5, 6, 7), 0.4, 2.0, seed=1), 0.3, seed=2, radius=1.5)
    pio.write_ind(str(tmp_path / "ind"), ind)
    assert np.array_equal(pio.read_ind(str(tmp_path / "ind"), 7, 6, 5), ind)
    rng = np.random.default_rng(0)
    fr, fb = rng.random((11, 19)), rng.random((11, 19))
    pio.write_lbm_restart(str(tmp_path / "it2"), 4, fr, fb)
    it1, a, b = pio.read_lbm_restart(str(tmp_path / "it2"), 11)
    assert it1 == 4 and np.array_equal(a, fr) and np.array_equal(b, fb)
    r, v, acc = rng.random((9, 3)), rng.random((9, 3)), rng.random((9, 3))
    pio.write_lsm_restart(str(tmp_path / "it3"), r, v, acc)
    for x, y in zip(pio.read_lsm_restart(str(tmp_path / "it3"), 9), (r, v, acc)):
        assert np.array_equal(x, y)


def test_fortran_number_parsing():
    assert list(pio.parse_fortran_numbers("1.5D-3 2*0.25 -3E2")) == [1.5e-3, 0.25, 0.25, -300.0]


def test_sphere_pack_deterministic_and_on_target():
    a = geo.sphere_pack((32, 32, 32), 0.35, 6.0, seed=0)
    b = geo.sphere_pack((32, 32, 32), 0.35, 6.0, seed=0)
    assert np.array_equal(a, b)
    assert 0.35 - 0.03 < (a == 1).mean() <= 0.35  # one radius-6 sphere is 2.8% of 32^3
    c = geo.sphere_pack((32, 32, 32), 0.35, 6.0, seed=1)
    assert not np.array_equal(a, c)
    blue = geo.phase_blobs(a, 0.5, seed=0, radius=3.0)
    assert abs(blue.sum() / (a == 1).sum() - 0.5) < 0.1 and not np.any(blue & (a == 0))
