"""psx_lsm_stress_noalloc, the allocation-free calc_stress, against
psx_lsm_stress on the same displacement (GPU): two calls in a row, the
second on a smaller volume, both through the C-ABI with device buffers the
caller owns; and the size limit of its scratch (CPU-side check)."""
import ctypes as C

import numpy as np
import pytest

from porescale_amd import _lib
from porescale_amd import geometry as geo
from porescale_amd.lsm import LSM

pytestmark = pytest.mark.gpu


def _springs(ind, stiff, ab):
    a1, b1, a2, b2 = ab
    ka = np.where(ind == 0, a1, np.where(ind == 3, a2, 0.0)) * stiff
    kb = np.where(ind == 0, b1, np.where(ind == 3, b2, 0.0)) * stiff
    return ka.reshape(-1), kb.reshape(-1)


def _noalloc(ind, ka, kb, u, case, g=0.01):
    import torch
    nz, ny, nx = ind.shape
    dka = torch.tensor(ka, dtype=torch.float64, device="cuda")
    dkb = torch.tensor(kb, dtype=torch.float64, device="cuda")
    du = torch.tensor(np.ascontiguousarray(u).reshape(-1), dtype=torch.float64, device="cuda")
    out = torch.full((9,), float("nan"), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    lc = _lib.LOAD_CASES.index(case)
    _lib.check(_lib.lib().psx_lsm_stress_noalloc(nx, ny, nz, C.c_void_p(dka.data_ptr()), C.c_void_p(dkb.data_ptr()),
                                                 C.c_void_p(du.data_ptr()), lc, g, C.c_void_p(out.data_ptr()),
                                                 None))
    torch.cuda.synchronize()
    return out.cpu().numpy().reshape(3, 3, order="F")


def test_noalloc_stress_equals_handle_stress_twice_smaller_second():
    import torch
    assert torch.cuda.is_available()
    ab = (1.0, 0.1, 0.3, 0.02)
    for n, case, seed in ((24, "xx", 1), (12, "yz", 2)):
        ind = geo.two_component(geo.sphere_pack((n, n, n), 0.35, 3.0, seed=seed), 0.4, seed=seed + 10, radius=2.0)
        stiff = np.exp(0.3 * np.random.default_rng(seed).standard_normal(ind.shape))
        print(f"case noalloc stress {n}^3 {case}", flush=True)
        with LSM(ind, *ab, precision="f64", stiff=stiff) as s:
            s.solve(case, 1e-12, 20000)
            want, _ = s.calc_stress(case)
            u = s.get_u(case)
        ka, kb = _springs(ind, stiff, ab)
        got = _noalloc(ind, ka, kb, u, case)
        np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-15 * np.abs(want).max())


def test_noalloc_rejects_volumes_beyond_its_scratch():
    import torch
    small = torch.zeros(16, dtype=torch.float64, device="cuda")
    p = C.c_void_p(small.data_ptr())
    L = _lib.lib()
    rc = L.psx_lsm_stress_noalloc(300, 300, 300, p, p, p, 0, 0.01, p, None)
    assert rc == _lib.PSX_EINVAL and b"scratch" in L.psx_last_error()
    rc = L.psx_lsm_stress_noalloc(4, 4, 4, p, p, p, 7, 0.01, p, None)
    assert rc == _lib.PSX_EINVAL
