"""torch.autograd harness of the LSM moduli (GPU): the hand-written adjoint
against finite differences of re-solves, and a calibration to target moduli."""
import numpy as np
import pytest
import torch

from porescale_amd import geometry as geo
from porescale_amd.autograd import LSMModuli, calibrate

pytestmark = pytest.mark.gpu


def _ind():
    return geo.two_component(geo.sphere_pack((8, 8, 9), 0.35, 2.5, seed=31), 0.4, seed=32, radius=1.5)


def test_gradcheck_spring_constants():
    ind = _ind()
    ab = torch.tensor([1.0, 0.1, 0.3, 0.02], dtype=torch.float64, requires_grad=True)
    print("case gradcheck ab", flush=True)
    f = lambda x: LSMModuli.apply(x, None, ind, 0.01, 1e-12, "f64")  # noqa: E731
    assert torch.autograd.gradcheck(f, (ab,), eps=1e-6, atol=1e-7, rtol=1e-5)


def test_stiffness_gradient_against_resolve():
    ind = _ind()
    rng = np.random.default_rng(1)
    st = torch.tensor(np.exp(0.3 * rng.standard_normal(ind.shape)), dtype=torch.float64, requires_grad=True)
    ab = torch.tensor([1.0, 0.1, 0.3, 0.02], dtype=torch.float64)
    KG = LSMModuli.apply(ab, st, ind, 0.01, 1e-12, "f64")
    (KG[0] + 2.0 * KG[1]).backward()
    g = st.grad.numpy()
    h = 1e-6
    solid = np.argwhere(ind != 1)
    for idx in solid[[1, len(solid) // 3, -2]]:
        e = np.zeros(ind.shape)
        e[tuple(idx)] = h
        vals = []
        for sgn in (1, -1):
            s2 = torch.tensor(st.detach().numpy() + sgn * e)
            kg = LSMModuli.apply(ab, s2, ind, 0.01, 1e-12, "f64")
            vals.append(float(kg[0] + 2.0 * kg[1]))
        fd = (vals[0] - vals[1]) / (2 * h)
        assert g[tuple(idx)] == pytest.approx(fd, rel=1e-4, abs=1e-9)
    # pore voxels carry no springs
    assert np.all(g[ind == 1] == 0.0)


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_calibration_reaches_target_moduli(precision):
    """Targets generated by known spring constants; Gauss-Newton steps from
    a start 30-50 % off (two adjoint solves per step) fit (K, G) to 1e-8
    relative well inside 50 steps."""
    ind = _ind()
    true = np.array([1.0, 0.1, 0.3, 0.02])
    with torch.no_grad():
        tgt = LSMModuli.apply(torch.tensor(true), None, ind, 0.01, 1e-11, "f64").numpy()
    print(f"case calibration {precision} target {tgt}", flush=True)
    hist, losses = calibrate(ind, true * np.array([1.5, 0.6, 1.3, 1.4]), tgt, steps=12, rtol=1e-11,
                             precision=precision)
    assert losses[-1] < 1e-14 and losses[0] > 1e-3


def test_adam_calibration_decreases_loss():
    ind = _ind()
    true = np.array([1.0, 0.1, 0.3, 0.02])
    with torch.no_grad():
        tgt = LSMModuli.apply(torch.tensor(true), None, ind, 0.01, 1e-10, "f64").numpy()
    hist, losses = calibrate(ind, true * np.array([1.5, 0.6, 1.3, 1.4]), tgt, steps=15, lr=0.05, method="adam")
    assert losses[-1] < 0.25 * losses[0]


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_set_springs_equals_fresh_handle(precision):
    """psx_lsm_set_springs refreshes the per-voxel constants, the
    preconditioners and the multigrid operators in place: moduli and the
    adjoint after a swap equal those of a handle built with the new
    constants (16^3, three multigrid levels in FP32)."""
    from porescale_amd.lsm import LSM
    ind = geo.two_component(geo.sphere_pack((16, 16, 16), 0.35, 3.0, seed=41), 0.4, seed=42, radius=2.0)
    a0, a1 = (1.0, 0.1, 0.3, 0.02), (1.4, 0.05, 0.2, 0.03)
    print(f"case set_springs {precision}", flush=True)
    with LSM(ind, *a1, precision=precision) as s:
        C1, K1, G1 = s.moduli(1e-11, 20000)
        d1, _ = s.adjoint(1.0, 0.5, 1e-11, 20000, want_stiff=False)
    with LSM(ind, *a0, precision=precision) as s:
        s.moduli(1e-11, 20000)
        s.set_springs(*a1)
        C2, K2, G2 = s.moduli(1e-11, 20000)
        d2, _ = s.adjoint(1.0, 0.5, 1e-11, 20000, want_stiff=False)
    np.testing.assert_allclose(C2, C1, rtol=1e-9, atol=1e-12)
    assert K2 == pytest.approx(K1, rel=1e-10) and G2 == pytest.approx(G1, rel=1e-10)
    np.testing.assert_allclose(d2, d1, rtol=1e-8, atol=1e-12)
