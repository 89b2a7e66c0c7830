"""Closed forms of the D3Q19 LBM used as checks (test infrastructure)."""
import numpy as np

# D3Q19 velocities in the reference's order (velocities19,
# lsm_lbm_static_alpha_beta_multipore.f:505-520)
CV = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
               [1, 1, 0], [-1, 1, 0], [1, -1, 0], [-1, -1, 0], [1, 0, 1], [-1, 0, 1], [1, 0, -1],
               [-1, 0, -1], [0, 1, 1], [0, -1, 1], [0, 1, -1], [0, -1, -1]])
KQ = np.array([1.0] + [1.0 / 6.0] * 6 + [1.0 / 12.0] * 12)


def bgk_one_step(inp, fr):
    """One step of a single-phase, all-fluid, triply periodic domain whose
    non-conserved moments all relax with rate 1: the post-collision state is
    f_eq(rho, j, u) + Fp, then streamed.
    f_eq_q = rho (t_q - k_q |u|^2 / 2) + k_q c_q.j + 3/2 k_q rho (c_q.u)^2,
    u = (j + F/2) / rho, t = (1 - 2 c^2, c^2/6, c^2/12), k = (1, 1/6, 1/12),
    Fp_q = k_q c_q.F (dt = 1).  ``fr`` is (nl, 19) with nodes x outer, z
    fastest."""
    nx, ny, nz = inp.ncx, inp.ncy, inp.ncz
    t = np.array([1 - 2 * inp.cr2] + [inp.cr2 / 6] * 6 + [inp.cr2 / 12] * 12)
    F = np.array(inp.bfr, float)
    rho = fr.sum(1)
    j = fr @ CV
    u = (j + 0.5 * F) / rho[:, None]
    u2 = (u * u).sum(1)
    cu, cj = u @ CV.T, j @ CV.T
    feq = rho[:, None] * (t[None, :] - KQ[None, :] * 0.5 * u2[:, None]) + KQ[None, :] * cj \
        + 1.5 * KQ[None, :] * rho[:, None] * cu * cu
    post = feq + (KQ * (CV @ F))[None, :] * (np.arange(19) > 0)
    grid = post.reshape(nx, ny, nz, 19)
    out = np.empty_like(grid)
    for q in range(19):
        out[..., q] = np.roll(grid[..., q], shift=tuple(CV[q]), axis=(0, 1, 2))
    return out.reshape(-1, 19)


def bgk_case(nx=6, ny=5, nz=7, seed=11):
    """All-fluid periodic box, tau = 1, body force off the axes, perturbed
    rest populations (all red): (ind, inp, fr, fb)."""
    from porescale_amd import io as pio
    ind = np.ones((nz, ny, nx), np.int32)
    inp = pio.LBMInput(sig=0.0, vir=1.0 / 6.0, vib=1.0 / 6.0, idir=3, ncx=nx, ncy=ny, ncz=nz, dt=1.0,
                       bfr=[2e-5, -1e-5, 3e-5], bfb=[2e-5, -1e-5, 3e-5])
    t = np.array([1 - 2 * inp.cr2] + [inp.cr2 / 6] * 6 + [inp.cr2 / 12] * 12)
    rng = np.random.default_rng(seed)
    fr = t[None, :] * (1.0 + 0.02 * rng.standard_normal((nx * ny * nz, 19)))
    return ind, inp, fr, np.zeros_like(fr)


def slit_case(nx=16, ny=4, nz=4, F=1e-5, nu=1.0 / 6.0):
    """Body-force slit: one solid plane x = 0, periodic, H = nx - 1 fluid
    layers, force F along z.  Returns (ind, inp, k) with k the permeability
    of the discrete Poiseuille profile between half-way bounce-back walls:
    u(s) = F/(2 nu) s (H - s) at s = 1/2 ... H - 1/2 has mean
    F/nu (H^2/12 + 1/24), and check_state divides the flux by all nx ny nz
    voxels, so k = (H/nx)(H^2/12 + 1/24).  A wall displaced by d changes k
    by about d/H relative."""
    from porescale_amd import io as pio
    ind = np.ones((nz, ny, nx), np.int32)
    ind[:, :, 0] = 0
    inp = pio.LBMInput(ncx=nx, ncy=ny, ncz=nz, sat=0.0, sig=0.0, vir=nu, vib=nu, idir=3, bfr=[0, 0, F],
                       bfb=[0, 0, F], dt=1.0)
    H = nx - 1
    return ind, inp, H / nx * (H * H / 12.0 + 1.0 / 24.0)


def int32_nsum_wraps(ind, iipore=1):
    """Diagonal links whose destination and axis projections are all fluid
    but whose int32 product of node ids m1*m2*m3 (L1:1118, 1293) wraps to 0:
    the reference would bounce them back.  Node ids are 1-based in the
    reference's order (x outer, z fastest).  Returns the count."""
    nz, ny, nx = ind.shape
    fl = np.transpose(ind, (2, 1, 0)) == iipore  # [x, y, z]
    mijk = np.zeros(fl.shape, np.uint64)
    mijk[fl] = np.arange(1, int(fl.sum()) + 1, dtype=np.uint64)
    xs, ys, zs = np.nonzero(fl)
    n = 0
    for c in CV[7:]:
        dest = mijk[(xs + c[0]) % nx, (ys + c[1]) % ny, (zs + c[2]) % nz]
        m1 = mijk[(xs + c[0]) % nx, ys, zs]
        m2 = mijk[xs, (ys + c[1]) % ny, zs]
        m3 = mijk[xs, ys, (zs + c[2]) % nz]
        ok = (dest != 0) & (m1 != 0) & (m2 != 0) & (m3 != 0)
        prod = (m1 * m2) % np.uint64(1 << 32)
        prod = (prod * m3) % np.uint64(1 << 32)
        n += int(np.count_nonzero(ok & (prod == 0)))
    return n
