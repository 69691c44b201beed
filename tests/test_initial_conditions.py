"""The 3D bubble-cloud initial condition: bubble placement and the slab-wise
generator against a whole-domain evaluation of the same definition."""
import numpy as np

from microfc_amd import ic


def test_cloud_centres_are_seeded_and_disjoint():
    c, r = ic.bubble_cloud_centres(64, seed=0)
    c2, r2 = ic.bubble_cloud_centres(64, seed=0)
    assert np.array_equal(c, c2) and np.array_equal(r, r2)
    assert len(r) == 64 and r.min() >= 0.02 and r.max() <= 0.05
    assert np.all(c[:, 0] - r >= 0.2) and np.all(c + r[:, None] <= 0.95) and np.all(c[:, 1:] - r[:, None] >= 0.05)
    d = np.sqrt(((c[:, None, :] - c[None, :, :]) ** 2).sum(-1))
    gap = d - r[:, None] - r[None, :]
    np.fill_diagonal(gap, np.inf)
    assert gap.min() >= 0.01
    assert not np.array_equal(ic.bubble_cloud_centres(64, seed=1)[0], c)


def test_cloud_slabs_equal_whole_domain_evaluation():
    N, nb = 36, 64
    prm, grids, q = ic.bubble_cloud_3d(N=N, nbubbles=nb, slab=5)
    xc, yc, zc = ic.cell_centres(prm, grids)
    Z, Y, X = np.meshgrid(zc, yc, xc, indexing="ij")
    eta = np.zeros((N, N, N))
    for c, R in zip(*ic.bubble_cloud_centres(nb)):
        r = np.sqrt((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2)
        eta = np.maximum(eta, ic._smooth_indicator(r, R, 1.0 / N))
    a_air = 1e-6 + (1.0 - 2e-6) * eta
    b = prm.buff_size
    inner = q[:, b:b + N, b:b + N, b:b + N]
    assert np.array_equal(inner[7], a_air)
    assert np.array_equal(inner[6], 1.0 - a_air)
    # ghost layers are left for the boundary fill
    assert not q[:, :b].any() and not q[:, :, :, :b].any()
    # pressure recovered from the conservative state (stiffened-gas mixture)
    rho = inner[0] + inner[1]
    ke = 0.5 * (inner[2] ** 2 + inner[3] ** 2 + inner[4] ** 2) / rho
    gam = inner[6] * prm.gammas[0] + inner[7] * prm.gammas[1]
    pinf = inner[6] * prm.pi_infs[0] + inner[7] * prm.pi_infs[1]
    p = (inner[5] - ke - pinf) / gam
    shocked = (xc < 0.15)[None, None, :]
    expect = np.broadcast_to(np.where(shocked, 1e9, 1e5), p.shape)
    assert np.allclose(p, expect, rtol=1e-6, atol=0)


def test_viscous_shear_matches_the_reference_example():
    """examples/2D_viscous/case.py: lower half both fluids at alpha 0.5 and
    alpha_rho 1000, u = 500; upper half fluid 2 only, alpha_rho 1, u = -500;
    p = 1e6; Re(1:2) = 1e-4 for both fluids; periodic x."""
    prm, grids, q = ic.viscous_shear_2d(N=32)
    assert prm.Re == [(1e-4, 1e-4), (1e-4, 1e-4)] and prm.viscous and prm.weno_Re_flux
    assert prm.bc_beg[0] == prm.bc_end[0] == -1
    qi = prm.interior(q)[:, 0]
    low, high = qi[:, 4, 4], qi[:, 28, 4]  # y = -0.36 and 0.39
    assert np.allclose(low[[0, 1, 2, 5, 6]], [1000.0, 1000.0, 2000.0 * 500.0, 0.5, 0.5])
    assert np.allclose(high[[0, 1, 2, 5, 6]], [0.0, 1.0, -500.0, 0.0, 1.0])
    # E = Gamma p + Pi + rho u^2 / 2 with Gamma = 1/3.4, Pi = 4.4 * 6e8 / 3.4
    G, P = 1 / 3.4, 4.4 * 6e8 / 3.4
    assert np.isclose(high[4], G * 1e6 + P + 0.5 * 500.0 ** 2)
    dt = ic.viscous_dt(prm, grids, q)
    assert 0 < dt < 0.2 * (1 / 32) ** 2 * 1.0 / (4 / 3 * 1e4) * 1.0001


def test_cloud_on_a_sub_box_equals_the_region_of_the_whole_box():
    """The cloud written on a sub-box at the whole box's spacing is that
    box's region, bit for bit (the refined sub-box judge of the cloud
    configuration relies on it)."""
    prm, g, q = ic.bubble_cloud_3d(N=32)
    sprm, sg, sq = ic.bubble_cloud_3d(N=16, lo=(0.0, 14 / 32, 4 / 32), hi=(0.5, 30 / 32, 20 / 32))
    np.testing.assert_array_equal(sprm.interior(sq), prm.interior(q)[:, 4:20, 14:30, 0:16])
