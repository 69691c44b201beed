"""Pins of the oracle's weno_Re_flux branch (viscous velocity gradients by the
divergence theorem on WENO face values, reconstructed to the faces with WENO:
src/simulation/m_viscous.fpp:200-214, :372-591, m_rhs.fpp:514-528) beyond the
closed forms of tests/test_oracle_viscous.py, which hold for it too:

* every boundary code and WENO order gives a finite RHS (the gradient cells
  the reference reads but never writes hold NaN in the oracle, so a NaN
  reaching an output would show);
* WENO1 face values are the cell value, so every gradient is 0 and the RHS is
  bitwise the inviscid one;
* with periodic boundaries the viscous fluxes telescope (conservation);
* exchanging x and y (with the velocity components) exchanges the RHS;
* on rough data the branch differs from the finite differences (it is not
  silently the other branch).
"""
import numpy as np
import pytest

from oracle import oracle as O

GAMMA = 2.5


def _case(m, n, Re, bc_beg=(-3, -3), bc_end=(-3, -3), wo=5, wrf=True, nf=1):
    return O.Case(m=m, n=n, num_fluids=nf, weno_order=wo, gammas=[GAMMA] * nf, pi_infs=[0.0] * nf,
                  bc_beg=tuple(bc_beg) + (-3,), bc_end=tuple(bc_end) + (-3,), Re=Re, weno_Re_flux=wrf)


def _rough_state(case, seed):
    """Ghosted conservative state: rho, p near 1, velocities O(1) with
    cell-to-cell noise, single fluid (alpha = 1)."""
    rng = np.random.default_rng(seed)
    D = case.num_dims
    shp = case.interior_shape[1:]
    prim = np.zeros((case.sys_size,) + shp)
    prim[0] = 1.0 + 0.2 * rng.random(shp)
    for d in range(D):
        prim[1 + d] = rng.normal(size=shp)
    prim[1 + D] = 1.0 + 0.1 * rng.random(shp)
    prim[2 + D] = 1.0
    q = np.zeros(case.ghost_shape)
    b = case.buff_size
    sl = (slice(None), slice(0, 1)) + ((slice(b, -b),) if case.n > 0 else (slice(0, 1),)) + (slice(b, -b),)
    q[sl] = O.prim_to_cons(case, prim)
    return q


BCS = [((-1, -1), (-1, -1)), ((-2, -3), (-2, -3)), ((-4, -5), (-4, -6)), ((-6, -2), (-3, -4)), ((-5, -1), (-5, -1))]


@pytest.mark.parametrize("wo", [3, 5])
@pytest.mark.parametrize("bcs", BCS, ids=[f"x{b[0][0]}{b[1][0]}_y{b[0][1]}{b[1][1]}" for b in BCS])
def test_finite_for_every_boundary(bcs, wo):
    bb, be = bcs
    case = _case(25, 30, Re=[(0.05, 0.08)], bc_beg=(bb[0], bb[1]), bc_end=(be[0], be[1]), wo=wo)
    r = O.compute_rhs(case, O.uniform_grid(case), _rough_state(case, 3))
    assert np.all(np.isfinite(r))


@pytest.mark.parametrize("bc", [-1, -2, -3, -4])
def test_finite_1d(bc):
    case = O.Case(m=40, num_fluids=1, gammas=[GAMMA], pi_infs=[0.0], bc_beg=(bc, -3, -3), bc_end=(bc, -3, -3),
                  Re=[(0.05, 0.08)], weno_Re_flux=True)
    r = O.compute_rhs(case, O.uniform_grid(case), _rough_state(case, 4))
    assert np.all(np.isfinite(r))


def test_weno1_gradients_vanish():
    """WENO1 gives vL = vR = v, so every gradient is 0 * (1/ds) and the
    viscous fluxes add exact zeros: the RHS equals the inviscid one (same
    ghost width 2 for WENO1 with and without viscosity)."""
    case = _case(30, 20, Re=[(0.05, 0.08)], wo=1)
    inv = O.Case(**{**case.__dict__, "Re": None})
    assert case.buff_size == inv.buff_size == 2
    grid = O.uniform_grid(case)
    q = _rough_state(case, 5)
    np.testing.assert_array_equal(O.compute_rhs(case, grid, q.copy()), O.compute_rhs(inv, grid, q.copy()))


def test_periodic_conservation():
    """Every face flux, viscous ones included, enters its two cells with
    opposite signs: with periodic boundaries the momentum and energy RHS sum
    to zero up to rounding."""
    case = _case(31, 23, Re=[(0.05, 0.08)], bc_beg=(-1, -1), bc_end=(-1, -1))
    r = O.compute_rhs(case, O.uniform_grid(case), _rough_state(case, 6))
    for v in (1, 2, 3):
        assert abs(np.sum(r[v])) <= 1e-13 * np.sum(np.abs(r[v])), v


def test_transpose_exchanges_the_rhs():
    """A state transposed in x and y, with u and v exchanged, on the
    transposed grid (unequal cell widths), gives the transposed RHS with the
    momentum components exchanged, to rounding: the x- and y-gradient
    planes and their transverse uses are not mixed up."""
    c1 = _case(27, 19, Re=[(0.05, 0.08)], bc_beg=(-3, -2), bc_end=(-4, -6))
    c2 = _case(19, 27, Re=[(0.05, 0.08)], bc_beg=(-2, -3), bc_end=(-6, -4))
    q1 = _rough_state(c1, 7)
    q2 = np.ascontiguousarray(np.swapaxes(q1, 2, 3)[[0, 2, 1, 3, 4]])
    r1 = O.compute_rhs(c1, O.uniform_grid(c1), q1)
    r2 = O.compute_rhs(c2, O.uniform_grid(c2), q2)
    t = np.swapaxes(r2, 2, 3)[[0, 2, 1, 3, 4]]
    for v in range(5):
        # the single fluid's volume fraction has a zero RHS up to rounding:
        # measured on the mass scale
        scale = np.max(np.abs(r1[v if v < 4 else 0]))
        assert np.max(np.abs(t[v] - r1[v])) <= 1e-11 * scale, v


def test_differs_from_finite_differences():
    """On rough data the two branches give different viscous terms (and the
    same mass RHS, which has no viscous part)."""
    cT = _case(26, 22, Re=[(0.05, 0.08)], wrf=True)
    cF = _case(26, 22, Re=[(0.05, 0.08)], wrf=False)
    grid = O.uniform_grid(cT)
    q = _rough_state(cT, 8)
    rT, rF = O.compute_rhs(cT, grid, q.copy()), O.compute_rhs(cF, grid, q.copy())
    np.testing.assert_array_equal(rT[0], rF[0])
    assert np.max(np.abs(rT[1] - rF[1])) > 1e-3 * np.max(np.abs(rF[1]))
