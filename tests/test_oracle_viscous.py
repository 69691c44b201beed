"""Pins the oracle's viscous terms (finite-difference velocity gradients,
src/simulation/m_viscous.fpp:159-370, m_riemann_solvers.fpp:169-231 and
:683-902, m_rhs.fpp:591-603 and :639-651) by what the mathematics fixes.

For a velocity field that is polynomial of degree <= 2 (bilinear for the cross
terms), at uniform density and pressure, the one-sided differences and their
four-point averages are exact, so the viscous part of the RHS -- the oracle's
RHS with Reynolds numbers minus the same RHS without them -- equals the
Navier-Stokes stress divergence in closed form:
  momentum  d_j tau_ij,  tau = mu_s (grad u + grad u^T - 2/3 div u I) + mu_b div u I
  energy    d_j (u_i tau_ij)
with mu_s = 1/Re(1), mu_b = 1/Re(2).  A wrong 4/3 or 2/3, a sign, a
transposed gradient or a dropped bulk term fails one of these."""
import numpy as np
import pytest

from oracle import oracle as O

from helpers import reghost

GAMMA = 2.5  # fluid_pp%gamma = 1/(1.4 - 1)
INTERIOR = np.s_[3:-3]


def _case(D, N, Re, nf=1, bc=-3, wrf=False, Ny=None):
    n = (Ny or N) - 1 if D > 1 else 0
    return O.Case(m=N - 1, n=n, num_fluids=nf, gammas=[GAMMA] * nf, pi_infs=[0.0] * nf,
                  bc_beg=(bc,) * 3, bc_end=(bc,) * 3, Re=Re, weno_Re_flux=wrf)


# Both gradient branches: finite differences (weno_Re_flux = F) and the
# divergence theorem on WENO5 face values reconstructed to the faces with
# WENO5 (weno_Re_flux = T).  For velocities polynomial of degree <= 2 the
# latter is exact too: every 3-cell candidate of WENO5 reproduces a
# quadratic's face values whatever the weights, so vR - vL over ds is the
# derivative at the cell centre and its reconstruction (linear data) the
# derivative at the face.  The 2D cases use unequal cell widths in x and y.
WRF = pytest.mark.parametrize("wrf", [False, True], ids=["fd_gradients", "weno_re_flux"])


def _centres(case, grid):
    out = []
    b = case.buff_size
    for d, s in enumerate(case.extents()):
        cb, ds = grid[d]
        out.append(cb[b: b + s + 1] + ds[b: b + s + 1] / 2.0)
    return out


def _state(case, grid, vel, alphas=None):
    """Ghosted conservative state with rho = p = 1 and the given velocity
    functions of the cell centres (ghost cells are filled by the RHS)."""
    nf, D = case.num_fluids, case.num_dims
    c = _centres(case, grid)
    X, Y = (np.meshgrid(c[0], c[1], indexing="xy") if D > 1 else (c[0][None, :], None))
    prim = np.zeros((case.sys_size,) + X.shape)
    al = alphas or [1.0 / nf] * nf
    for i in range(nf):
        prim[i] = al[i]           # alpha_i rho_i with rho_i = 1
        prim[nf + D + 1 + i] = al[i]
    for d in range(D):
        prim[nf + d] = vel[d](X, Y)
    prim[nf + D] = 1.0
    q = np.zeros(case.ghost_shape)
    b = case.buff_size
    cons = O.prim_to_cons(case, prim)
    if D > 1:
        q[:, 0, b:-b, b:-b] = cons
    else:
        q[:, 0, 0, b:-b] = cons[:, 0, :]
    return q, X, Y


def _viscous_part(case, vel, alphas=None):
    grid = O.uniform_grid(case)
    q, X, Y = _state(case, grid, vel, alphas)
    inviscid = O.Case(**{**case.__dict__, "Re": None})
    gi, qi = reghost(case, inviscid, grid, q)
    r = O.compute_rhs(case, grid, q.copy()) - O.compute_rhs(inviscid, gi, qi)
    return r[:, 0], X, Y


def _inner(wrf):
    """Cells clear of the extrapolated ghost layers: the face gradients of
    weno_Re_flux reach 5 cells further (velocity stencil, then gradient
    stencil) than the finite differences."""
    return np.s_[6:-6] if wrf else INTERIOR


def _check(got, want, scale, tol=1e-9, inner=INTERIOR):
    g = got[inner, inner] if got.ndim == 2 and got.shape[0] > 1 else got[..., inner]
    w = want[inner, inner] if np.ndim(want) == 2 and np.shape(want)[0] > 1 else \
        (want[..., inner] if np.ndim(want) else want)
    err = np.max(np.abs(g - w)) / scale
    assert err < tol, err


@WRF
def test_1d_quadratic_velocity(wrf):
    """u = a x^2: momentum (4/3 mu_s + mu_b) u_xx; the energy term adds the
    HLLC face velocity times the stress, close to d_x(u tau_xx)."""
    a, R1, R2 = 0.3, 0.01, 0.02
    case = _case(1, 64, Re=[(R1, R2)], wrf=wrf)
    r, X, _ = _viscous_part(case, [lambda x, y: a * x * x])
    mu = 4.0 / 3.0 / R1 + 1.0 / R2
    _check(r[1], np.full_like(X, mu * 2 * a), mu * 2 * a, inner=_inner(wrf))
    # d_x(u mu u_x) = mu (u_x^2 + u u_xx) = 6 mu a^2 x^2, to the O(dx^2) of the
    # reconstructed face velocity
    _check(r[2], 6 * mu * a * a * X * X, 6 * mu * a * a, tol=2e-3, inner=_inner(wrf))
    assert np.all(r[0] == 0.0) and np.all(r[3] == 0.0)  # mass and volume fraction untouched


def test_1d_quartic_velocity_weno_re_flux():
    """u = a x^4 with weno_Re_flux: the theorem on WENO5 face values gives
    G_j = u'(x_j) (point values read as cell averages), whose WENO5 face
    value is q(x_{j+1/2}) with q = u' - h^2/24 u^(3), so the momentum term is
    mu (q(x_{j+1/2}) - q(x_{j-1/2}))/h = 12 mu a x^2 exactly at the ideal
    weights; the nonlinear weights add O(h^2) with a small constant.  The
    average of the wrong face values, 5e-1 (L(j) + R(j+1)), would add
    12 mu a h^2, some fifty times that."""
    a, R1, R2, N = 0.3, 0.01, 0.02, 64
    case = _case(1, N, Re=[(R1, R2)], wrf=True)
    r, X, _ = _viscous_part(case, [lambda x, y: a * x ** 4])
    mu = 4.0 / 3.0 / R1 + 1.0 / R2
    offset = 12 * mu * a / N ** 2
    _check(r[1], 12 * mu * a * X * X, offset, tol=0.1, inner=_inner(True))


@WRF
@pytest.mark.parametrize("which", ["couette_x", "couette_y"])
def test_2d_couette_heating(which, wrf):
    """u = a y (or v = a x): stresses are uniform, so the momentum terms
    vanish and the energy gains the viscous heating mu_s a^2 exactly."""
    a, R1 = 0.7, 0.05
    case = _case(2, 40, Re=[(R1, -1.0)], wrf=wrf, Ny=28)
    vel = [lambda x, y: a * y, lambda x, y: 0.0 * x] if which == "couette_x" else \
          [lambda x, y: 0.0 * x, lambda x, y: a * x]
    r, X, Y = _viscous_part(case, vel)
    heat = a * a / R1
    _check(r[1], 0.0, heat, inner=_inner(wrf))
    _check(r[2], 0.0, heat, inner=_inner(wrf))
    _check(r[3], np.full_like(X, heat), heat, inner=_inner(wrf))


@WRF
def test_2d_quadratic_shear_and_normal(wrf):
    """u = a y^2 + c x^2, v = e x^2 + f y^2: momentum
    x: mu_s (2a) + (4/3 mu_s + mu_b)(2c);  y: mu_s (2e) + (4/3 mu_s + mu_b)(2f)."""
    a, c, e, f = 0.4, -0.25, 0.15, 0.35
    R1, R2 = 0.02, 0.03
    case = _case(2, 40, Re=[(R1, R2)], wrf=wrf, Ny=28)
    r, X, _ = _viscous_part(case, [lambda x, y: a * y * y + c * x * x, lambda x, y: e * x * x + f * y * y])
    ms, mb = 1.0 / R1, 1.0 / R2
    wx = ms * 2 * a + (4.0 / 3.0 * ms + mb) * 2 * c
    wy = ms * 2 * e + (4.0 / 3.0 * ms + mb) * 2 * f
    _check(r[1], np.full_like(X, wx), abs(wx), inner=_inner(wrf))
    _check(r[2], np.full_like(X, wy), abs(wy), inner=_inner(wrf))


@WRF
def test_2d_bilinear_cross_terms(wrf):
    """u = a x y: tau_xx = 4/3 mu_s a y + mu_b a y, tau_xy = mu_s a x,
    tau_yy = -2/3 mu_s a y + mu_b a y; momentum x: 0, y: mu_s a/3 + mu_b a."""
    a, R1, R2 = 0.6, 0.02, 0.05
    case = _case(2, 40, Re=[(R1, R2)], wrf=wrf, Ny=28)
    r, X, _ = _viscous_part(case, [lambda x, y: a * x * y, lambda x, y: 0.0 * x])
    ms, mb = 1.0 / R1, 1.0 / R2
    scale = ms * a
    _check(r[1], np.zeros_like(X), scale, inner=_inner(wrf))
    _check(r[2], np.full_like(X, ms * a / 3.0 + mb * a), scale, inner=_inner(wrf))


@WRF
def test_two_fluid_reynolds_mixture(wrf):
    """Two fluids at alpha = (0.25, 0.75) with different shear Reynolds
    numbers: the face Reynolds number is 1/(0.25/R_1 + 0.75/R_2), so the
    Couette heating is a^2 (0.25/R_1 + 0.75/R_2); a fluid without Re(1)
    drops out of the sum."""
    a = 0.5
    for Re, mu in (([(0.01, -1.0), (0.04, -1.0)], 0.25 / 0.01 + 0.75 / 0.04),
                   ([(0.01, -1.0), (-1.0, -1.0)], 0.25 / 0.01)):
        case = _case(2, 32, Re=Re, nf=2, wrf=wrf)
        r, X, _ = _viscous_part(case, [lambda x, y: a * y, lambda x, y: 0.0 * x], alphas=[0.25, 0.75])
        heat = a * a * mu
        _check(r[2 + 2], np.full_like(X, heat), heat, inner=_inner(wrf))


@WRF
def test_2d_y_uniform_equals_1d(wrf):
    """A 2D field that does not vary in y gives the 1D viscous RHS in every
    row, bit for bit (the extra 2D terms add exact zeros; 16 rows make the y
    spacing, and so every y coefficient, exact)."""
    rng = np.random.default_rng(5)
    N = 48
    c1 = _case(1, N, Re=[(0.02, 0.03)], wrf=wrf)
    c2 = _case(2, N, Re=[(0.02, 0.03)], wrf=wrf)
    c2.n = 15
    g1, g2 = O.uniform_grid(c1), O.uniform_grid(c2)
    b = c1.buff_size
    prim1 = np.zeros((c1.sys_size, N))
    prim1[0] = 1.0 + 0.2 * rng.random(N)
    prim1[1] = rng.normal(size=N)
    prim1[2] = 1.0 + 0.1 * rng.random(N)
    prim1[3] = 1.0
    q1 = np.zeros(c1.ghost_shape)
    q1[:, 0, 0, b:-b] = O.prim_to_cons(c1, prim1)
    prim2 = np.zeros((c2.sys_size, c2.n + 1, N))
    prim2[0] = prim1[0]
    prim2[1] = prim1[1]
    prim2[3] = prim1[2]
    prim2[4] = 1.0
    q2 = np.zeros(c2.ghost_shape)
    q2[:, 0, b:-b, b:-b] = O.prim_to_cons(c2, prim2)
    r1 = O.compute_rhs(c1, g1, q1)[:, 0, 0]
    r2 = O.compute_rhs(c2, g2, q2)[:, 0]
    for j in range(c2.n + 1):
        np.testing.assert_array_equal(r2[[0, 1, 3, 4], j], r1)


def test_inviscid_entries_leave_rhs_unchanged():
    """Reynolds numbers <= 0 (the reference's default -1e6 included) are not
    viscous: the RHS is bitwise the inviscid one."""
    case = _case(2, 24, Re=[(-1e6, 0.0)])
    grid = O.uniform_grid(case)
    q, _, _ = _state(case, grid, [lambda x, y: x * y, lambda x, y: x - y])
    inv = O.Case(**{**case.__dict__, "Re": None})
    np.testing.assert_array_equal(O.compute_rhs(case, grid, q.copy()), O.compute_rhs(inv, grid, q.copy()))


def test_shear_wave_decays_at_the_discrete_viscous_rate():
    """v = A sin(2 pi x), u = 0 at rho = p = 1, periodic: the HLLC flux of
    the tangential momentum vanishes (equal pressure and zero normal
    velocity), so d(rho v)/dt is the viscous term alone, mu times the
    three-point second difference of v.  After RK3 steps the amplitude
    follows exp(-lambda t) with lambda = nu (2 - 2 cos(k dx)) / dx^2 (nu = mu
    here), up to the RK3 error and O(A^2) terms."""
    N, A, Re, dt, nsteps = 32, 1e-3, 10.0, 1e-3, 200
    case = _case(2, N, Re=[(Re, -1.0)], bc=-1)
    case.n = 15
    grid = O.uniform_grid(case)
    q, X, _ = _state(case, grid, [lambda x, y: 0.0 * x, lambda x, y: A * np.sin(2 * np.pi * x)])
    q2 = np.zeros_like(q)
    for t in range(nsteps):
        O.tvd_rk_step(case, grid, dt, q, q2, t_step=t)
    b = case.buff_size
    v = q[2, 0, b:-b, b:-b] / q[0, 0, b:-b, b:-b]
    amp = 2.0 / N * np.sum(v * np.sin(2 * np.pi * X), axis=1)
    dx = 1.0 / N
    lam = (1.0 / Re) * (2.0 - 2.0 * np.cos(2 * np.pi * dx)) / dx ** 2
    want = A * np.exp(-lam * dt * nsteps)
    assert np.max(np.abs(amp - want)) / want < 1e-5, (amp, want)


def _uniform_prim(case, alphas, rho_i, u, v, p):
    nf, D = case.num_fluids, case.num_dims
    prim = np.zeros(case.ghost_shape)
    for i in range(nf):
        prim[i] = alphas[i] * rho_i[i]
        prim[nf + D + 1 + i] = alphas[i]
    prim[nf] = u
    if D > 1:
        prim[nf + 1] = v
    prim[nf + D] = p
    return prim


def test_stability_criteria_closed_form():
    """ICFL, VCFL and Rc of a uniform state (m_data_output.fpp:214-263) with
    the harmonic mixture Reynolds numbers of m_variables_conversion.fpp:
    210-223, on a grid with dx != dy: ICFL = dt / min(dx/(|u|+c), dy/(|v|+c)),
    VCFL = max_i(dt/Re_i) / min(dx, dy)^2, Rc = min(dx(|u|+c), dy(|v|+c)) /
    max_i(1/Re_i), Re_i = 1/sum_f(alpha_f/Re_f,i) (1e16 for a kind no fluid
    has)."""
    gam = [2.5, 1.0 / 3.4]
    pinf = [0.0, 4.4 * 6e8 / 3.4 / 1e4]
    case = O.Case(m=31, n=15, num_fluids=2, gammas=gam, pi_infs=pinf, Re=[(100.0, 50.0), (200.0, -1.0)])
    grid = O.uniform_grid(case)
    dx, dy = 1.0 / 32, 1.0 / 16
    al, rr, u, v, p = (0.3, 0.7), (1.2, 1000.0), 3.0, -2.0, 4e4
    prim = _uniform_prim(case, al, rr, u, v, p)
    dt = 1e-6
    icfl, vcfl, rc = O.stability(case, grid, prim, dt)
    rho = sum(a * r for a, r in zip(al, rr))
    G = sum(a * g for a, g in zip(al, gam))
    P = sum(a * q for a, q in zip(al, pinf))
    c = np.sqrt(((G + 1) * p + P) / (G * rho))
    re1 = 1.0 / (al[0] / 100.0 + al[1] / 200.0)
    re2 = 1.0 / (al[0] / 50.0)
    assert icfl == pytest.approx(dt / min(dx / (abs(u) + c), dy / (abs(v) + c)), rel=1e-14)
    assert vcfl == pytest.approx(max(dt / re1, dt / re2) / min(dx, dy) ** 2, rel=1e-14)
    assert rc == pytest.approx(min(dx * (abs(u) + c), dy * (abs(v) + c)) / max(1 / re1, 1 / re2), rel=1e-14)
    # one fluid, shear only: the bulk Reynolds number is 1/sgm_eps
    one = O.Case(m=31, n=0, num_fluids=1, gammas=[2.5], pi_infs=[0.0], Re=[(100.0, -1.0)])
    prim1 = _uniform_prim(one, (1.0,), (1.5,), 3.0, 0.0, 4.0)
    icfl, vcfl, rc = O.stability(one, O.uniform_grid(one), prim1, dt)
    c1 = np.sqrt(1.4 * 4.0 / 1.5)
    assert vcfl == pytest.approx(max(dt / 100.0, dt / 1e16) / dx ** 2, rel=1e-14)
    assert rc == pytest.approx(dx * (3.0 + c1) * 100.0, rel=1e-14)
    assert icfl == pytest.approx(dt / dx * (3.0 + c1), rel=1e-14)
    # a NaN pressure makes the sound speed, ICFL and Rc of its cell NaN,
    # and the extrema NaN; an inviscid case reports zeros
    prim[4, 0, 7, 9] = np.nan
    icfl, vcfl, rc = O.stability(case, grid, prim, dt)
    assert np.isnan(icfl) and np.isnan(rc)
    inv = O.Case(m=31, n=15, num_fluids=2, gammas=gam, pi_infs=pinf)
    assert O.stability(inv, O.uniform_grid(inv), _uniform_prim(inv, al, rr, u, v, p), dt)[1:] == (0.0, 0.0)
