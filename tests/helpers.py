"""Shared helpers for the parity tests: case builders and the comparison.

The comparison metric is the per-variable relative L-infinity error
max|a - b| / max|b| (a zero reference variable compares absolutely), with
the FP64 bar 1e-12 per step; the number of differing bits is reported too.
"""
import sys

import numpy as np

from microfc_amd import ic
from microfc_amd.rhs import Params
from oracle import oracle as O

FP64_TOL = 1e-12


def to_oracle(prm: Params) -> O.Case:
    return O.Case(m=prm.m, n=prm.n, p=prm.p, num_fluids=prm.num_fluids, weno_order=prm.weno_order,
                  time_stepper=prm.time_stepper, t_step_stop=prm.t_step_stop, weno_eps=prm.weno_eps,
                  gammas=list(prm.gammas), pi_infs=list(prm.pi_infs), bc_beg=tuple(prm.bc_beg),
                  bc_end=tuple(prm.bc_end), Re=[tuple(r) for r in prm.Re] if prm.Re else None,
                  weno_Re_flux=prm.weno_Re_flux)


def rel_linf(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    out = []
    for v in range(a.shape[0]):
        den = np.max(np.abs(b[v]))
        num = np.max(np.abs(a[v] - b[v]))
        out.append(num / den if den > 0 else num)
    return np.array(out)


def state_linf(prm, a, b, q):
    """Per-variable L-inf error of a state a against b on the scales of the
    state q the step started from: max|alpha_i rho_i| for the partial
    densities, max(rho) * max(c) for every momentum component (a momentum
    that is zero in q, as in a fluid at rest, is measured on the momentum a
    sound wave would carry), max|E| for the energy and 1 for the volume
    fractions."""
    nf, D = prm.num_fluids, prm.num_dims
    qi = np.asarray(prm.interior(q), dtype=np.float64)
    rho = sum(qi[i] for i in range(nf))
    gam = sum(qi[nf + D + 1 + i] * prm.gammas[i] for i in range(nf))
    pinf = sum(qi[nf + D + 1 + i] * prm.pi_infs[i] for i in range(nf))
    dyn = sum(0.5 * qi[nf + d] ** 2 for d in range(D)) / rho
    pres = (qi[nf + D] - dyn - pinf) / gam
    c = np.sqrt(((gam + 1.0) * pres + pinf) / (gam * rho))
    mom = float(np.max(rho)) * float(np.max(c))
    den = [float(np.max(np.abs(qi[i]))) for i in range(nf)] + [mom] * D + [float(np.max(np.abs(qi[nf + D])))] + \
        [1.0] * nf
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.array([np.max(np.abs(a[v] - b[v])) / den[v] for v in range(a.shape[0])])


def scaled_linf(prm, a, b):
    """Per-variable L-inf error of a against b, normalised by a physical
    scale: max|b_v| for partial densities and energy, the largest momentum
    component for every momentum component, and for the volume fractions
    (whose reference rhs is round-off small in single-fluid regions) the
    larger of their own max|b_v| and the largest partial-density rhs
    max|b_mass|."""
    nf, D = prm.num_fluids, prm.num_dims
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    mom = np.max(np.abs(b[nf:nf + D]))
    div = max(np.max(np.abs(b[i])) for i in range(nf))
    out = []
    for v in range(a.shape[0]):
        if nf <= v < nf + D:
            den = mom
        elif v > nf + D:
            den = max(np.max(np.abs(b[v])), div)
        else:
            den = np.max(np.abs(b[v]))
        out.append(np.max(np.abs(a[v] - b[v])) / den)
    return np.array(out)


def ulp_diff(a, b):
    ai = np.asarray(a, dtype=np.float64).view(np.int64)
    bi = np.asarray(b, dtype=np.float64).view(np.int64)
    ai = np.where(ai < 0, np.int64(-2**63) - ai, ai)
    bi = np.where(bi < 0, np.int64(-2**63) - bi, bi)
    return int(np.max(np.abs(ai - bi))) if ai.size else 0


def assert_parity(a, b, tol=FP64_TOL, what=""):
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape, (a.shape, b.shape)
    assert np.all(np.isfinite(a)) and np.all(np.isfinite(b)), f"{what}: non-finite values"
    err = rel_linf(a, b)
    bitwise = bool(np.array_equal(a.view(np.int64), b.view(np.int64)))
    print(f"  {what}: rel_linf max {err.max():.3e} bitwise={bitwise} max_ulp={ulp_diff(a, b)}", flush=True)
    assert err.max() <= tol, f"{what}: rel L-inf {err} > {tol}"
    return bitwise


def perturbed(prm, q, seed=1, amp=1e-3):
    """Multiplies the interior partial densities and energy by 1 + amp*U(-1,1)
    so every cell carries distinct values (all variables stay physical)."""
    rng = np.random.default_rng(seed)
    qi = prm.interior(q)
    nf, D = prm.num_fluids, prm.num_dims
    for v in list(range(nf)) + [nf + D]:
        qi[v] *= 1.0 + amp * rng.uniform(-1.0, 1.0, qi[v].shape)
    return q


def stretched_grids(prm: Params, amps=(0.5, 0.3, 0.4)):
    """Non-uniform grids: s_cb(xi) = xi + a sin(2 pi xi) / (2 pi) on [0, 1]
    (cell widths varying by a factor (1+a)/(1-a)), ghosted by the case's BCs."""
    from microfc_amd.rhs import populate_grid_buffers, uniform_cell_boundaries
    grids = []
    for d, s in enumerate(prm.extents()):
        xi = uniform_cell_boundaries(s, 0.0, 1.0)
        cb = xi + amps[d] * np.sin(2.0 * np.pi * xi) / (2.0 * np.pi)
        grids.append(populate_grid_buffers(cb, prm.buff_size, prm.bc_beg[d], prm.bc_end[d]))
    return grids


def multi_fluid_2d(nf: int, N: int = 40, seed: int = 3):
    """nf stiffened-gas fluids in smooth random layers with a pressure pulse:
    every partial density, volume fraction and velocity component varies."""
    rng = np.random.default_rng(seed)
    gam = rng.uniform(1.3, 4.4, nf)
    pinf = np.where(np.arange(nf) % 2 == 0, 0.0, rng.uniform(1e3, 1e5, nf))
    prm = Params(m=N - 1, n=N - 1, num_fluids=nf, weno_order=5, gammas=list(1.0 / (gam - 1.0)),
                 pi_infs=list(gam * pinf / (gam - 1.0)), bc_beg=(-3, -1, -3), bc_end=(-3, -1, -3))
    grids = ic.uniform_grids(prm)
    X, Y, _ = ic.mesh(prm, grids)
    w = np.stack([1.0 + 0.5 * np.sin(2 * np.pi * (k + 1) * X + rng.uniform(0, 6)) *
                  np.cos(2 * np.pi * Y + rng.uniform(0, 6)) for k in range(nf)])
    alpha = w / w.sum(axis=0)
    prim = np.zeros(prm.interior_shape)
    prim[:nf] = alpha * rng.uniform(1.0, 1000.0, nf)[:, None, None, None]
    prim[nf] = 10.0 * np.sin(2 * np.pi * Y)
    prim[nf + 1] = 5.0 * np.cos(2 * np.pi * X)
    prim[nf + 2] = 1e5 * (1.0 + 0.5 * np.exp(-((X - 0.4) ** 2 + (Y - 0.6) ** 2) / 0.01))
    prim[nf + 3:] = alpha
    return prm, grids, ic.ghosted(prm, ic.prim_to_cons(prm, prim))


def multi_fluid_3d(nf: int, N: int = 26, seed: int = 4):
    """nf stiffened-gas fluids in smooth random 3D layers with a pressure
    pulse and all three velocity components, on a stretched grid with
    periodic x, reflective y and Riemann-extrapolation z boundaries."""
    rng = np.random.default_rng(seed)
    gam = rng.uniform(1.3, 4.4, nf)
    pinf = np.where(np.arange(nf) % 2 == 0, 0.0, rng.uniform(1e3, 1e5, nf))
    prm = Params(m=N - 1, n=N, p=N + 1, num_fluids=nf, weno_order=5, gammas=list(1.0 / (gam - 1.0)),
                 pi_infs=list(gam * pinf / (gam - 1.0)), bc_beg=(-1, -2, -4), bc_end=(-1, -2, -4))
    grids = stretched(prm, (0.3, 0.4, 0.5), (0.1, 0.9, 2.0))
    X, Y, Z = ic.mesh(prm, grids)
    w = np.stack([1.0 + 0.5 * np.sin(2 * np.pi * (k + 1) * X + rng.uniform(0, 6)) *
                  np.cos(2 * np.pi * (Y + Z) + rng.uniform(0, 6)) for k in range(nf)])
    alpha = w / w.sum(axis=0)
    prim = np.zeros(prm.interior_shape)
    prim[:nf] = alpha * rng.uniform(1.0, 1000.0, nf)[:, None, None, None]
    prim[nf] = 10.0 * np.sin(2 * np.pi * Y)
    prim[nf + 1] = 5.0 * np.cos(2 * np.pi * X) * np.sin(np.pi * Y)
    prim[nf + 2] = 3.0 * np.sin(2 * np.pi * (X + Z))
    prim[nf + 3] = 1e5 * (1.0 + 0.5 * np.exp(-((X - 0.4) ** 2 + (Y - 0.6) ** 2 + (Z - 0.5) ** 2) / 0.01))
    prim[nf + 4:] = alpha
    return prm, grids, ic.ghosted(prm, ic.prim_to_cons(prm, prim))


def stretched(case, amps, phases):
    """Non-uniform grids s_cb(xi) = f(xi) - f(0), f(xi) = xi + a sin(2 pi xi + phi)/(2 pi),
    per direction, ghosted by the case's BCs (case: an oracle Case or Params)."""
    out = []
    for d, s in enumerate(case.extents()):
        xi = np.arange(s + 2, dtype=np.float64) / (s + 1)
        f = lambda t: t + amps[d] * np.sin(2 * np.pi * t + phases[d]) / (2 * np.pi)  # noqa: E731
        cbi = f(xi) - f(0.0)
        cb, ds, _ = O.grid_buffers(cbi, case.buff_size, case.bc_beg[d], case.bc_end[d])
        out.append((cb, ds))
    return out


def smooth_prim_3d(case, seed, w_zero=False):
    """Two-fluid primitive field with a material interface (a sphere of
    water in air), a pressure pulse and non-zero velocities in every
    component (w optionally zero); every term is periodic or negligible at
    the domain faces, so periodic wraps carry no jump."""
    rng = np.random.default_rng(seed)
    m, n, p = case.m, case.n, case.p
    Z, Y, X = np.meshgrid((np.arange(p + 1) + 0.5) / (p + 1), (np.arange(n + 1) + 0.5) / (n + 1),
                          (np.arange(m + 1) + 0.5) / (m + 1), indexing="ij")
    r = np.sqrt((X - 0.5) ** 2 + (Y - 0.5) ** 2 + (Z - 0.5) ** 2)
    a1 = 1e-6 + (1 - 2e-6) * 0.5 * (1 + np.tanh((0.2 - r) / 0.1))
    prim = np.zeros(case.interior_shape)
    prim[0] = a1 * 1000.0 * (1 + 0.01 * np.sin(2 * np.pi * (X + 2 * Y + 3 * Z)))
    prim[1] = (1 - a1) * 1.2 * (1 + 0.02 * np.cos(2 * np.pi * (2 * X - Y + Z)))
    phase = rng.uniform(0, 2 * np.pi, 3)
    prim[2] = 40.0 + 30.0 * np.sin(2 * np.pi * Y + phase[0])
    prim[3] = -25.0 + 20.0 * np.cos(2 * np.pi * Z + phase[1])
    prim[4] = 0.0 if w_zero else 15.0 + 35.0 * np.sin(2 * np.pi * X + phase[2])
    prim[5] = 1e5 * (1.0 + 3.0 * np.exp(-((X - 0.45) ** 2 + (Y - 0.55) ** 2 + (Z - 0.5) ** 2) / 0.01))
    prim[6] = a1
    prim[7] = 1 - a1
    q = np.zeros(case.ghost_shape)
    b = case.buff_size
    q[:, b:b + p + 1, b:b + n + 1, b:b + m + 1] = O.prim_to_cons(case, prim)
    return q


def small_cases():
    """(name, params, grids, q) for the parity matrix; sizes keep the oracle
    to well under a second per RHS."""
    out = []
    prm, g, q = ic.sod_1d(m=199)
    out.append(("sod1d_w5_rk3", prm, g, perturbed(prm, q)))
    prm, g, q = ic.sod_1d(m=99, weno_order=3)
    prm.time_stepper = 2
    out.append(("sod1d_w3_rk2", prm, g, perturbed(prm, q)))
    prm, g, q = ic.sod_1d(m=49, weno_order=1)
    prm.time_stepper = 1
    out.append(("sod1d_w1_rk1", prm, g, perturbed(prm, q)))
    prm, g, q = ic.sod_1d(m=127)
    prm.bc_beg = (-2, -3, -3)
    prm.bc_end = (-4, -3, -3)
    g = ic.uniform_grids(prm)
    out.append(("sod1d_reflect_riemannextrap", prm, g, perturbed(prm, q)))
    prm, g, q = ic.shock_bubble_2d(N=64)
    out.append(("bubble2d_64", prm, g, perturbed(prm, q)))
    prm, g, q = ic.shock_bubble_2d(N=48)
    prm.bc_beg = (-1, -2, -3)
    prm.bc_end = (-1, -4, -3)
    g = ic.uniform_grids(prm)
    out.append(("bubble2d_periodic_reflect", prm, g, perturbed(prm, q)))
    prm, g, q = ic.shear_layer_2d(N=64)
    out.append(("shear2d_64_periodic", prm, g, perturbed(prm, q)))
    prm, g, q = ic.shock_droplet_3d(N=28)
    out.append(("droplet3d_28", prm, g, perturbed(prm, q)))
    prm, g, q = ic.shock_bubble_2d(N=48)
    out.append(("bubble2d_stretched", prm, stretched_grids(prm), perturbed(prm, q)))
    prm, g, q = ic.shock_droplet_3d(N=28)
    out.append(("droplet3d_stretched", prm, stretched_grids(prm), perturbed(prm, q)))
    prm, g, q = ic.bubble_cloud_3d(N=32, nbubbles=16)
    out.append(("cloud3d_32", prm, g, perturbed(prm, q)))
    prm, g, q = ic.shock_bubble_2d(N=40, weno_order=3)
    prm.time_stepper = 2
    out.append(("bubble2d_w3_rk2", prm, g, perturbed(prm, q)))
    prm, g, q = ic.shock_bubble_2d(N=40, weno_order=1)
    prm.bc_beg = (-1, -2, -3)
    prm.bc_end = (-1, -4, -3)
    g = ic.uniform_grids(prm)
    out.append(("bubble2d_w1_periodic_reflect", prm, g, perturbed(prm, q)))
    prm, g, q = ic.shock_droplet_3d(N=20, weno_order=3)
    prm.bc_beg = (-3, -1, -2)
    prm.bc_end = (-4, -1, -2)
    g = ic.uniform_grids(prm)
    out.append(("droplet3d_w3_mixed_bc", prm, g, perturbed(prm, q)))
    prm, g, q = ic.shock_droplet_3d(N=20, weno_order=1)
    prm.time_stepper = 1
    out.append(("droplet3d_w1_rk1", prm, g, perturbed(prm, q)))
    for nf in (3, 4, 6, 10):
        prm, g, q = multi_fluid_2d(nf)
        out.append((f"fluids{nf}_2d", prm, g, q))
    prm, g, q = multi_fluid_3d(5)
    out.append(("fluids5_3d_stretched", prm, g, q))
    out.extend(walls_3d_cases())
    out.extend(xy_tile_cases())
    return out


def one_fluid_3d_case(m, n, p, bcb, bce, seed=9):
    """One gas in 3D on a uniform grid: density, pressure and every velocity
    component varying smoothly, periodic in all three directions."""
    ga, pia = ic.AIR
    prm = Params(m=m, n=n, p=p, num_fluids=1, weno_order=5, gammas=[ga], pi_infs=[pia], bc_beg=tuple(bcb),
                 bc_end=tuple(bce))
    g = ic.uniform_grids(prm)
    rng = np.random.default_rng(seed)
    Z, Y, X = np.meshgrid((np.arange(p + 1) + 0.5) / (p + 1), (np.arange(n + 1) + 0.5) / (n + 1),
                          (np.arange(m + 1) + 0.5) / (m + 1), indexing="ij")
    ph = rng.uniform(0, 2 * np.pi, 5)
    prim = np.zeros(prm.interior_shape)
    prim[0] = 1.2 * (1 + 0.3 * np.sin(2 * np.pi * (X + Y) + ph[0]))
    prim[1] = 30.0 * np.sin(2 * np.pi * Y + ph[1])
    prim[2] = -20.0 * np.cos(2 * np.pi * Z + ph[2])
    prim[3] = 25.0 * np.sin(2 * np.pi * X + ph[3])
    prim[4] = 1e5 * (1.0 + 0.5 * np.cos(2 * np.pi * (X - Z) + ph[4]))
    prim[5] = 1.0
    q = np.zeros(prm.ghost_shape)
    b = prm.buff_size
    q[:, b:b + p + 1, b:b + n + 1, b:b + m + 1] = O.prim_to_cons(prm, prim)
    return prm, g, q


def xy_tile_cases():
    """Grids uniform bit for bit in x and y (power-of-two cell counts), so the
    contracted form takes its fused x and y sweep (one or two fluids): cell
    counts that leave partial 30 x 14 tiles, every boundary code in x and y,
    RK2, one fluid in 3D; three fluids on such a grid, which keep the
    separate sweeps; and grids uniform up to the rounding of their cell
    boundaries (WENO coefficients 1e-14 to 4e-14 apart), on which the
    contracted and FP32 forms take the uniform-grid kernels with the first
    cell's coefficients and the plain form the per-cell ones."""
    out = []
    prm, g, q = one_fluid_3d_case(31, 31, 31, (-1, -1, -1), (-1, -1, -1))
    out.append(("xy_tiles_3d_one_fluid", prm, g, q))
    prm, g, q = field_3d_case(31, 31, 31, (-2, -4, -1), (-4, -2, -1))
    out.append(("xy_tiles_3d_walls", prm, g, q))
    prm, g, q = field_3d_case(63, 31, 31, (-1, -1, -3), (-1, -1, -3), time_stepper=2)
    out.append(("xy_tiles_3d_periodic_rk2", prm, g, q))
    prm, g, q = ic.shock_bubble_2d(N=64)
    prm.bc_beg = (-2, -4, -3)
    prm.bc_end = (-4, -2, -3)
    g = ic.uniform_grids(prm)
    out.append(("xy_tiles_2d_walls", prm, g, perturbed(prm, q)))
    prm, g, q = multi_fluid_2d(3, N=64)
    out.append(("xy_tiles_fluids3_2d", prm, g, q))
    prm, g, q = field_3d_case(47, 39, 29, (-1, -1, -1), (-1, -1, -1))
    out.append(("xy_tiles_3d_rounded_grid", prm, g, q))
    prm, g, q = ic.shock_bubble_2d(N=96)
    out.append(("xy_tiles_2d_rounded_grid", prm, g, perturbed(prm, q)))
    return out


def field_3d_case(m, n, p, bcb, bce, weno_order=5, grid="uniform", seed=7, time_stepper=3):
    """Two fluids in 3D with non-zero u, v, w at every boundary (the smooth
    field of smooth_prim_3d), on a uniform or a stretched grid."""
    gw, piw = ic.WATER
    ga, pia = ic.AIR
    prm = Params(m=m, n=n, p=p, num_fluids=2, weno_order=weno_order, time_stepper=time_stepper,
                 gammas=[gw, ga], pi_infs=[piw, pia], bc_beg=tuple(bcb), bc_end=tuple(bce))
    g = ic.uniform_grids(prm) if grid == "uniform" else stretched(prm, (0.5, 0.3, 0.4), (0.2, 1.3, 2.9))
    return prm, g, smooth_prim_3d(prm, seed)


def walls_3d_cases():
    """3D cases whose reflective y and z walls carry non-zero transverse and
    normal momentum, and z boundaries that are periodic (-1) and Riemann
    extrapolation (-4), in the uniform-grid and the per-cell-coefficient
    kernels."""
    out = []
    prm, g, q = field_3d_case(27, 25, 26, (-3, -2, -2), (-4, -2, -2))
    out.append(("field3d_walls_yz", prm, g, q))
    prm, g, q = field_3d_case(25, 29, 27, (-4, -3, -1), (-3, -4, -1), grid="stretched")
    out.append(("field3d_z_periodic_stretched", prm, g, q))
    prm, g, q = field_3d_case(19, 15, 23, (-1, -2, -4), (-1, -2, -4), weno_order=3, time_stepper=2)
    out.append(("field3d_w3_z_riemann", prm, g, q))
    prm, g, q = field_3d_case(26, 25, 28, (-2, -2, -2), (-2, -2, -2), grid="stretched")
    out.append(("field3d_all_walls_stretched", prm, g, q))
    return out


def reghost(src, dst, grids, q):
    """The same interior cells and cell boundaries laid out with dst's ghost
    width (src and dst are oracle Cases or Params of one grid, e.g. a viscous
    case and its inviscid twin, whose buff_size differs)."""
    from microfc_amd.rhs import populate_grid_buffers
    bs, bd = src.buff_size, dst.buff_size
    new_grids = []
    for d, s in enumerate(src.extents()):
        cbi = grids[d][0][bs:bs + s + 2]  # s_cb(-1..s)
        new_grids.append(populate_grid_buffers(cbi, bd, dst.bc_beg[d], dst.bc_end[d]))
    js, ks = (bs if src.n > 0 else 0), (bs if src.p > 0 else 0)
    jd, kd = (bd if dst.n > 0 else 0), (bd if dst.p > 0 else 0)
    qn = np.zeros(dst.ghost_shape)
    qn[:, kd:kd + dst.p + 1, jd:jd + dst.n + 1, bd:bd + dst.m + 1] = \
        q[:, ks:ks + src.p + 1, js:js + src.n + 1, bs:bs + src.m + 1]
    return new_grids, qn


def relayout(prm, grids, q):
    """Grids and field laid out with prm's current ghost width, whatever
    width they were built with (the width is read from the array shapes):
    for cases made inviscid and then given Reynolds numbers."""
    from microfc_amd.rhs import populate_grid_buffers
    bd = prm.buff_size
    new_grids = []
    for d, s in enumerate(prm.extents()):
        cb = grids[d][0]
        bg = (cb.size - (s + 2)) // 2
        new_grids.append(populate_grid_buffers(cb[bg:bg + s + 2], bd, prm.bc_beg[d], prm.bc_end[d]))
    bq = (q.shape[-1] - prm.m - 1) // 2
    jq, kq = (bq if prm.n > 0 else 0), (bq if prm.p > 0 else 0)
    jd, kd = (bd if prm.n > 0 else 0), (bd if prm.p > 0 else 0)
    qn = np.zeros(prm.ghost_shape)
    qn[:, kd:kd + prm.p + 1, jd:jd + prm.n + 1, bd:bd + prm.m + 1] = \
        q[:, kq:kq + prm.p + 1, jq:jq + prm.n + 1, bq:bq + prm.m + 1]
    return new_grids, qn


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()
