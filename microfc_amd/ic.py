"""Initial conditions for the benchmark cases, written straight into the
ghosted conservative layout the library consumes.

prim_to_cons restates s_convert_primitive_to_conservative_variables
(src/common/m_variables_conversion.fpp:385-454), the pre_process direction.
The smoothed circle/sphere patch follows the tanh smoothing of
src/pre_process/m_create_patches.fpp:122-128 (eta = tanh(smooth_coeff/min(dx)
*(r - R))*(-0.5) + 0.5) and the mixing rule of m_assign_patches.fpp:54-165
(a smoothed patch mixes the patch and background primitive states by eta).
Configurations are the benchmark's cases; each returns (params, grids, q)
with q a float64 numpy array of the ghosted shape (ghost cells zero).
"""
from __future__ import annotations

import math

import numpy as np

from .rhs import Params, populate_grid_buffers, uniform_cell_boundaries

# Stiffened-gas inputs in MFC's convention (examples/1D_kapilashocktube/case.py:60-63):
# gamma = 1/(g-1), pi_inf = g*p_inf/(g-1).
def eos_inputs(g: float, p_inf: float = 0.0):
    return 1.0 / (g - 1.0), g * p_inf / (g - 1.0)


WATER = eos_inputs(4.4, 6.0e8)
AIR = eos_inputs(1.4, 0.0)


def prim_to_cons(params: Params, prim):
    """Primitive -> conservative, vectorised over the trailing axes."""
    nf, D = params.num_fluids, params.num_dims
    q = np.array(prim, dtype=np.float64, copy=True)
    rho = np.zeros(q.shape[1:])
    gam = np.zeros(q.shape[1:])
    pinf = np.zeros(q.shape[1:])
    for i in range(nf):
        rho = rho + prim[i]
        gam = gam + prim[nf + D + 1 + i] * params.gammas[i]
        pinf = pinf + prim[nf + D + 1 + i] * params.pi_infs[i]
    dyn = np.zeros(q.shape[1:])
    for i in range(nf, nf + D):
        q[i] = rho * prim[i]
        dyn = dyn + q[i] * prim[i] / 2.0
    q[nf + D] = gam * prim[nf + D] + dyn + pinf
    return q


def uniform_grids(params: Params, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0)):
    grids = []
    for d, s in enumerate(params.extents()):
        cbi = uniform_cell_boundaries(s, lo[d], hi[d])
        grids.append(populate_grid_buffers(cbi, params.buff_size, params.bc_beg[d], params.bc_end[d]))
    return grids


def cell_centres(params: Params, grids):
    b = params.buff_size
    out = []
    for d, s in enumerate(params.extents()):
        cb = grids[d][0]
        out.append(0.5 * (cb[b: b + s + 1] + cb[b + 1: b + s + 2]))
    return out


def mesh(params: Params, grids):
    """Cell-centre coordinate arrays broadcast to (p+1, n+1, m+1)."""
    cc = cell_centres(params, grids)
    shp = params.interior_shape[1:]
    X = np.broadcast_to(cc[0][None, None, :], shp)
    Y = np.broadcast_to(cc[1][None, :, None], shp) if params.num_dims > 1 else None
    Z = np.broadcast_to(cc[2][:, None, None], shp) if params.num_dims > 2 else None
    return X, Y, Z


def ghosted(params: Params, interior_cons):
    q = np.zeros(params.ghost_shape)
    b = params.buff_size
    bj = b if params.n > 0 else 0
    bk = b if params.p > 0 else 0
    q[:, bk:bk + params.p + 1, bj:bj + params.n + 1, b:b + params.m + 1] = interior_cons
    return q


def sod_1d(m: int = 999, weno_order: int = 5):
    """1D Sod: x < 0.5 rho=1, p=1; x >= 0.5 rho=0.125, p=0.1; gamma 1.4."""
    g, pi = AIR
    prm = Params(m=m, num_fluids=1, weno_order=weno_order, gammas=[g], pi_infs=[pi])
    grids = uniform_grids(prm)
    X, _, _ = mesh(prm, grids)
    prim = np.zeros(prm.interior_shape)
    left = X < 0.5
    prim[0] = np.where(left, 1.0, 0.125)
    prim[1] = 0.0
    prim[2] = np.where(left, 1.0, 0.1)
    prim[3] = 1.0
    return prm, grids, ghosted(prm, prim_to_cons(prm, prim))


def stiffened_hugoniot_density(rho1, p1, p2, g, p_inf):
    """Post-shock density of a stiffened gas from the Rankine-Hugoniot relation."""
    a = (p2 + p_inf) / (p1 + p_inf)
    return rho1 * ((g + 1.0) * a + (g - 1.0)) / ((g - 1.0) * a + (g + 1.0))


def _smooth_indicator(r, R, dxmin, smooth_cells=2.0):
    """eta = 1 inside, 0 outside, tanh profile of about smooth_cells cells."""
    return 0.5 - 0.5 * np.tanh((r - R) / (0.5 * smooth_cells * dxmin))


def shock_bubble_2d(N: int = 2048, weno_order: int = 5, alpha_min: float = 1e-6, precision: int = 0):
    """Water column with an air bubble (R = 0.1 at (0.5, 0.5)) hit by a planar
    water shock at x = 0.3 (post-shock p = 1e9, u = 200; p = 1e5 ahead).
    Fluid 1 water, fluid 2 air; extrapolation boundaries."""
    gw, piw = WATER
    ga, pia = AIR
    prm = Params(m=N - 1, n=N - 1, num_fluids=2, weno_order=weno_order, gammas=[gw, ga], pi_infs=[piw, pia],
                 bc_beg=(-3, -3, -3), bc_end=(-3, -3, -3), precision=precision)
    grids = uniform_grids(prm)
    X, Y, _ = mesh(prm, grids)
    dxmin = 1.0 / N
    r = np.sqrt((X - 0.5) ** 2 + (Y - 0.5) ** 2)
    eta = _smooth_indicator(r, 0.1, dxmin)
    a_air = alpha_min + (1.0 - 2.0 * alpha_min) * eta
    a_wat = 1.0 - a_air
    rho_w1, p1 = 1000.0, 1.0e5
    rho_w2 = stiffened_hugoniot_density(rho_w1, p1, 1.0e9, 4.4, 6.0e8)
    shocked = X < 0.3
    prim = np.zeros(prm.interior_shape)
    prim[0] = a_wat * np.where(shocked, rho_w2, rho_w1)
    prim[1] = a_air * 1.0
    prim[2] = np.where(shocked, 200.0, 0.0)
    prim[3] = 0.0
    prim[4] = np.where(shocked, 1.0e9, p1)
    prim[5] = a_wat
    prim[6] = a_air
    return prm, grids, ghosted(prm, prim_to_cons(prm, prim))


def shock_droplet_3d(N: int = 512, weno_order: int = 5, alpha_min: float = 1e-6, precision: int = 0,
                     window=None, shock_x: float = 0.2):
    """Water droplet (R = 0.1, centre of [0,1]^3) in air (rho 1.2, p 1e5) hit
    by a Mach 2.4 air shock at x = shock_x (Rankine-Hugoniot post-shock state
    of the air, imposed on every fluid behind the shock).
    window = ((i0, j0, k0), (mx, my, mz)) makes only that box of the N^3 grid,
    as sub_box would cut it out of the whole case (same cell boundaries,
    extrapolation at the cut faces)."""
    gw, piw = WATER
    ga, pia = AIR
    prm = Params(m=N - 1, n=N - 1, p=N - 1, num_fluids=2, weno_order=weno_order, gammas=[gw, ga],
                 pi_infs=[piw, pia], precision=precision)
    grids = uniform_grids(prm)
    if window is not None:
        (lo, shape) = window
        full = prm
        prm = Params(m=shape[0] - 1, n=shape[1] - 1, p=shape[2] - 1, num_fluids=2, weno_order=weno_order,
                     gammas=[gw, ga], pi_infs=[piw, pia], precision=precision)
        b = full.buff_size
        grids = [populate_grid_buffers(grids[d][0][lo[d] + b:lo[d] + b + shape[d] + 1], prm.buff_size,
                                       prm.bc_beg[d], prm.bc_end[d]) for d in range(3)]
    X, Y, Z = mesh(prm, grids)
    r = np.sqrt((X - 0.5) ** 2 + (Y - 0.5) ** 2 + (Z - 0.5) ** 2)
    eta = _smooth_indicator(r, 0.1, 1.0 / N)
    a_wat = alpha_min + (1.0 - 2.0 * alpha_min) * eta
    a_air = 1.0 - a_wat
    g, M, rho1, p1 = 1.4, 2.4, 1.2, 1.0e5
    c1 = math.sqrt(g * p1 / rho1)
    p2 = p1 * (1.0 + 2.0 * g / (g + 1.0) * (M * M - 1.0))
    rho2 = rho1 * (g + 1.0) * M * M / ((g - 1.0) * M * M + 2.0)
    u2 = M * c1 * (1.0 - rho1 / rho2)
    shocked = X < shock_x
    prim = np.zeros(prm.interior_shape)
    prim[0] = a_wat * 1000.0
    prim[1] = a_air * np.where(shocked, rho2, rho1)
    prim[2] = np.where(shocked, u2, 0.0)
    prim[5] = np.where(shocked, p2, p1)
    prim[6] = a_wat
    prim[7] = a_air
    return prm, grids, ghosted(prm, prim_to_cons(prm, prim))


def shear_layer_2d(N: int = 4096, weno_order: int = 5, seed: int = 0, modes: int = 16, delta: float = 0.01,
                   precision: int = 0):
    """Kelvin-Helmholtz shear layer, one gas (gamma 1.4): u = tanh((y-0.5)/delta),
    rho = 1 below / 0.5 above y = 0.5 (smoothed by the same tanh), p = 1, and a
    seeded multimode v perturbation (amplitudes U(0,1e-3), random phases,
    Gaussian envelope around the layer).  Periodic in x, extrapolation in y."""
    g, pi = AIR
    prm = Params(m=N - 1, n=N - 1, num_fluids=1, weno_order=weno_order, gammas=[g], pi_infs=[pi],
                 bc_beg=(-1, -3, -3), bc_end=(-1, -3, -3), precision=precision)
    grids = uniform_grids(prm)
    X, Y, _ = mesh(prm, grids)
    rng = np.random.default_rng(seed)
    amp = rng.uniform(0.0, 1e-3, modes)
    ph = rng.uniform(0.0, 2.0 * math.pi, modes)
    v = np.zeros(prm.interior_shape[1:])
    for k in range(modes):
        v = v + amp[k] * np.sin(2.0 * math.pi * (k + 1) * X + ph[k])
    v = v * np.exp(-((Y - 0.5) / (4.0 * delta)) ** 2)
    t = np.tanh((Y - 0.5) / delta)
    prim = np.zeros(prm.interior_shape)
    prim[0] = 1.0 - 0.25 * (1.0 + t)
    prim[1] = t
    prim[2] = v
    prim[3] = 1.0
    prim[4] = 1.0
    return prm, grids, ghosted(prm, prim_to_cons(prm, prim))


def viscous_shear_case(N: int = 2048, weno_order: int = 5, Re: float = 1e-4, weno_Re_flux: bool = True) -> dict:
    """The reference's examples/2D_viscous/case.py as a case dictionary at
    N^2 cells (N = 51 is the example itself): [-0.5, 0.5]^2, two water-like
    stiffened gases (gamma 4.4, p_inf 6e8) with shear and volume Reynolds
    numbers Re; patch 1 fills the domain with both fluids at alpha = 0.5,
    alpha_rho = 1000 each, u = +500; patch 2 (centred at y = 0.25, length
    0.5, allowed to overwrite patch 1) holds fluid 2 alone at alpha_rho = 1,
    u = -500; p = 1e6 everywhere.  Periodic in x; bc -6 in y (extrapolation
    in the ghost fill).  The example's logistics and algorithm keys come
    along: dt 5e-10, 10 RK3 steps saved every step, run_time_info,
    parallel_io and weno_Re_flux (WENO-reconstructed velocity gradients;
    False selects finite differences)."""
    g, pi = eos_inputs(4.4, 6.0e8)
    c = {"case_dir": "'.'", "run_time_info": "T",
         "x_domain%beg": -0.5, "x_domain%end": 0.5, "y_domain%beg": -0.5, "y_domain%end": 0.5,
         "m": N - 1, "n": N - 1, "dt": 5e-10, "t_step_start": 0, "t_step_stop": 10, "t_step_save": 1,
         "num_patches": 2, "num_fluids": 2, "time_stepper": 3, "weno_order": weno_order, "weno_eps": 1e-16,
         "weno_Re_flux": "T" if weno_Re_flux else "F",
         "bc_x%beg": -1, "bc_x%end": -1, "bc_y%beg": -6, "bc_y%end": -6,
         "format": 1, "precision": 2, "prim_vars_wrt": "T", "parallel_io": "T"}
    patches = [dict(geometry=3, x_centroid=0.0, y_centroid=0.0, length_x=1.0, length_y=1.0,
                    **{"vel(1)": 500.0, "vel(2)": 0.0, "pres": 1e6, "alpha_rho(1)": 1000.0,
                       "alpha_rho(2)": 1000.0, "alpha(1)": 0.5, "alpha(2)": 0.5}),
               dict(geometry=3, x_centroid=0.0, y_centroid=0.25, length_x=1.0, length_y=0.5,
                    **{"alter_patch(1)": "T", "vel(1)": -500.0, "vel(2)": 0.0, "pres": 1e6,
                       "alpha_rho(1)": 0.0, "alpha_rho(2)": 1.0, "alpha(1)": 0.0, "alpha(2)": 1.0})]
    for k, p in enumerate(patches, start=1):
        for key, v in p.items():
            c[f"patch_icpp({k})%{key}"] = v
    for f in (1, 2):
        c[f"fluid_pp({f})%gamma"], c[f"fluid_pp({f})%pi_inf"] = g, pi
        c[f"fluid_pp({f})%Re(1)"] = c[f"fluid_pp({f})%Re(2)"] = Re
    return c


def viscous_shear_2d(N: int = 2048, weno_order: int = 5, Re: float = 1e-4, weno_Re_flux: bool = True):
    """The reference's viscous example at N^2 cells (viscous_shear_case),
    laid out by the pre-processor's patch rules."""
    from .pre_process import PreProcess
    return PreProcess(viscous_shear_case(N, weno_order, Re, weno_Re_flux)).state()


def viscous_dt(prm: Params, grids, q, safety: float = 0.2) -> float:
    """Explicit viscous time-step bound safety * min(dx)^2 / nu_max, with
    nu = (4/3 mu_s + mu_b) / rho and mu = 1/Re of the most viscous fluid (the
    CFL reduction bounds only the inviscid wave speeds)."""
    nf = prm.num_fluids
    mu = 0.0
    for pair in prm.Re or []:
        ms = 1.0 / pair[0] if pair[0] > 0 else 0.0
        mb = 1.0 / pair[1] if pair[1] > 0 else 0.0
        mu = max(mu, 4.0 / 3.0 * ms + mb)
    rho = sum(prm.interior(q)[i] for i in range(nf))
    dx = min(float(np.min(g[1][prm.buff_size: -prm.buff_size])) for g in grids)
    return safety * dx * dx * float(np.min(rho)) / mu if mu > 0 else math.inf


def bubble_cloud_centres(nbubbles: int = 64, seed: int = 0, r_lo: float = 0.02, r_hi: float = 0.05,
                         gap: float = 0.01, x_lo: float = 0.2, margin: float = 0.05):
    """Seeded, non-overlapping bubble centres and radii in [0,1]^3: radius
    U(r_lo, r_hi), centre uniform in [x_lo+R, 1-margin-R] x [margin+R,
    1-margin-R]^2 (ahead of the shock), redrawn until its surface is at
    least `gap` from every earlier bubble."""
    rng = np.random.default_rng(seed)
    cs, rs = [], []
    for _ in range(1000 * nbubbles):
        if len(rs) == nbubbles:
            break
        R = rng.uniform(r_lo, r_hi)
        c = np.array([rng.uniform(x_lo + R, 1.0 - margin - R), rng.uniform(margin + R, 1.0 - margin - R),
                      rng.uniform(margin + R, 1.0 - margin - R)])
        if all(np.sqrt(np.sum((c - c2) ** 2)) >= R + R2 + gap for c2, R2 in zip(cs, rs)):
            cs.append(c)
            rs.append(R)
    if len(rs) != nbubbles:
        raise ValueError(f"could not place {nbubbles} non-overlapping bubbles")
    return np.array(cs), np.array(rs)


def bubble_cloud_3d(N: int = 768, nbubbles: int = 64, seed: int = 0, weno_order: int = 5,
                    alpha_min: float = 1e-6, x_shock: float = 0.15, precision: int = 0, slab: int = 32,
                    lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0)):
    """Cloud of air bubbles (rho 1, p 1e5; radii U(0.02, 0.05), seeded
    non-overlapping centres ahead of the shock) in water (rho 1000, p 1e5)
    hit by a planar water shock at x = x_shock (post-shock p = 1e9, u = 200,
    Rankine-Hugoniot density), [0,1]^3, extrapolation boundaries.  Each
    bubble's smoothed indicator is evaluated within R + 24 dx of its centre
    (beyond it tanh rounds to 1 and the indicator is exactly 0 in FP64); the
    cloud's indicator is the maximum over bubbles.  Written slab by slab in z, so the host holds
    one ghosted state and a slab of temporaries (768^3: 30 GB + 1.2 GB).
    lo, hi: a sub-box of the unit cube to grid instead (same cloud and shock;
    N cells per direction, cubic cells when the box is)."""
    gw, piw = WATER
    ga, pia = AIR
    prm = Params(m=N - 1, n=N - 1, p=N - 1, num_fluids=2, weno_order=weno_order, gammas=[gw, ga],
                 pi_infs=[piw, pia], precision=precision)
    grids = uniform_grids(prm, lo=lo, hi=hi)
    xc, yc, zc = cell_centres(prm, grids)
    dxmin = min(h - l for l, h in zip(lo, hi)) / N
    centres, radii = bubble_cloud_centres(nbubbles, seed)
    rho_w1, p1 = 1000.0, 1.0e5
    rho_w2 = stiffened_hugoniot_density(rho_w1, p1, 1.0e9, 4.4, 6.0e8)
    shocked = (xc < x_shock)[None, None, :]
    q = np.zeros(prm.ghost_shape)
    b = prm.buff_size
    reach = 24.0 * dxmin


    for k0 in range(0, N, slab):
        k1 = min(N, k0 + slab)
        eta = np.zeros((k1 - k0, N, N))
        for c, R in zip(centres, radii):
            r_reach = R + reach
            za, zb = np.searchsorted(zc, c[2] - r_reach), np.searchsorted(zc, c[2] + r_reach, side="right")
            za, zb = max(za, k0), min(zb, k1)
            if za >= zb:
                continue
            ya, yb = np.searchsorted(yc, c[1] - r_reach), np.searchsorted(yc, c[1] + r_reach, side="right")
            xa, xb = np.searchsorted(xc, c[0] - r_reach), np.searchsorted(xc, c[0] + r_reach, side="right")
            Z = zc[za:zb, None, None]
            Y = yc[None, ya:yb, None]
            X = xc[None, None, xa:xb]
            r = np.sqrt((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2)
            sub = eta[za - k0:zb - k0, ya:yb, xa:xb]
            np.maximum(sub, _smooth_indicator(r, R, dxmin), out=sub)
        a_air = alpha_min + (1.0 - 2.0 * alpha_min) * eta
        a_wat = 1.0 - a_air
        shp = eta.shape
        prim = np.zeros((prm.sys_size,) + shp)
        prim[0] = a_wat * np.where(shocked, rho_w2, rho_w1)
        prim[1] = a_air * 1.0
        prim[2] = np.broadcast_to(np.where(shocked, 200.0, 0.0), shp)
        prim[5] = np.broadcast_to(np.where(shocked, 1.0e9, p1), shp)
        prim[6] = a_wat
        prim[7] = a_air
        q[:, b + k0:b + k1, b:b + N, b:b + N] = prim_to_cons(prm, prim)
    return prm, grids, q


def sub_box(prm: Params, grids, q, lo, shape):
    """A window of a case as a case of its own: interior cells lo[d] ..
    lo[d]+shape[d]-1 in each direction, the same cell boundaries (so the same
    cell widths and WENO coefficients), the same fluids and form, and the
    window's conservative values; a face the window cuts through the domain
    gets an extrapolation boundary (-3), a face on the domain's boundary keeps
    its boundary condition."""
    import copy
    D = prm.num_dims
    sub = copy.deepcopy(prm)
    ext = [prm.m, prm.n, prm.p]
    bcb, bce = list(prm.bc_beg), list(prm.bc_end)
    new_ext = [0, 0, 0]
    new_grids = []
    for d in range(D):
        l0, s = lo[d], shape[d]
        if l0 < 0 or s < 1 or l0 + s > ext[d] + 1:
            raise ValueError(f"window {l0}+{s} outside 0..{ext[d]} in direction {d}")
        new_ext[d] = s - 1
        if l0 > 0 or bcb[d] == -1 and s < ext[d] + 1:
            bcb[d] = -3
        if l0 + s < ext[d] + 1 or bce[d] == -1 and s < ext[d] + 1:
            bce[d] = -3
        cb = grids[d][0]
        b = prm.buff_size
        cbi = cb[l0 + b:l0 + b + s + 1]  # s_cb(l0-1 .. l0+s-1)
        new_grids.append(cbi)
    sub.m, sub.n, sub.p = new_ext
    sub.bc_beg, sub.bc_end = tuple(bcb), tuple(bce)
    new_grids = [populate_grid_buffers(new_grids[d], sub.buff_size, bcb[d], bce[d]) for d in range(D)]
    qi = prm.interior(q)
    sl = [slice(None)] + [slice(None)] * 3
    idx = [2, 1, 0]  # array axes of x, y, z
    for d in range(D):
        sl[1 + idx[d]] = slice(lo[d], lo[d] + shape[d])
    return sub, new_grids, ghosted(sub, qi[tuple(sl)])
