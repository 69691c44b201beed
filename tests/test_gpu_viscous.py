"""Viscous terms on the GPU against the CPU FP64 oracle, whose viscous
restatement tests/test_oracle_viscous.py and tests/test_oracle_weno_re_flux.py
pin: every case with finite-difference velocity gradients and again with
weno_Re_flux (gradients by the divergence theorem on WENO face values,
reconstructed to the faces).  Bar as for the inviscid path: relative
L-infinity <= 1e-12 per RHS and per step (bitwise equality is printed)."""
import copy

import numpy as np
import pytest

from microfc_amd import RhsModule, ic
from microfc_amd.rhs import MfcError, PREC_FP32_BULK
from oracle import oracle as O

from helpers import assert_parity, multi_fluid_2d, perturbed, relayout, reghost, say, stretched_grids, to_oracle

pytestmark = pytest.mark.gpu


def with_velocity(prm, grids, q, amp):
    """Adds a smooth velocity field of amplitude amp to every component at
    unchanged density and pressure (the kinetic energy goes into E), so the
    viscous terms act everywhere."""
    q = q.copy()
    nf, D = prm.num_fluids, prm.num_dims
    X, Y, _ = ic.mesh(prm, grids)
    qi = prm.interior(q)
    rho = sum(qi[i] for i in range(nf))
    ke0 = sum(0.5 * qi[nf + d] ** 2 for d in range(D)) / rho
    for d in range(D):
        phase = 2 * np.pi * (X + (0.0 if Y is None else (d + 1) * Y))
        qi[nf + d] = qi[nf + d] + rho * amp * np.sin(phase + d)
    qi[nf + D] = qi[nf + D] + sum(0.5 * qi[nf + d] ** 2 for d in range(D)) / rho - ke0
    return q


def _visc(prm, g, q, Re):
    """The case with Reynolds numbers Re, grids and field laid out with the
    viscous ghost width."""
    prm.Re = Re
    return (prm, *relayout(prm, g, q))


def _cases():
    out = []
    prm, g, q = _visc(*ic.sod_1d(m=199), [(1e3, 2e3)])
    out.append(("visc_sod1d_w5_rk3", prm, g, with_velocity(prm, g, perturbed(prm, q), 0.2)))
    prm, g, q = ic.sod_1d(m=63, weno_order=1)
    prm.time_stepper = 1
    prm.bc_beg = (-2, -3, -3)
    prm.bc_end = (-4, -3, -3)
    prm, g, q = _visc(prm, ic.uniform_grids(prm), q, [(50.0, -1.0)])
    out.append(("visc_sod1d_w1_rk1_reflect_extrap", prm, g, with_velocity(prm, g, perturbed(prm, q), 0.2)))
    prm, g, q = _visc(*ic.shock_bubble_2d(N=48), [(0.05, 0.1), (0.1, 0.2)])
    out.append(("visc_bubble2d_48", prm, g, with_velocity(prm, g, perturbed(prm, q), 20.0)))
    prm, g, q = ic.shock_bubble_2d(N=40, weno_order=3)
    prm.time_stepper = 2
    prm.bc_beg = (-1, -2, -3)
    prm.bc_end = (-1, -4, -3)
    prm, g, q = _visc(prm, ic.uniform_grids(prm), q, [(0.05, -1.0), (0.1, -1.0)])
    out.append(("visc_bubble2d_w3_rk2_periodic_reflect", prm, g, with_velocity(prm, g, perturbed(prm, q), 20.0)))
    prm, g, q = _visc(*ic.shock_bubble_2d(N=40), [(-1.0, 0.05), (-1.0, 0.1)])
    g = stretched_grids(prm)
    out.append(("visc_bubble2d_stretched_bulk_only", prm, g, with_velocity(prm, g, perturbed(prm, q), 20.0)))
    prm, g, q = _visc(*ic.shock_bubble_2d(N=40), [(0.05, 0.1), (0.1, -1.0)])
    g = stretched_grids(prm)
    out.append(("visc_bubble2d_stretched_shear_bulk", prm, g, with_velocity(prm, g, perturbed(prm, q), 20.0)))
    prm, g, q = _visc(*multi_fluid_2d(3, N=32), [(0.3, -1.0), (-1.0, 0.7), (0.9, 1.1)])
    out.append(("visc_fluids3_2d", prm, g, q))
    prm, g, q = _visc(*multi_fluid_2d(4, N=32), [(0.3, 0.5), (0.4, -1.0), (-1.0, 0.8), (0.9, 1.1)])
    out.append(("visc_fluids4_2d", prm, g, q))
    prm, g, q = ic.shock_bubble_2d(N=36, weno_order=1)
    prm.time_stepper = 1
    prm.bc_beg = (-4, -1, -3)
    prm.bc_end = (-2, -1, -3)
    prm, g, q = _visc(prm, ic.uniform_grids(prm), q, [(0.05, 0.1), (0.1, 0.2)])
    out.append(("visc_bubble2d_w1_rk1_extrap_periodic", prm, g, with_velocity(prm, g, perturbed(prm, q), 20.0)))
    prm, g, q = _visc(*ic.shear_layer_2d(N=48), [(200.0, -1.0)])
    out.append(("visc_shear2d_48", prm, g, perturbed(prm, q)))
    return out


CASES = _cases()
IDS = [c[0] for c in CASES]


def _wrf(case):
    name, prm, g, q = case
    prm = copy.deepcopy(prm)
    prm.weno_Re_flux = True
    return (name + "_weno_re_flux", prm, g, q)


ALL = CASES + [_wrf(c) for c in CASES]
ALL_IDS = [c[0] for c in ALL]


@pytest.mark.parametrize("case", ALL, ids=ALL_IDS)
def test_viscous_rhs_parity(gpu, case):
    torch = gpu
    name, prm, grids, q = case
    say(f"[viscous rhs] {name} shape={prm.ghost_shape} Re={prm.Re} weno_Re_flux={prm.weno_Re_flux}")
    oc = to_oracle(prm)
    ref = O.compute_rhs(oc, grids, q.copy())
    oi = O.Case(**{**oc.__dict__, "Re": None})
    gi, qi = reghost(oc, oi, grids, q)
    inv = O.compute_rhs(oi, gi, qi)
    mod = RhsModule(prm, grids)
    try:
        qd = torch.from_numpy(q.copy()).cuda()
        rd = mod.new_field(ghosted=False)
        mod.s_compute_rhs(qd, None, rd, 0)
        mod.sync()
        got = rd.cpu().numpy()
        assert_parity(got, ref, what=f"{name} rhs")
        # the viscous terms are not a rounding-level change here
        nf, D = prm.num_fluids, prm.num_dims
        moved = np.max(np.abs(ref[nf:nf + D + 1] - inv[nf:nf + D + 1])) / np.max(np.abs(inv[nf:nf + D + 1]))
        say(f"  viscous share of the momentum/energy rhs: {moved:.2e}")
        if prm.weno_Re_flux and prm.weno_order == 1:
            assert moved == 0.0  # WENO1 face values are the cell value: every gradient is 0
        else:
            assert moved > 1e-8
    finally:
        mod.finalize()


@pytest.mark.parametrize("case", ALL, ids=ALL_IDS)
def test_viscous_tvd_rk_parity(gpu, case):
    torch = gpu
    name, prm, grids, q = case
    nsteps = 3
    say(f"[viscous tvd_rk] {name} stepper={prm.time_stepper} steps={nsteps}")
    oc = to_oracle(prm)
    mod = RhsModule(prm, grids)
    try:
        qd = torch.from_numpy(q.copy()).cuda()
        dt = 0.2 * mod.cfl_dt(qd, 1.0)
        qo = q.copy()
        q2o = np.zeros_like(q)
        q2d = mod.new_field()
        rd = mod.new_field(ghosted=False)
        for t in range(nsteps):
            O.tvd_rk_step(oc, grids, dt, qo, q2o, t_step=t)
            mod.s_tvd_rk(t, (qd, q2d), None, rd, dt)
            mod.sync()
            assert_parity(prm.interior(qd.cpu().numpy()), prm.interior(qo), what=f"{name} step {t}")
    finally:
        mod.finalize()


def test_viscous_rejections(gpu):
    """Viscous terms are 1D/2D and FP64, as the library documents."""
    prm, g, _ = ic.shock_droplet_3d(N=28)
    prm.Re = [(1.0, -1.0), (1.0, -1.0)]
    with pytest.raises(MfcError, match="1D and 2D"):
        RhsModule(prm, g)
    prm, g, _ = ic.shock_bubble_2d(N=32)
    prm.Re = [(1.0, -1.0), (1.0, -1.0)]
    prm.precision = PREC_FP32_BULK
    with pytest.raises(MfcError, match="FP64"):
        RhsModule(prm, g)


@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_stability_criteria_parity(gpu, case):
    """ICFL max, VCFL max and Rc min of the run-time information
    (m_data_output.fpp:197-263) equal the oracle's bit for bit."""
    torch = gpu
    name, prm, grids, q = case
    say(f"[stability criteria] {name}")
    oc = to_oracle(prm)
    prim = O.cons_to_prim(oc, O.bc_fill(oc, q.copy()))
    mod = RhsModule(prm, grids)
    try:
        pd = torch.from_numpy(prim).cuda()
        for dt in (1e-7, 2e-4):
            got = mod.stability_criteria(pd, dt)
            ref = O.stability(oc, grids, prim, dt)
            say(f"  dt={dt}: icfl {got[0]:.6e} vcfl {got[1]:.6e} Rc {got[2]:.6e}")
            assert got == ref, (got, ref)
            assert got[1] > 0 and got[2] > 0
    finally:
        mod.finalize()
