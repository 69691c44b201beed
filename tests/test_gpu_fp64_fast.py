"""Contracted FP64 form (MFC_PREC_FP64_FAST: FMA contraction, divisions as a
multiply by a Newton-refined reciprocal, WENO weights normalised by one
reciprocal of their sum) against the FP64 oracle on the whole parity matrix.

Bars, with the error printed per case:
  * one RK step from identical inputs (the oracle's state before the step):
    relative L-infinity of the new state <= 1e-12 per variable, the
    reference's FP64 requirement, checked on three consecutive steps;
  * one RHS: relative L-infinity <= 1e-12 per variable, except the volume
    fractions, whose reference RHS is a near-cancellation of the flux
    difference and the alpha*div(u) source (zero up to rounding in
    single-fluid regions and away from interfaces): they are measured
    against the larger of their own largest RHS and the largest
    partial-density RHS (helpers.scaled_linf), <= 1e-12;
  * 40 steps of the 2D shock-bubble and the 3D shock-droplet: the two
    trajectories stay within 1e-9 relative L-infinity (per-step rounding
    differences that the shocks and interfaces amplify over time);
  * discrete conservation of mass, momentum and energy with periodic
    boundaries to FP64 round-off.
The form is not bit-equal to the oracle."""
import numpy as np
import pytest

from microfc_amd import PREC_FP64_FAST, RhsModule, ic
from oracle import oracle as O

from helpers import perturbed, rel_linf, say, scaled_linf, small_cases, to_oracle, ulp_diff

pytestmark = pytest.mark.gpu

TOL = 1e-12
CASES = small_cases()
IDS = [c[0] for c in CASES]


def _check(got, ref, what, tol=TOL):
    err = rel_linf(got, ref)
    say(f"  {what}: rel L-inf per variable {np.array2string(err, precision=2)} max_ulp={ulp_diff(got, ref)}")
    assert np.all(np.isfinite(got))
    assert err.max() <= tol, f"{what}: {err}"


@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_fp64_fast_rhs(gpu, case):
    torch = gpu
    name, prm, grids, q = case
    say(f"[fp64 fast rhs] {name}")
    ref = O.compute_rhs(to_oracle(prm), grids, q)
    prm.precision = PREC_FP64_FAST
    mod = RhsModule(prm, grids)
    try:
        qd = torch.from_numpy(q.copy()).cuda()
        rd = mod.new_field(ghosted=False)
        mod.s_compute_rhs(qd, None, rd, 0)
        mod.sync()
        got = rd.cpu().numpy()
        rel = rel_linf(got, ref)
        err = scaled_linf(prm, got, ref)
        adv = prm.num_fluids + prm.num_dims + 1
        err[:adv] = rel[:adv]
        say(f"  {name} rhs: rel L-inf per variable (volume fractions on the partial-density rhs scale) "
            f"{np.array2string(err, precision=2)}; plain rel {np.array2string(rel, precision=2)}")
        assert np.all(np.isfinite(got)) and err.max() <= TOL, err
    finally:
        prm.precision = 0
        mod.finalize()


@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_fp64_fast_steps(gpu, case):
    torch = gpu
    name, prm, grids, q = case
    nsteps = 3
    oc = to_oracle(prm)
    prm.precision = PREC_FP64_FAST
    mod = RhsModule(prm, grids)
    try:
        qd = torch.from_numpy(q.copy()).cuda()
        dt = 0.3 * mod.cfl_dt(qd, 1.0)
        say(f"[fp64 fast steps] {name} stepper={prm.time_stepper} dt={dt:.3e}")
        q2d = mod.new_field()
        rd = mod.new_field(ghosted=False)
        qo = q.copy()
        q2o = np.zeros_like(qo)
        for t in range(nsteps):
            # each step starts from the oracle's state: identical inputs
            qd.copy_(torch.from_numpy(qo))
            O.tvd_rk_step(oc, grids, dt, qo, q2o, t_step=t)
            mod.s_tvd_rk(t, (qd, q2d), None, rd, dt)
            mod.sync()
            _check(prm.interior(qd.cpu().numpy()), prm.interior(qo), f"{name} step {t}")
    finally:
        prm.precision = 0
        mod.finalize()


def test_fp64_fast_rounded_grid_768(gpu):
    """The 2D shock-bubble at 768^2 on [0, 1]^2 (every cell perturbed): its
    cell widths differ by the rounding of the boundaries, its WENO
    coefficients by up to 2.6e-13 relative, and the contracted form takes the
    fused sweep with the first cell's coefficients.  One RHS and one RK3 step
    from the oracle's state within 1e-12 of the oracle's (the RHS's volume
    fractions on the partial-density scale, as above), and 20 further steps
    of both within the 1e-9 trajectory bar."""
    torch = gpu
    prm, grids, q = ic.shock_bubble_2d(N=768)
    q = perturbed(prm, q)
    oc = to_oracle(prm)
    ref = O.compute_rhs(oc, grids, q)
    prm.precision = PREC_FP64_FAST
    mod = RhsModule(prm, grids)
    try:
        qd = torch.from_numpy(q.copy()).cuda()
        rd = mod.new_field(ghosted=False)
        mod.timing(enable=True, reset=True)
        mod.s_compute_rhs(qd, None, rd, 0)
        mod.sync()
        calls = {k: v[1] for k, v in mod.timing(enable=False).items() if v[1]}
        say(f"[fp64 fast rounded grid] bubble2d_768 launches {calls}")
        assert calls.get("sweep_xy") == 1
        got = rd.cpu().numpy()
        rel = rel_linf(got, ref)
        err = scaled_linf(prm, got, ref)
        adv = prm.num_fluids + prm.num_dims + 1
        err[:adv] = rel[:adv]
        say(f"  rhs: {np.array2string(err, precision=2)}")
        assert np.all(np.isfinite(got)) and err.max() <= TOL, err
        dt = 0.3 * mod.cfl_dt(qd, 1.0)
        q2d = mod.new_field()
        qo = q.copy()
        q2o = np.zeros_like(qo)
        O.tvd_rk_step(oc, grids, dt, qo, q2o, t_step=0)
        mod.s_tvd_rk(0, (qd, q2d), None, rd, dt)
        mod.sync()
        _check(prm.interior(qd.cpu().numpy()), prm.interior(qo), "bubble2d_768 step")
        for t in range(1, 21):
            O.tvd_rk_step(oc, grids, dt, qo, q2o, t_step=t)
            mod.s_tvd_rk(t, (qd, q2d), None, rd, dt)
        mod.sync()
        _check(prm.interior(qd.cpu().numpy()), prm.interior(qo), "bubble2d_768 after 21 steps", tol=1e-9)
    finally:
        prm.precision = 0
        mod.finalize()


def test_fp64_fast_conserves_to_roundoff(gpu):
    torch = gpu
    prm, grids, q = ic.shear_layer_2d(N=64)
    prm.bc_beg = (-1, -1, -1)
    prm.bc_end = (-1, -1, -1)
    grids = ic.uniform_grids(prm)
    prm.precision = PREC_FP64_FAST
    say("[fp64 fast conservation] shear2d_64 periodic")
    mod = RhsModule(prm, grids)
    try:
        qd = torch.from_numpy(q.copy()).cuda()
        rd = mod.new_field(ghosted=False)
        mod.s_compute_rhs(qd, None, rd, 0)
        mod.sync()
        r = rd.cpu().numpy()
        for v in range(4):
            tot, mag = abs(np.sum(r[v])), np.sum(np.abs(r[v]))
            say(f"  var {v}: |sum rhs| / sum|rhs| = {tot / mag:.2e}")
            assert tot <= 1e-13 * mag
    finally:
        mod.finalize()


def test_fp64_fast_long_run_on_bench_shapes(gpu):
    """40 RK3 steps of the 2D shock-bubble and the 3D shock-droplet (the
    benchmark workloads at reduced size) from the same initial state: the
    two trajectories stay within 1e-9 relative L-infinity."""
    torch = gpu
    for name, (prm, grids, q) in [("bubble2d_96", ic.shock_bubble_2d(N=96)),
                                  ("droplet3d_32", ic.shock_droplet_3d(N=32))]:
        oc = to_oracle(prm)
        prm.precision = PREC_FP64_FAST
        mod = RhsModule(prm, grids)
        try:
            qd = torch.from_numpy(q.copy()).cuda()
            dt = mod.cfl_dt(qd, 0.4)
            say(f"[fp64 fast 40 steps] {name} dt={dt:.3e}")
            q2d = mod.new_field()
            rd = mod.new_field(ghosted=False)
            qo = q.copy()
            q2o = np.zeros_like(qo)
            for t in range(40):
                O.tvd_rk_step(oc, grids, dt, qo, q2o, t_step=t)
                mod.s_tvd_rk(t, (qd, q2d), None, rd, dt)
            mod.sync()
            _check(prm.interior(qd.cpu().numpy()), prm.interior(qo), f"{name} after 40 steps", tol=1e-9)
        finally:
            prm.precision = 0
            mod.finalize()
