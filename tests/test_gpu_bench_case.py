"""The benchmark workload itself: the 512^3 two-fluid 3D shock-droplet.

* Windows of the 512^3 grid through the Mach 2.4 air shock (its cell
  boundaries and initial state, cut faces extrapolated; one in the middle of
  the shock plane, one at a corner where it meets the domain's boundaries):
  one RHS and one RK3 step against the oracle, bitwise in the plain FP64
  form; in the contracted form (the benchmark's default) the step is within
  1e-12 per variable on the state's scales (helpers.state_linf) and the RHS
  of the mass, x-momentum and energy within 1e-12 of its largest value.  At t = 0 the droplet sits
  in air at rest and uniform pressure, where the RHS is zero up to rounding,
  so the windows are on the shock.
* The whole 512^3 grid, where the oracle would take minutes, through
  properties of the case, with the shock moved into the droplet (x = 0.42)
  and a spherical pressure pulse on the droplet, so the RHS of every
  variable is large wherever it is not zero by symmetry: the state is
  mirror-symmetric in y and z and symmetric under exchanging y and z, so the
  RHS must be too, with the normal momentum negated under a mirror and the
  transverse momenta exchanged under the swap, to rounding (the sweeps add
  the x, y, z contributions in that order and reconstruct the two faces of a
  cell with differently ordered formulas).
"""
import numpy as np
import pytest

from microfc_amd import PREC_FP64, PREC_FP64_FAST, RhsModule, ic
from oracle import oracle as O

from helpers import rel_linf, say, state_linf, to_oracle, ulp_diff

pytestmark = pytest.mark.gpu

N = 512
WINDOWS = [("shock_middle", (70, 224, 224), (64, 48, 40)),
           ("shock_corner", (80, 0, 0), (48, 40, 44))]


@pytest.mark.parametrize("win", WINDOWS, ids=[w[0] for w in WINDOWS])
@pytest.mark.parametrize("prec", [PREC_FP64, PREC_FP64_FAST], ids=["fp64", "fp64_fast"])
def test_bench_grid_window(gpu, win, prec):
    torch = gpu
    name, lo, shape = win
    prm, grids, q = ic.shock_droplet_3d(N=N, window=(lo, shape))
    say(f"[512^3 window] {name} cells {lo} + {shape} form {prec}")
    oc = to_oracle(prm)
    ref = O.compute_rhs(oc, grids, q.copy())
    prm.precision = prec
    mod = RhsModule(prm, grids)
    try:
        qd = torch.from_numpy(q.copy()).cuda()
        rd = mod.new_field(ghosted=False)
        mod.s_compute_rhs(qd, None, rd, 0)
        mod.sync()
        got = rd.cpu().numpy()
        err = rel_linf(got, ref)
        say(f"  rhs rel L-inf {np.array2string(err, precision=2)} max_ulp={ulp_diff(got, ref)}")
        dt = 4.8e-7  # the benchmark's CFL 0.4 step at 512^3
        qo = q.copy()
        O.tvd_rk_step(oc, grids, dt, qo, np.zeros_like(qo), t_step=0)
        q2 = mod.new_field()
        mod.s_tvd_rk(0, (qd, q2), None, rd, dt)
        mod.sync()
        sg, sr = prm.interior(qd.cpu().numpy()), prm.interior(qo)
        serr = state_linf(prm, sg, sr, q)
        say(f"  one RK3 step: rel L-inf {np.array2string(serr, precision=2)} max_ulp={ulp_diff(sg, sr)}")
        if prec == PREC_FP64:
            assert np.array_equal(got.view(np.int64), ref.view(np.int64))
            assert np.array_equal(sg.view(np.int64), sr.view(np.int64))
        else:
            assert serr.max() <= 1e-12
            # the planar shock moves mass, x momentum and energy; the
            # transverse momenta and the volume fractions have a zero RHS up
            # to rounding here (measured in the step on the state's scales)
            assert err[[0, 1, 2, 5]].max() <= 1e-12
    finally:
        mod.finalize()


@pytest.mark.parametrize("prec", [PREC_FP64, PREC_FP64_FAST], ids=["fp64", "fp64_fast"])
def test_bench_grid_full_symmetries(gpu, prec):
    torch = gpu
    prm, grids, q = ic.shock_droplet_3d(N=N, precision=prec, shock_x=0.42)
    say(f"[512^3 symmetries] form {prec}")
    # the shock's pressure depends on x alone and no fluid moves in y or z, so
    # a spherical pressure pulse on the droplet makes the transverse momenta
    # move: E += Gamma_mix * dp, symmetric like the rest of the state
    qi = prm.interior(q)
    X, Y, Z = ic.mesh(prm, grids)
    dp = 2e5 * np.exp(-((X - 0.5) ** 2 + (Y - 0.5) ** 2 + (Z - 0.5) ** 2) / 0.004)
    qi[5] += (qi[6] * prm.gammas[0] + qi[7] * prm.gammas[1]) * dp
    del X, Y, Z, dp, qi
    rho_max = float(np.max(q[0] + q[1]))
    mod = RhsModule(prm, grids)
    try:
        qd = torch.from_numpy(q).cuda()
        del q
        rd = mod.new_field(ghosted=False)
        mod.s_compute_rhs(qd, None, rd, 0)
        mod.sync()
        del qd
        r = rd  # (S, z, y, x) on the device; compare there, move scalars only
        scale = torch.amax(torch.abs(r).reshape(r.shape[0], -1), dim=1)
        assert torch.isfinite(scale).all().item()
        # volume fractions on the velocity-divergence scale (their rhs is a
        # near-cancellation away from moving interfaces)
        div = float(torch.max(scale[0], scale[1])) / rho_max
        scale[6:] = torch.clamp(scale[6:], min=div)

        def worst(a, b):
            d = torch.amax(torch.abs(a - b).reshape(a.shape[0], -1), dim=1)
            return (d / scale).cpu().numpy()

        # mirror in z: rhs(x, y, N-1-z) = rhs(x, y, z), rho*w negated
        m = torch.flip(r, dims=[1])
        m[4] = -m[4]
        ez = worst(m, r)
        # mirror in y, rho*v negated
        m = torch.flip(r, dims=[2])
        m[3] = -m[3]
        ey = worst(m, r)
        del m
        # exchange y and z, with rho*v and rho*w exchanged
        t = r.transpose(1, 2)[[0, 1, 2, 4, 3, 5, 6, 7]]
        es = worst(t, r)
        say(f"  mirror z {np.array2string(ez, precision=2)}")
        say(f"  mirror y {np.array2string(ey, precision=2)}")
        say(f"  y<->z    {np.array2string(es, precision=2)}")
        for e in (ez, ey, es):
            assert e.max() <= 1e-12, e
    finally:
        mod.finalize()
