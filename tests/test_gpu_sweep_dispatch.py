"""Which sweep kernels a case runs, read from the library's per-kernel HIP-event
timing (mfc_timing_read): the fused x and y sweep on grids uniform in x and y
with WENO5 and one or two fluids, in the plain, contracted and FP32-bulk
forms; the separate sweeps on stretched grids, with three fluids and in the
learned form; on grids uniform up to the rounding of their cell boundaries
the fused sweep in the contracted and FP32-bulk forms and the separate
sweeps (per-cell coefficients) in the plain form.  The results of every path are checked against the oracle by
the parity, contracted-form and FP32-bulk tests; this pins that the fused
path is the one those tests exercise on the uniform cases."""
import numpy as np
import pytest

from microfc_amd import PREC_FP32_BULK, PREC_FP64, PREC_FP64_FAST, RhsModule, ic
from microfc_amd.rhs import populate_grid_buffers, uniform_cell_boundaries
from oracle import oracle as O

from helpers import field_3d_case, multi_fluid_2d, say, stretched_grids

pytestmark = pytest.mark.gpu


def _launches(torch, prm, grids, q, precision):
    prm.precision = precision
    mod = RhsModule(prm, grids)
    try:
        qd = torch.from_numpy(q.copy()).cuda()
        rd = mod.new_field(ghosted=False)
        mod.timing(enable=True, reset=True)
        mod.s_compute_rhs(qd, None, rd, 0)
        mod.sync()
        t = mod.timing(enable=False)
        return {k: v[1] for k, v in t.items() if v[1]}
    finally:
        prm.precision = 0
        mod.finalize()


@pytest.mark.parametrize("precision", [PREC_FP64, PREC_FP64_FAST, PREC_FP32_BULK], ids=["fp64", "fast", "fp32"])
def test_uniform_grid_runs_the_fused_sweep(gpu, precision):
    prm, g, q = field_3d_case(31, 31, 31, (-2, -4, -1), (-4, -2, -1))
    n = _launches(gpu, prm, g, q, precision)
    say(f"[dispatch] uniform 3D, precision {precision}: {n}")
    assert n.get("sweep_xy") == 1 and n.get("sweep_z") == 1
    assert "sweep_x" not in n and "sweep_y" not in n


@pytest.mark.parametrize("precision", [PREC_FP64, PREC_FP64_FAST], ids=["fp64", "fast"])
def test_stretched_grid_and_three_fluids_run_separate_sweeps(gpu, precision):
    prm, g, q = field_3d_case(31, 31, 31, (-2, -4, -1), (-4, -2, -1))
    n = _launches(gpu, prm, stretched_grids(prm), q, precision)
    say(f"[dispatch] stretched 3D, precision {precision}: {n}")
    assert "sweep_xy" not in n and n.get("sweep_x") == 1 and n.get("sweep_y") == 1
    prm, g, q = multi_fluid_2d(3, N=64)
    n = _launches(gpu, prm, g, q, precision)
    say(f"[dispatch] three fluids 2D, precision {precision}: {n}")
    assert "sweep_xy" not in n and n.get("sweep_x") == 1 and n.get("sweep_y") == 1


@pytest.mark.parametrize("precision", [PREC_FP64, PREC_FP64_FAST, PREC_FP32_BULK], ids=["fp64", "fast", "fp32"])
def test_rounded_uniform_grid(gpu, precision):
    # 47 x 39 x 29 cells on [0, 1]^3: equal widths up to the rounding of the
    # boundaries (WENO coefficients up to 2.5e-14 apart)
    prm, g, q = field_3d_case(47, 39, 29, (-1, -1, -1), (-1, -1, -1))
    n = _launches(gpu, prm, g, q, precision)
    say(f"[dispatch] rounded uniform 3D, precision {precision}: {n}")
    if precision == PREC_FP64:
        assert "sweep_xy" not in n and n.get("sweep_x") == 1 and n.get("sweep_y") == 1
    else:
        assert n.get("sweep_xy") == 1 and n.get("sweep_z") == 1
    prm, g, q = ic.shock_bubble_2d(N=96)
    n = _launches(gpu, prm, g, q, precision)
    say(f"[dispatch] rounded uniform 2D, precision {precision}: {n}")
    assert ("sweep_xy" in n) == (precision != PREC_FP64)


@pytest.mark.parametrize("eps,fused", [(1e-13, True), (4e-12, False)], ids=["6e-14", "1.8e-12"])
def test_uniform_tolerance_of_the_contracted_form(gpu, eps, fused):
    """x cell boundaries of the 64^2 shock-bubble moved by eps * dx * sin: WENO
    coefficients 6e-14 apart (relative) are taken as uniform by the contracted
    form, 1.8e-12 apart are not (the bound is 5e-13)."""
    prm, g, q = ic.shock_bubble_2d(N=64)
    b, s = prm.buff_size, prm.m
    cbi = uniform_cell_boundaries(s, 0.0, 1.0).copy()
    k = np.arange(1, s + 1)
    cbi[1:s + 1] += eps / (s + 1) * np.sin(6 * np.pi * k / (s + 1))
    gx = populate_grid_buffers(cbi, b, prm.bc_beg[0], prm.bc_end[0])
    c = O.weno_coefficients(gx[0], s, b, 5)[:, b - 1:b + s + 2]
    dev = float(np.max(np.abs(c - c[:, :1]) / np.maximum(np.abs(c[:, :1]), 1e-300)))
    assert (dev <= 5e-13) == fused, dev
    n = _launches(gpu, prm, [gx, g[1]], q, PREC_FP64_FAST)
    say(f"[dispatch] x boundaries moved by {eps:.0e} dx (coefficients {dev:.1e} apart): {n}")
    assert ("sweep_xy" in n) == fused
    if not fused:
        assert n.get("sweep_x") == 1 and n.get("sweep_y") == 1
