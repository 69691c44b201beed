"""The pre-processor (microfc_amd/pre_process.py) pinned by what the
geometry and the formulas fix, independent of its own code: brute-force
cell counts and index ranges of each patch shape on grids whose cell
centres are exact binary fractions, patch precedence through alter_patch,
the smoothing blend at single cells, the stretched grid's symmetry and
monotonicity, the conservative conversion's closed form, and round trips
through the serial and parallel file formats.  The benchmark generators of
microfc_amd/ic.py for the reference's Sod and 2D_viscous examples are the
same fields, bit for bit."""
import math
from fractions import Fraction

import numpy as np
import pytest

from microfc_amd import case as C
from microfc_amd import ic
from microfc_amd.pre_process import DFLT_REAL, PreProcess

G_AIR = 1.0 / (1.4 - 1.0)


def _base2d(m=63, n=31, patches=(), fluids=1, **extra):
    c = {"m": m, "n": n, "x_domain%beg": 0.0, "x_domain%end": 1.0, "y_domain%beg": 0.0, "y_domain%end": 0.5,
         "num_fluids": fluids, "num_patches": len(patches)}
    for f in range(fluids):
        c[f"fluid_pp({f + 1})%gamma"] = G_AIR
        c[f"fluid_pp({f + 1})%pi_inf"] = 0.0
    for k, p in enumerate(patches, start=1):
        for key, v in p.items():
            c[f"patch_icpp({k})%{key}"] = v
    c.update(extra)
    return c


def _state(rho, u, v, p, alpha=1.0):
    return {"alpha_rho(1)": rho, "alpha(1)": alpha, "vel(1)": u, "vel(2)": v, "pres": p}


BG = dict(geometry=3, x_centroid=0.5, y_centroid=0.25, length_x=1.0, length_y=0.5, **_state(1.0, 0.0, 0.0, 1.0))


def _centres(s, lo, hi):
    h = Fraction(hi - lo) / (s + 1)
    return [Fraction(lo) + h * Fraction(2 * i + 1, 2) for i in range(s + 1)]


def test_uniform_grid_is_the_simulations():
    pp = PreProcess(_base2d(patches=[BG]))
    from microfc_amd.rhs import uniform_cell_boundaries
    assert np.array_equal(pp.x_cb, uniform_cell_boundaries(63, 0.0, 1.0))
    assert np.array_equal(pp.y_cb, uniform_cell_boundaries(31, 0.0, 0.5))
    assert [Fraction(x) for x in pp.x_cc] == _centres(63, 0, 1)


def test_rectangle_covers_exactly_its_index_range():
    rect = dict(geometry=3, x_centroid=0.25, y_centroid=0.25, length_x=0.25, length_y=0.125,
                alter_patch=None, **_state(2.0, 1.0, -1.0, 3.0))
    rect.pop("alter_patch")
    rect["alter_patch(1)"] = "T"
    pp = PreProcess(_base2d(patches=[BG, rect]))
    prim, _ = pp.initial_condition()
    xs, ys = _centres(63, 0, 1), _centres(31, 0, Fraction(1, 2))
    want = np.array([[Fraction(1, 8) <= x <= Fraction(3, 8) and Fraction(3, 16) <= y <= Fraction(5, 16)
                      for x in xs] for y in ys])
    assert np.array_equal(pp.patch_id_fp == 2, want) and want.sum() == 16 * 8
    assert np.all(prim[0][want] == 2.0) and np.all(prim[0][~want] == 1.0)
    assert np.all(prim[1][want] == 1.0) and np.all(prim[2][want] == -1.0) and np.all(prim[3][want] == 3.0)


def test_circle_cells_by_brute_force_and_alter_patch():
    circ = dict(geometry=2, x_centroid=0.5, y_centroid=0.25, radius=0.15, **_state(5.0, 0.0, 0.0, 1.0))
    xs, ys = _centres(63, 0, 1), _centres(31, 0, Fraction(1, 2))
    R = Fraction(0.15)
    inside = np.array([[(x - Fraction(1, 2)) ** 2 + (y - Fraction(1, 4)) ** 2 <= R * R for x in xs] for y in ys])
    # may not overwrite patch 1 (alter_patch(1) defaults to F): nothing changes
    pp = PreProcess(_base2d(patches=[BG, circ]))
    prim, _ = pp.initial_condition()
    assert np.all(pp.patch_id_fp == 1) and np.all(prim[0] == 1.0)
    # with permission it takes exactly the cells whose centre is inside
    circ["alter_patch(1)"] = "T"
    pp = PreProcess(_base2d(patches=[BG, circ]))
    prim, _ = pp.initial_condition()
    assert np.array_equal(pp.patch_id_fp == 2, inside)
    assert abs(inside.sum() * (1 / 64) * (0.5 / 32) - math.pi * 0.15 ** 2) < 0.02 * math.pi * 0.15 ** 2


def test_ellipse_sweep_line_varcircle_and_vortex_regions():
    xs, ys = _centres(63, 0, 1), _centres(31, 0, Fraction(1, 2))
    cases = [
        (dict(geometry=5, x_centroid=0.5, y_centroid=0.25, **{"radii(1)": 0.25, "radii(2)": 0.125}),
         lambda x, y: ((x - Fraction(1, 2)) / Fraction(1, 4)) ** 2 + ((y - Fraction(1, 4)) / Fraction(1, 8)) ** 2 <= 1),
        (dict(geometry=4, x_centroid=0.5, y_centroid=0.25, **{"normal(1)": 1.0, "normal(2)": 2.0}),
         lambda x, y: x + 2 * y - Fraction(1, 2) - Fraction(1, 2) >= 0),
        (dict(geometry=18, x_centroid=0.5, y_centroid=0.25, radius=0.125, epsilon=0.0625),
         lambda x, y: Fraction(3, 32) <= ((x - Fraction(1, 2)) ** 2 + (y - Fraction(1, 4)) ** 2) ** 0.5 <= Fraction(5, 32)),
        (dict(geometry=6, x_centroid=0.5, y_centroid=0.25, radius=0.125),
         lambda x, y: (x - Fraction(1, 2)) ** 2 + (y - Fraction(1, 4)) ** 2 <= Fraction(1, 64)),
    ]
    for geo, inside in cases:
        p = dict(geo, **_state(3.0, 0.0, 0.0, 1.0))
        p["alter_patch(1)"] = "T"
        pp = PreProcess(_base2d(patches=[BG, p]))
        pp.initial_condition()
        want = np.array([[bool(inside(x, y)) for x in xs] for y in ys])
        assert want.sum() > 0 and np.array_equal(pp.patch_id_fp == 2, want), geo["geometry"]


def test_smoothed_circle_blends_by_the_tanh_pseudo_volume_fraction():
    circ = dict(geometry=2, x_centroid=0.5, y_centroid=0.25, radius=0.1, smoothen="T", smooth_patch_id=1,
                smooth_coeff=0.5, **_state(1000.0, 0.0, 0.0, 1.0, alpha=1.0))
    bg = dict(BG, **{"alpha_rho(1)": 1.0})
    pp = PreProcess(_base2d(patches=[bg, circ]))
    prim, _ = pp.initial_condition()
    # every cell of the smoothing patch takes the blend; a cell changes hands
    # only where 1 - eta < 1e-16, which this gentle smoothing never reaches
    dx = min(1 / 64, 0.5 / 32)
    for j, i in [(16, 32), (16, 40), (20, 38), (3, 3)]:
        x, y = pp.x_cc[i], pp.y_cc[j]
        eta = math.tanh(0.5 / dx * (math.sqrt((x - 0.5) ** 2 + (y - 0.25) ** 2) - 0.1)) * -0.5 + 0.5
        assert prim[0, j, i] == pytest.approx(eta * 1000.0 + (1 - eta) * 1.0, rel=1e-13)
    assert np.all(pp.patch_id_fp == 1)
    # a sharp smoothing (eta = 1 to rounding inside) hands the centre over
    circ["smooth_coeff"] = 50.0
    pp = PreProcess(_base2d(patches=[bg, circ]))
    pp.initial_condition()
    assert pp.patch_id_fp[16, 32] == 2 and pp.patch_id_fp[3, 3] == 1


def test_unassigned_cells_keep_dflt_real():
    rect = dict(geometry=3, x_centroid=0.25, y_centroid=0.25, length_x=0.5, length_y=0.5, **_state(1, 0, 0, 1))
    pp = PreProcess(_base2d(patches=[rect]))
    prim, _ = pp.initial_condition()
    assert np.all(prim[:, :, 32:] == DFLT_REAL) and np.all(pp.patch_id_fp[:, 32:] == 0)


def test_conservative_conversion_closed_form():
    g2, pi2 = 1 / 3.4, 4.4 * 6e8 / 3.4
    p = {"alpha_rho(1)": 0.3, "alpha_rho(2)": 700.0, "alpha(1)": 0.3, "alpha(2)": 0.7, "vel(1)": 3.0,
         "vel(2)": -4.0, "pres": 2e5, "geometry": 3, "x_centroid": 0.5, "y_centroid": 0.25, "length_x": 1.0,
         "length_y": 0.5}
    c = _base2d(patches=[p], fluids=2)
    c["fluid_pp(2)%gamma"], c["fluid_pp(2)%pi_inf"] = g2, pi2
    _, cons = PreProcess(c).initial_condition()
    rho = 700.3
    assert np.allclose(cons[2], rho * 3.0, rtol=1e-15) and np.allclose(cons[3], -rho * 4.0, rtol=1e-15)
    E = (0.3 * G_AIR + 0.7 * g2) * 2e5 + 0.5 * rho * 25.0 + 0.7 * pi2
    assert np.allclose(cons[4], E, rtol=1e-15)


def test_1d_analytical_pressure_bump_uses_the_right_cell_boundary():
    c = {"m": 99, "n": 0, "x_domain%beg": 0.0, "x_domain%end": 1.0, "num_fluids": 1, "num_patches": 1,
         "fluid_pp(1)%gamma": G_AIR, "fluid_pp(1)%pi_inf": 0.0, "patch_icpp(1)%geometry": 15,
         "patch_icpp(1)%x_centroid": 0.5, "patch_icpp(1)%length_x": 1.0, "patch_icpp(1)%vel(1)": 0.0,
         "patch_icpp(1)%pres": 1.0, "patch_icpp(1)%alpha_rho(1)": 1.0, "patch_icpp(1)%alpha(1)": 1.0}
    prim, _ = PreProcess(c).initial_condition()
    xb = (np.arange(100) + 1) / 100.0
    assert np.allclose(prim[2, 0], 1.0 + 0.2 * np.exp(-((xb - 0.5) ** 2) / 0.01), rtol=1e-14)


def test_stretched_grid_is_symmetric_monotone_and_clusters_between_x_a_and_x_b():
    c = _base2d(m=127, patches=[BG], **{"x_domain%beg": -0.5, "x_domain%end": 0.5, "stretch_x": "T",
                                        "a_x": 4.0, "x_a": -0.2, "x_b": 0.2, "loops_x": 2})
    pp = PreProcess(c)
    cb, dx = pp.x_cb, np.diff(pp.x_cb)
    assert np.all(dx > 0)
    assert np.allclose(cb, -cb[::-1], atol=1e-15)          # odd symmetry about the centre
    assert cb[64] == 0.0                                    # the centre stays
    # the cells are finest at the centre, between x_a and x_b, and grow
    # monotonically from there to several times that at the ends
    half = dx[64:]
    assert np.all(np.diff(half) > 0) and dx[0] > 5 * dx.min() and pp.dx == dx.min()
    assert np.argmin(dx) in (63, 64)
    assert np.array_equal(pp.x_cc, (cb[1:] + cb[:-1]) / 2.0)


def test_reference_examples_equal_the_benchmark_generators():
    """examples/1D_sodshocktube (two line segments) and examples/2D_viscous
    (two rectangles, the second overwriting the first's upper half), laid out
    here, equal ic.sod_1d and ic.viscous_shear_2d bit for bit."""
    sod = {"m": 999, "n": 0, "x_domain%beg": 0.0, "x_domain%end": 1.0, "num_fluids": 1, "num_patches": 2,
           "fluid_pp(1)%gamma": 1.0 / (1.4 - 1.0), "fluid_pp(1)%pi_inf": 0.0,
           "patch_icpp(1)%geometry": 1, "patch_icpp(1)%x_centroid": 0.25, "patch_icpp(1)%length_x": 0.5,
           "patch_icpp(1)%vel(1)": 0.0, "patch_icpp(1)%pres": 1.0, "patch_icpp(1)%alpha_rho(1)": 1.0,
           "patch_icpp(1)%alpha(1)": 1.0,
           "patch_icpp(2)%geometry": 1, "patch_icpp(2)%x_centroid": 0.75, "patch_icpp(2)%length_x": 0.5,
           "patch_icpp(2)%vel(1)": 0.0, "patch_icpp(2)%pres": 0.1, "patch_icpp(2)%alpha_rho(1)": 0.125,
           "patch_icpp(2)%alpha(1)": 1.0}
    prm, g, q = PreProcess(sod).state()
    prm_r, g_r, q_r = ic.sod_1d(m=999)
    assert np.array_equal(q, q_r) and np.array_equal(g[0][0], g_r[0][0])
    gam, pinf = 1.0 / (4.4 - 1.0), 4.4 * 6.0e8 / (4.4 - 1.0)
    visc = {"case_dir": "'.'", "run_time_info": "T", "x_domain%beg": -0.5, "x_domain%end": 0.5,
            "y_domain%beg": -0.5, "y_domain%end": 0.5, "m": 50, "n": 50, "dt": 5e-10, "t_step_start": 0,
            "t_step_stop": 10, "t_step_save": 1, "num_patches": 2, "num_fluids": 2, "time_stepper": 3,
            "weno_order": 5, "weno_eps": 1e-16, "weno_Re_flux": "T",
            "bc_x%beg": -1, "bc_x%end": -1, "bc_y%beg": -6, "bc_y%end": -6,
            "format": 1, "precision": 2, "prim_vars_wrt": "T", "parallel_io": "T",
            "patch_icpp(1)%geometry": 3, "patch_icpp(1)%x_centroid": 0.0, "patch_icpp(1)%y_centroid": 0.0,
            "patch_icpp(1)%length_x": 1.0, "patch_icpp(1)%length_y": 1.0, "patch_icpp(1)%vel(1)": 500.0,
            "patch_icpp(1)%vel(2)": 0.0, "patch_icpp(1)%pres": 1e6, "patch_icpp(1)%alpha_rho(1)": 1000.0,
            "patch_icpp(1)%alpha_rho(2)": 1000.0, "patch_icpp(1)%alpha(1)": 0.5, "patch_icpp(1)%alpha(2)": 0.5,
            "patch_icpp(2)%geometry": 3, "patch_icpp(2)%x_centroid": 0.0, "patch_icpp(2)%y_centroid": 0.25,
            "patch_icpp(2)%length_x": 1.0, "patch_icpp(2)%length_y": 0.5, "patch_icpp(2)%alter_patch(1)": "T",
            "patch_icpp(2)%vel(1)": -500.0, "patch_icpp(2)%vel(2)": 0.0, "patch_icpp(2)%pres": 1e6,
            "patch_icpp(2)%alpha_rho(1)": 0.0, "patch_icpp(2)%alpha_rho(2)": 1.0, "patch_icpp(2)%alpha(1)": 0.0,
            "patch_icpp(2)%alpha(2)": 1.0}
    for f in (1, 2):
        visc[f"fluid_pp({f})%gamma"], visc[f"fluid_pp({f})%pi_inf"] = gam, pinf
        visc[f"fluid_pp({f})%Re(1)"] = visc[f"fluid_pp({f})%Re(2)"] = 1e-4
    prm, g, q = PreProcess(visc).state()
    assert prm.buff_size == 6 and prm.weno_Re_flux and visc == ic.viscous_shear_case(N=51)
    # patch 2 holds the rows whose centre lies in [0, 0.5] (y_cc from the
    # reference's formula -0.5 + 0.5 dy (2j + 1): exact for j = 25, where it
    # is 0), patch 1 the rest
    qi = prm.interior(q)
    yc = [Fraction(-1, 2) + Fraction(1, 51) * Fraction(2 * j + 1, 2) for j in range(51)]
    upper = np.array([0 <= y <= Fraction(1, 2) for y in yc])
    assert upper.sum() == 26
    assert np.all(qi[1, 0][upper] == 1.0) and np.all(qi[0, 0][upper] == 0.0)
    assert np.all(qi[0, 0][~upper] == 1000.0) and np.all(qi[2, 0][~upper] == 2000.0 * 500.0)


@pytest.mark.parametrize("parallel", [False, True], ids=["serial", "parallel_io"])
def test_written_files_read_back(tmp_path, parallel):
    circ = dict(geometry=2, x_centroid=0.5, y_centroid=0.25, radius=0.1, smoothen="T", smooth_patch_id=1,
                smooth_coeff=0.5, **_state(1000.0, 0.0, 0.0, 1.0))
    c = _base2d(patches=[BG, circ], parallel_io="T" if parallel else "F")
    pp = PreProcess(c)
    prm, grids, q = pp.write(str(tmp_path))
    if parallel:
        back = C.read_parallel(str(tmp_path), 0, prm)
        cb = C.read_parallel_grid(str(tmp_path), prm)[1]
    else:
        back = C.read_serial(str(tmp_path), 0, prm)
        cb = C.read_record(str(C.t_step_dir(str(tmp_path), 0)) + "/y_cb.dat")
    assert np.array_equal(back, prm.interior(q))
    assert np.array_equal(cb, pp.y_cb)
