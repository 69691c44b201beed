"""The Fortran host (microfc_amd/fortran, built from the reference's module
structure) running whole cases on the GPU through ISO_C_BINDING, checked
against the CPU FP64 oracle stepped with the driver's own dt sequence."""
import os
import subprocess

import numpy as np
import pytest

from microfc_amd import case as C
from microfc_amd import ic
from microfc_amd.pre_process import PreProcess
from oracle import oracle as O

from helpers import assert_parity, perturbed, reghost, say, to_oracle

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SIM = os.path.join(ROOT, "microfc_amd", "fortran", "simulation")


def _run(tmp_path, prm, grids, q, dt, nsteps, save, run_time_info=False, **extra):
    case = C.case_from_params(prm, dt=dt, t_step_stop=nsteps, t_step_save=save, run_time_info=run_time_info)
    case.update(extra)
    C.write_inp(case, str(tmp_path / "simulation.inp"))
    C.write_serial(str(tmp_path), 0, prm, grids, q)
    say(f"[fortran host] {tmp_path.name} steps={nsteps}")
    r = subprocess.run([SIM], cwd=str(tmp_path), capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return r


def _oracle_steps(prm, grids, q, dt, nsteps, upto):
    oc = to_oracle(prm)
    qo = q.copy()
    q2 = np.zeros_like(qo)
    for t, d in enumerate(C.dt_sequence(dt, 0, nsteps)[:upto]):
        O.tvd_rk_step(oc, grids, d, qo, q2, t_step=t)
    return prm.interior(qo)


def test_fortran_driver_sod(gpu, tmp_path):
    prm, grids, q = ic.sod_1d(m=199)
    q = perturbed(prm, q)
    dt, nsteps = 4.5e-4, 40
    _run(tmp_path, prm, grids, q, dt, nsteps, 20)
    for t in (20, 40):
        got = C.read_serial(str(tmp_path), t, prm)
        assert_parity(got, _oracle_steps(prm, grids, q, dt, nsteps, t), what=f"sod step {t}")
    assert os.path.exists(tmp_path / "time_data.dat")


def test_fortran_driver_bubble_with_run_time_info(gpu, tmp_path):
    prm, grids, q = ic.shock_bubble_2d(N=64)
    q = perturbed(prm, q)
    dt, nsteps = 1e-7, 12
    _run(tmp_path, prm, grids, q, dt, nsteps, 6, run_time_info=True)
    got = C.read_serial(str(tmp_path), 12, prm)
    assert_parity(got, _oracle_steps(prm, grids, q, dt, nsteps, 12), what="bubble step 12")
    text = open(tmp_path / "run_time.inf").read()
    header, rows, footer = _run_time_table(text)
    assert header[0].startswith("Description: Stability information at each time-step")
    assert header[5].startswith("Date: ")
    assert len(rows) == nsteps + 1 and [int(r[0]) for r in rows] == list(range(nsteps + 1))
    oc = to_oracle(prm)
    prim = O.cons_to_prim(oc, O.bc_fill(oc, q.copy()))
    icfl0 = O.icfl_max(oc, grids, prim, dt)
    assert abs(float(rows[0][2]) - icfl0) < 1e-6
    assert footer[0].startswith("ICFL Max: ") and abs(float(footer[0].split()[2]) - max(float(r[2]) for r in rows)) < 1e-6
    assert any(ln.startswith("Run-time: ") for ln in footer)


def _run_time_table(text):
    """(header lines, table rows split into fields, footer lines) of a
    run_time.inf written by s_open/s_write/s_close_run_time_information_file."""
    lines = text.splitlines()
    t = next(i for i, ln in enumerate(lines) if ln.startswith("===") and "Time-steps" in ln)
    f = next(i for i, ln in enumerate(lines) if ln.startswith("-----"))
    return lines[:t + 1], [ln.split() for ln in lines[t + 1:f]], [ln for ln in lines[f + 1:] if ln.strip()]


def test_fortran_compute_rhs_dropin(gpu, tmp_path):
    """s_compute_rhs(q_cons_vf, q_prim_vf, rhs_vf, t_step) called from Fortran
    with the reference's type(scalar_field), dimension(sys_size) arguments:
    rhs and q_prim_vf equal the oracle's bit for bit (2D and 1D)."""
    for name, (prm, grids, q) in [("bubble2d_48", ic.shock_bubble_2d(N=48)), ("sod1d_100", ic.sod_1d(m=99))]:
        q = perturbed(prm, q)
        d = tmp_path / name
        d.mkdir()
        case = C.case_from_params(prm, dt=1e-7, t_step_stop=5, t_step_save=5, run_time_info=True)
        C.write_inp(case, str(d / "simulation.inp"))
        C.write_serial(str(d), 0, prm, grids, q)
        say(f"[fortran s_compute_rhs] {name}")
        r = subprocess.run([os.path.join(ROOT, "microfc_amd", "fortran", "compute_rhs_check")], cwd=str(d),
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        oc = to_oracle(prm)
        rhs_ref, prim_ref = O.compute_rhs(oc, grids, q, t_step=0, want_prim=True)
        got = np.fromfile(d / "rhs.dat", dtype="<f8").reshape(prm.interior_shape)
        assert_parity(got, rhs_ref, tol=0.0, what=f"{name} s_compute_rhs rhs_vf")
        gp = np.fromfile(d / "q_prim.dat", dtype="<f8").reshape(prm.ghost_shape)
        assert_parity(gp, prim_ref, tol=0.0, what=f"{name} s_compute_rhs q_prim_vf")


def test_fortran_driver_nan_guard(gpu, tmp_path):
    """A NaN in the initial condition stops the run with "Error" at the first
    save (p_main.fpp:291-302), after printing the cell's indices."""
    prm, grids, q = ic.sod_1d(m=99)
    b = prm.buff_size
    q[0, 0, 0, b + 37] = np.nan
    case = C.case_from_params(prm, dt=1e-4, t_step_stop=4, t_step_save=2)
    C.write_inp(case, str(tmp_path / "simulation.inp"))
    C.write_serial(str(tmp_path), 0, prm, grids, q)
    r = subprocess.run([SIM], cwd=str(tmp_path), capture_output=True, text=True, timeout=120)
    # STOP "Error" ends the program with the message (its exit status is 0
    # under amdflang, as for the reference built the same way)
    assert "STOP: Error" in r.stderr, r.stderr[-500:]
    assert not (tmp_path / "p_all" / "p0" / "2").exists()
    assert not (tmp_path / "time_data.dat").exists()
    # the guard's line: j, k, rank, t_step, m, n of the first NaN
    loc = [ln.split() for ln in r.stdout.splitlines() if len(ln.split()) == 6 and "Time" not in ln]
    assert loc and loc[-1][1:] == ["0", "0", "2", "99", "0"], r.stdout[-500:]
    assert 0 <= int(loc[-1][0]) <= 37


def test_fortran_driver_viscous_run_time_information(gpu, tmp_path):
    """Viscous runs write the ICFL, VCFL and Rc columns and their extrema
    (m_data_output.fpp:136-141, 280-294, 693-697), the values equal to the
    oracle's criteria of the first stage's primitive field to the printed
    digits."""
    prm, grids, q = ic.viscous_shear_2d(N=51)
    q = perturbed(prm, q)
    dt, nsteps = 5e-10, 4
    _run(tmp_path, prm, grids, q, dt, nsteps, 4, run_time_info=True)
    header, rows, footer = _run_time_table(open(tmp_path / "run_time.inf").read())
    assert header[-1].startswith("==== Time-steps ====== Time ======= ICFL Max ==== VCFL Max ====== Rc Min")
    assert len(rows) == nsteps + 1 and all(len(r) == 5 for r in rows)
    oc = to_oracle(prm)
    prim = O.cons_to_prim(oc, O.bc_fill(oc, q.copy()))
    icfl, vcfl, rc = O.stability(oc, grids, prim, dt)
    assert abs(float(rows[0][2]) - icfl) <= 1e-6 and abs(float(rows[0][3]) - vcfl) <= 1e-6
    assert abs(float(rows[0][4]) - rc) <= 1e-6 * max(1.0, abs(rc))
    names = [ln.split(":")[0] for ln in footer]
    assert names[:3] == ["ICFL Max", "VCFL Max", "Rc Min"]


def test_fortran_driver_reference_viscous_example(gpu, tmp_path):
    """examples/2D_viscous/case.py run unchanged (m = n = 50, dt = 5e-10, 10
    RK3 steps saved every step, Re 1e-4 for both fluids, periodic x,
    weno_Re_flux, run_time_info, parallel_io): the whole dictionary through
    the pre-processor into restart_data/, its simulation keys into
    simulation.inp.  Every saved state equals the oracle's steps with
    WENO-reconstructed velocity gradients bit for bit, and run_time.inf has
    the viscous columns for every step."""
    case = ic.viscous_shear_case(N=51)
    prm, grids, q = PreProcess(case).write(str(tmp_path))
    assert prm.weno_Re_flux and prm.buff_size == 6
    C.write_inp(C.simulation_inputs(case), str(tmp_path / "simulation.inp"))
    say("[fortran host] examples/2D_viscous unchanged")
    r = subprocess.run([SIM], cwd=str(tmp_path), capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    dt, nsteps = case["dt"], case["t_step_stop"]
    oc = to_oracle(prm)
    qo = q.copy()
    q2 = np.zeros_like(qo)
    for t, d in enumerate(C.dt_sequence(dt, 0, nsteps)):
        O.tvd_rk_step(oc, grids, d, qo, q2, t_step=t)
        assert_parity(C.read_parallel(str(tmp_path), t + 1, prm), prm.interior(qo), tol=0.0,
                      what=f"2D_viscous step {t + 1}")
    header, rows, footer = _run_time_table(open(tmp_path / "run_time.inf").read())
    assert len(rows) == nsteps + 1 and all(len(row) == 5 for row in rows)
    # the WENO gradients are not the finite differences here
    fd = prm.__class__(**{**prm.__dict__, "weno_Re_flux": False})
    assert np.max(np.abs(_oracle_steps(fd, grids, q, dt, nsteps, nsteps)[2] - prm.interior(qo)[2])) > 0.0


def test_fortran_driver_viscous_fd_gradients(gpu, tmp_path):
    """The viscous example with finite-difference gradients (weno_Re_flux =
    F) on a perturbed state, serial files: the Fortran host's fluid_pp%Re
    reach the library, and the run equals the oracle's viscous steps."""
    prm, grids, q = ic.viscous_shear_2d(N=51, weno_Re_flux=False)
    assert not prm.weno_Re_flux
    q = perturbed(prm, q)
    dt, nsteps = 5e-10, 10
    _run(tmp_path, prm, grids, q, dt, nsteps, 5)
    for t in (5, 10):
        got = C.read_serial(str(tmp_path), t, prm)
        assert_parity(got, _oracle_steps(prm, grids, q, dt, nsteps, t), what=f"viscous example step {t}")
    # the viscous terms moved the state: the inviscid oracle differs
    prm_inv = prm.__class__(**{**prm.__dict__, "Re": None})
    gi, qi = reghost(prm, prm_inv, grids, q)
    inv = _oracle_steps(prm_inv, gi, qi, dt, nsteps, 10)
    ref = _oracle_steps(prm, grids, q, dt, nsteps, 10)
    assert np.max(np.abs(inv[2] - ref[2])) > 1e-6 * np.max(np.abs(ref[2]))


def test_fortran_driver_aborts_on_cfl_violation(gpu, tmp_path):
    prm, grids, q = ic.sod_1d(m=99)
    case = C.case_from_params(prm, dt=5e-2, t_step_stop=4, t_step_save=2, run_time_info=True)
    C.write_inp(case, str(tmp_path / "simulation.inp"))
    C.write_serial(str(tmp_path), 0, prm, grids, q)
    r = subprocess.run([SIM], cwd=str(tmp_path), capture_output=True, text=True, timeout=120)
    assert r.returncode != 0 and "ICFL is greater than 1.0" in r.stdout


def test_fortran_driver_learned_weights(gpu, tmp_path):
    """bulk_precision = 2 with mlp_weights naming a file of FP32 parameters:
    the driver loads the network, and its run equals the Python host's run
    of the same library bit for bit."""
    torch = gpu
    from microfc_amd import PREC_FP32_LEARNED, RhsModule, learned
    params = learned.init_params(3)
    params[learned.OFF["W3"]:learned.OFF["W3"] + 512] = np.random.default_rng(4).normal(0, 0.3, 512)
    params.astype("<f4").tofile(tmp_path / "mlp.bin")
    prm, grids, q = ic.shock_bubble_2d(N=64)
    q = perturbed(prm, q)
    prm.precision = PREC_FP32_LEARNED
    dt, nsteps = 1e-7, 8
    _run(tmp_path, prm, grids, q, dt, nsteps, 8, mlp_weights="'mlp.bin'")
    got = C.read_serial(str(tmp_path), 8, prm)
    learned.Mlp(params)
    mod = RhsModule(prm, grids)
    try:
        q1 = torch.from_numpy(q.copy()).cuda()
        q2 = mod.new_field()
        rhs = mod.new_field(ghosted=False)
        for t, d in enumerate(C.dt_sequence(dt, 0, nsteps)):
            mod.s_tvd_rk(t, (q1, q2), None, rhs, d)
        mod.sync()
        ref = prm.interior(q1.cpu().numpy())
    finally:
        mod.finalize()
    assert np.array_equal(got, ref)
    # and the network changed the result
    prm.precision = 1
    (tmp_path / "bulk").mkdir()
    _run(tmp_path / "bulk", prm, grids, q, dt, nsteps, 8)
    assert not np.array_equal(C.read_serial(str(tmp_path / "bulk"), 8, prm), got)


def test_fortran_driver_learned_needs_weights(gpu, tmp_path):
    prm, grids, q = ic.sod_1d(m=99)
    prm.precision = 2
    case = C.case_from_params(prm, dt=1e-4, t_step_stop=2, t_step_save=2)
    C.write_inp(case, str(tmp_path / "simulation.inp"))
    C.write_serial(str(tmp_path), 0, prm, grids, q)
    r = subprocess.run([SIM], cwd=str(tmp_path), capture_output=True, text=True, timeout=120)
    assert r.returncode != 0 and "mlp_weights" in r.stdout


def test_fortran_driver_parallel_io(gpu, tmp_path):
    """parallel_io = T reads restart_data/lustre_{x_cb,y_cb,<t>}.dat and
    writes lustre_<t>.dat (m_start_up.fpp:393-511, m_data_output.fpp:477-540):
    the saved states equal the oracle's steps bit for bit, and the run from
    a saved restart file continues exactly as the unbroken run."""
    prm, grids, q = ic.shock_bubble_2d(N=48)
    q = perturbed(prm, q)
    dt, nsteps = 1e-7, 8
    case = C.case_from_params(prm, dt=dt, t_step_stop=nsteps, t_step_save=4, parallel_io=True)
    C.write_inp(case, str(tmp_path / "simulation.inp"))
    C.write_parallel_grid(str(tmp_path), prm, grids)
    C.write_parallel(str(tmp_path), 0, prm, q)
    say("[fortran host] parallel_io")
    r = subprocess.run([SIM], cwd=str(tmp_path), capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert not (tmp_path / "p_all").exists()
    for t in (4, 8):
        got = C.read_parallel(str(tmp_path), t, prm)
        assert_parity(got, _oracle_steps(prm, grids, q, dt, nsteps, t), tol=0.0, what=f"parallel_io step {t}")
    cb = C.read_parallel_grid(str(tmp_path), prm)
    assert np.array_equal(cb[0], grids[0][0][prm.buff_size:prm.buff_size + prm.m + 2])
    # restart from the step-4 file: the oracle continued from that state with
    # the dt sequence a run starting at t_step 4 uses
    r2 = tmp_path / "restart"
    r2.mkdir()
    case2 = dict(case, t_step_start=4)
    C.write_inp(case2, str(r2 / "simulation.inp"))
    C.write_parallel_grid(str(r2), prm, grids)
    (r2 / "restart_data" / "lustre_4.dat").write_bytes((tmp_path / "restart_data" / "lustre_4.dat").read_bytes())
    r = subprocess.run([SIM], cwd=str(r2), capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    oc = to_oracle(prm)
    qo = q.copy()
    prm.interior(qo)[...] = C.read_parallel(str(tmp_path), 4, prm)
    q2 = np.zeros_like(qo)
    for t, d in enumerate(C.dt_sequence(dt, 4, nsteps)):
        O.tvd_rk_step(oc, grids, d, qo, q2, t_step=4 + t)
    assert_parity(C.read_parallel(str(r2), 8, prm), prm.interior(qo), tol=0.0, what="parallel_io restart step 8")
