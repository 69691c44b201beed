"""Case files and the Fortran host on the CPU: the namelist writer, the
Fortran unformatted records, the driver's dt sequence, and the Fortran
driver reading a case and rejecting an invalid one through the C ABI (the
input checks run before any GPU call)."""
import os
import subprocess

import numpy as np
import pytest

from microfc_amd import case as C
from microfc_amd import ic

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SIM = os.path.join(ROOT, "microfc_amd", "fortran", "simulation")


def test_namelist_format(tmp_path):
    txt = C.write_inp({"m": 99, "run_time_info": True, "dt": 1e-4, "bc_x%beg": -3,
                       "fluid_pp(1)%gamma": 2.5}, str(tmp_path / "simulation.inp"))
    assert txt == "&user_inputs\nm = 99\nrun_time_info = T\ndt = 0.0001\nbc_x%beg = -3\nfluid_pp(1)%gamma = 2.5\n&end/\n"


def test_namelist_rejects_unknown_key(tmp_path):
    with pytest.raises(KeyError):
        C.write_inp({"m": 99, "not_a_key": 1}, str(tmp_path / "simulation.inp"))


def test_fortran_record_round_trip(tmp_path):
    a = np.random.default_rng(0).normal(size=(3, 5, 7))
    p = str(tmp_path / "r.dat")
    C.write_record(p, a)
    raw = open(p, "rb").read()
    assert len(raw) == a.size * 8 + 8
    assert int.from_bytes(raw[:4], "little") == a.size * 8 == int.from_bytes(raw[-4:], "little")
    assert np.array_equal(C.read_record(p).reshape(a.shape), a)


def test_serial_files_round_trip(tmp_path):
    prm, grids, q = ic.shock_bubble_2d(N=32)
    C.write_serial(str(tmp_path), 0, prm, grids, q)
    back = C.read_serial(str(tmp_path), 0, prm)
    assert np.array_equal(back, prm.interior(q))
    cb = C.read_record(os.path.join(C.t_step_dir(str(tmp_path), 0), "x_cb.dat"))
    assert cb.size == prm.m + 2 and cb[0] == 0.0 and cb[-1] == 1.0


def test_dt_sequence_matches_driver_rule():
    seq = C.dt_sequence(0.1, 0, 7)
    assert len(seq) == 7
    assert all(s == 0.1 for s in seq[:-1])
    assert abs(seq[-1] - 0.1) < 1e-15


def _write_case(tmp_path, **over):
    prm, grids, q = ic.sod_1d(m=99)
    case = C.case_from_params(prm, dt=1e-4, t_step_stop=5, t_step_save=5)
    case.update(over)
    C.write_inp(case, str(tmp_path / "simulation.inp"))
    C.write_serial(str(tmp_path), 0, prm, grids, q)


@pytest.mark.skipif(not os.path.exists(SIM), reason="Fortran host not built")
def test_fortran_driver_rejects_invalid_case(tmp_path):
    _write_case(tmp_path, weno_order=4)
    r = subprocess.run([SIM], cwd=str(tmp_path), capture_output=True, text=True, timeout=60)
    assert r.returncode != 0
    assert "mfc_rhs_init" in r.stdout and "weno_order" in r.stdout


@pytest.mark.skipif(not os.path.exists(SIM), reason="Fortran host not built")
def test_fortran_driver_passes_reynolds_numbers(tmp_path):
    """fluid_pp%Re reaches mfc_rhs_init through the bind(C) struct: a viscous
    case in the FP32-bulk form is refused by the library's input check (which
    runs before any GPU call) with the viscous message."""
    _write_case(tmp_path, **{"fluid_pp(1)%Re(1)": 100.0, "bulk_precision": 1})
    r = subprocess.run([SIM], cwd=str(tmp_path), capture_output=True, text=True, timeout=60)
    assert r.returncode != 0
    assert "mfc_rhs_init" in r.stdout and "viscous terms run in the plain FP64 form only" in r.stdout
