"""Case files of the simulation host.

* ``write_inp`` turns a case dictionary into ``simulation.inp`` the way the
  reference's toolchain does (toolchain/mfc/run/input.py:14-32): one
  ``key = value`` line per known key inside ``&user_inputs ... &end/``,
  logicals as 'T'/'F', an unknown key is an error.
* ``write_serial`` / ``read_serial`` handle the serial data files
  ``<case_dir>/p_all/p0/<t_step>/{x_cb,y_cb,z_cb,q_cons_vf<i>}.dat``: Fortran
  unformatted sequential records (4-byte length markers around the payload),
  the format src/pre_process/m_data_output.f90 writes and
  src/simulation/m_start_up.fpp:298-390 reads.  q_cons_vf<i> holds the
  interior sf(0:m, 0:n, 0:p) of variable i, x fastest.
* ``write_parallel`` / ``read_parallel`` handle the restart files of
  parallel_io = T: ``<case_dir>/restart_data/lustre_{x_cb,y_cb,z_cb}.dat``
  (the global cell boundaries, raw little-endian doubles) and
  ``lustre_<t_step>.dat``, the interior of every variable as raw doubles,
  variable i at byte offset (m+1)(n+1)(p+1)*8*(i-1), x fastest -- what
  src/simulation/m_data_output.fpp:477-540 writes and
  src/simulation/m_start_up.fpp:393-511 reads through MPI-IO (with p+1 = 1 in
  the reference's 2D).
"""
from __future__ import annotations

import os
import struct

import numpy as np

from .rhs import Params

# Namelist variables of the simulation host: the reference's &user_inputs
# (src/simulation/m_start_up.fpp:73-83) plus p, bc_z and bulk_precision.
SIMULATION_KEYS = ["case_dir", "run_time_info", "m", "n", "p", "dt", "t_step_start", "t_step_stop",
                   "t_step_save", "num_fluids", "time_stepper", "weno_eps", "weno_flat", "riemann_flat",
                   "cu_mpi", "probe_wrt", "fd_order", "num_probes", "t_step_old", "weno_Re_flux",
                   "weno_order", "precision", "parallel_io", "bulk_precision", "mlp_weights"]
for _d in "xyz":
    SIMULATION_KEYS += [f"bc_{_d}%beg", f"bc_{_d}%end"]
for _f in range(1, 11):
    SIMULATION_KEYS += [f"fluid_pp({_f})%gamma", f"fluid_pp({_f})%pi_inf", f"fluid_pp({_f})%Re(1)",
                        f"fluid_pp({_f})%Re(2)"]
for _p in range(1, 11):
    SIMULATION_KEYS += [f"probe({_p})%x", f"probe({_p})%y"]


def _fmt(val):
    if isinstance(val, bool):
        return "T" if val else "F"
    if isinstance(val, float):
        return repr(val)
    return str(val)


def simulation_inputs(case: dict) -> dict:
    """The simulation's keys of a whole case dictionary (the domain, patches,
    stretching and output-format keys of the other programs dropped), as the
    toolchain writes each program's .inp from one dictionary
    (toolchain/mfc/run/input.py:14-32)."""
    return {k: v for k, v in case.items() if k in SIMULATION_KEYS}


def write_inp(case: dict, path: str) -> str:
    lines = []
    for key, val in case.items():
        if key not in SIMULATION_KEYS:
            raise KeyError(f"case parameter '{key}' is not a simulation input")
        lines.append(f"{key} = {_fmt(val)}\n")
    text = "&user_inputs\n" + "".join(lines) + "&end/\n"
    with open(path, "w") as fh:
        fh.write(text)
    return text


def case_from_params(prm: Params, dt: float, t_step_stop: int, t_step_save: int, case_dir: str = "'.'",
                     run_time_info: bool = False, t_step_start: int = 0, parallel_io: bool = False) -> dict:
    c = {"case_dir": case_dir, "run_time_info": run_time_info, "parallel_io": parallel_io,
         "m": prm.m, "n": prm.n, "p": prm.p,
         "dt": dt, "t_step_start": t_step_start, "t_step_stop": t_step_stop, "t_step_save": t_step_save,
         "num_fluids": prm.num_fluids, "time_stepper": prm.time_stepper, "weno_order": prm.weno_order,
         "weno_eps": prm.weno_eps, "bulk_precision": prm.precision}
    if prm.viscous:
        c["weno_Re_flux"] = bool(prm.weno_Re_flux)
    for d, nm in enumerate("xyz"[: prm.num_dims]):
        c[f"bc_{nm}%beg"] = prm.bc_beg[d]
        c[f"bc_{nm}%end"] = prm.bc_end[d]
    for i in range(prm.num_fluids):
        c[f"fluid_pp({i + 1})%gamma"] = float(prm.gammas[i])
        c[f"fluid_pp({i + 1})%pi_inf"] = float(prm.pi_infs[i])
        for r in range(2):
            if prm.Re and i < len(prm.Re) and prm.Re[i][r] > 0:
                c[f"fluid_pp({i + 1})%Re({r + 1})"] = float(prm.Re[i][r])
    return c


def write_record(path: str, arr) -> None:
    data = np.ascontiguousarray(arr, dtype="<f8").tobytes()
    with open(path, "wb") as fh:
        fh.write(struct.pack("<i", len(data)))
        fh.write(data)
        fh.write(struct.pack("<i", len(data)))


def read_record(path: str) -> np.ndarray:
    with open(path, "rb") as fh:
        raw = fh.read()
    (n,) = struct.unpack("<i", raw[:4])
    (n2,) = struct.unpack("<i", raw[4 + n:8 + n])
    if n != n2 or len(raw) != n + 8:
        raise ValueError(f"{path}: not a single Fortran unformatted record")
    return np.frombuffer(raw[4:4 + n], dtype="<f8").copy()


def t_step_dir(case_dir: str, t_step: int) -> str:
    return os.path.join(case_dir, "p_all", "p0", str(t_step))


def write_serial(case_dir: str, t_step: int, prm: Params, grids, q_ghosted) -> None:
    """Writes cell boundaries s_cb(-1:s) and the interior of each variable."""
    d = t_step_dir(case_dir, t_step)
    os.makedirs(d, exist_ok=True)
    b = prm.buff_size
    for k, nm in enumerate("xyz"[: prm.num_dims]):
        s = prm.extents()[k]
        write_record(os.path.join(d, f"{nm}_cb.dat"), grids[k][0][b: b + s + 2])
    qi = prm.interior(q_ghosted)
    for v in range(prm.sys_size):
        write_record(os.path.join(d, f"q_cons_vf{v + 1}.dat"), qi[v])


def read_serial(case_dir: str, t_step: int, prm: Params) -> np.ndarray:
    """Interior conservative field of shape (S, p+1, n+1, m+1)."""
    d = t_step_dir(case_dir, t_step)
    out = np.zeros(prm.interior_shape)
    for v in range(prm.sys_size):
        out[v] = read_record(os.path.join(d, f"q_cons_vf{v + 1}.dat")).reshape(prm.interior_shape[1:])
    return out


def restart_dir(case_dir: str) -> str:
    return os.path.join(case_dir, "restart_data")


def write_parallel_grid(case_dir: str, prm: Params, grids) -> None:
    """lustre_{x,y,z}_cb.dat: s_cb(-1:s) of each direction, raw doubles."""
    d = restart_dir(case_dir)
    os.makedirs(d, exist_ok=True)
    b = prm.buff_size
    for k, nm in enumerate("xyz"[: prm.num_dims]):
        s = prm.extents()[k]
        np.ascontiguousarray(grids[k][0][b: b + s + 2], dtype="<f8").tofile(os.path.join(d, f"lustre_{nm}_cb.dat"))


def write_parallel(case_dir: str, t_step: int, prm: Params, q_ghosted) -> None:
    """lustre_<t_step>.dat: the interior of every variable, one block each."""
    d = restart_dir(case_dir)
    os.makedirs(d, exist_ok=True)
    np.ascontiguousarray(prm.interior(q_ghosted), dtype="<f8").tofile(os.path.join(d, f"lustre_{t_step}.dat"))


def read_parallel(case_dir: str, t_step: int, prm: Params) -> np.ndarray:
    """Interior conservative field of shape (S, p+1, n+1, m+1) from lustre_<t_step>.dat."""
    path = os.path.join(restart_dir(case_dir), f"lustre_{t_step}.dat")
    raw = np.fromfile(path, dtype="<f8")
    if raw.size != int(np.prod(prm.interior_shape)):
        raise ValueError(f"{path}: {raw.size} doubles, expected {int(np.prod(prm.interior_shape))}")
    return raw.reshape(prm.interior_shape).copy()


def read_parallel_grid(case_dir: str, prm: Params):
    """s_cb(-1:s) of each direction from lustre_{x,y,z}_cb.dat."""
    out = []
    for k, nm in enumerate("xyz"[: prm.num_dims]):
        out.append(np.fromfile(os.path.join(restart_dir(case_dir), f"lustre_{nm}_cb.dat"), dtype="<f8"))
    return out


def dt_sequence(dt: float, t_step_start: int, t_step_stop: int):
    """The time-step sizes the driver uses (src/simulation/p_main.fpp:196-287):
    dt is fixed except that it is cut to finaltime - mytime once
    mytime + dt >= finaltime."""
    mytime = 0.0 if t_step_start == 0 else t_step_start * dt
    finaltime = t_step_stop * dt
    out = []
    t = t_step_start
    while True:
        mytime = mytime + dt
        if t == t_step_stop:
            break
        out.append(dt)
        if mytime + dt >= finaltime:
            dt = finaltime - mytime
        t += 1
    return out
