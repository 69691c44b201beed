"""numpy/ctypes front end of the CPU FP64 oracle (oracle/mfc_oracle.c).

TEST INFRASTRUCTURE ONLY: the checker the HIP path is compared against.  Only
tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py import
this module.  The product (microfc_amd/) never does.

Array conventions (same as the HIP library's C ABI):
  * a ghosted field is a float64 array of shape (S, nzg, nyg, nxg), C order,
    so x is fastest and element [v, k+bk, j+bj, i+b] is sf(i, j, k) of
    variable v (Fortran column-major sf(-b:m+b, -b:n+b, -b:p+b));
  * rhs is interior only, shape (S, p+1, n+1, m+1);
  * a grid direction is (cb, ds): cb has s+2b+2 entries starting at
    s_cb(-1-b), ds has s+2b+1 entries starting at ds(-b).
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

NUM_FLUIDS_MAX = 10


class _Case(ctypes.Structure):
    _fields_ = [
        ("m", ctypes.c_int), ("n", ctypes.c_int), ("p", ctypes.c_int),
        ("num_fluids", ctypes.c_int), ("weno_order", ctypes.c_int),
        ("time_stepper", ctypes.c_int), ("t_step_stop", ctypes.c_int),
        ("weno_eps", ctypes.c_double),
        ("gammas", ctypes.c_double * NUM_FLUIDS_MAX),
        ("pi_infs", ctypes.c_double * NUM_FLUIDS_MAX),
        ("bc_beg", ctypes.c_int * 3), ("bc_end", ctypes.c_int * 3),
        ("Re", (ctypes.c_double * 2) * NUM_FLUIDS_MAX),
        ("weno_Re_flux", ctypes.c_int),
    ]


def lib():
    """The compiled oracle.  MFC_ORACLE_LIB names another build of it (the
    mutation checks of tests/test_oracle_mutants.py load deliberately broken
    copies this way)."""
    global _LIB
    if _LIB is None:
        path = os.environ.get("MFC_ORACLE_LIB") or os.path.join(_HERE, "libmfc_oracle.so")
        if not os.path.exists(path):
            import subprocess
            subprocess.run(["make", "-C", _HERE], check=True, capture_output=True)
        L = ctypes.CDLL(path)
        dp = ctypes.POINTER(ctypes.c_double)
        cp = ctypes.POINTER(_Case)
        L.orc_grid_buffers.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, dp, dp]
        L.orc_weno_coefficients.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, dp]
        L.orc_cons_to_prim.argtypes = [cp, dp, dp]
        L.orc_bc_fill.argtypes = [cp, dp]
        arr3 = dp * 3
        L.orc_compute_rhs.argtypes = [cp, arr3, arr3, dp, dp, dp, ctypes.c_int]
        L.orc_tvd_rk_step.argtypes = [cp, arr3, arr3, ctypes.c_double, dp, dp, dp, dp, ctypes.c_int]
        L.orc_icfl_max.argtypes = [cp, arr3, dp, ctypes.c_double, dp]
        L.orc_stability.argtypes = [cp, arr3, dp, ctypes.c_double, dp]
        L.orc_weno_cell.argtypes = [ctypes.c_int, ctypes.c_double, dp, dp, dp, dp]
        L.orc_weno_cell.restype = None
        L.orc_hllc_face.argtypes = [cp, ctypes.c_int, dp, dp, dp, dp]
        L.orc_hllc_face.restype = None
        L.orc_set_learned.argtypes = [ctypes.POINTER(ctypes.c_float)]
        L.orc_set_learned.restype = None
        L.orc_weno5_learned_cell.argtypes = [ctypes.c_double, dp, dp, ctypes.c_int, dp, dp]
        L.orc_weno5_learned_cell.restype = None
        _LIB = L
    return _LIB


def _p(a):
    if a is None:
        return ctypes.cast(None, ctypes.POINTER(ctypes.c_double))
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


@dataclass
class Case:
    """Case parameters with the reference's meaning (m_global_parameters.fpp)."""
    m: int
    n: int = 0
    p: int = 0
    num_fluids: int = 1
    weno_order: int = 5
    time_stepper: int = 3
    t_step_stop: int = -100
    weno_eps: float = 1e-16
    gammas: list = field(default_factory=lambda: [2.5])
    pi_infs: list = field(default_factory=lambda: [0.0])
    bc_beg: tuple = (-3, -3, -3)
    bc_end: tuple = (-3, -3, -3)
    Re: list = None  # per fluid (Re(1), Re(2)); a value <= 0 leaves the fluid out
    weno_Re_flux: bool = False  # velocity gradients from WENO face values (True) or finite differences

    @property
    def viscous(self):
        return bool(self.Re) and any(r > 0 for pair in self.Re for r in pair)

    @property
    def num_dims(self):
        return 1 + (self.n > 0) + (self.p > 0)

    @property
    def sys_size(self):
        return 2 * self.num_fluids + self.num_dims + 1

    @property
    def buff_size(self):
        """m_global_parameters.fpp:355-359: 2*polyn + 2 ghost layers when any
        fluid has a Reynolds number, else polyn + 2."""
        polyn = (self.weno_order - 1) // 2
        return 2 * polyn + 2 if self.viscous else polyn + 2

    @property
    def ghost_shape(self):
        b = self.buff_size
        return (self.sys_size,
                self.p + 2 * b + 1 if self.p > 0 else 1,
                self.n + 2 * b + 1 if self.n > 0 else 1,
                self.m + 2 * b + 1)

    @property
    def interior_shape(self):
        return (self.sys_size, self.p + 1, self.n + 1, self.m + 1)

    def extents(self):
        return [self.m, self.n, self.p][: self.num_dims]

    def c(self):
        s = _Case()
        s.m, s.n, s.p = self.m, self.n, self.p
        s.num_fluids, s.weno_order = self.num_fluids, self.weno_order
        s.time_stepper, s.t_step_stop = self.time_stepper, self.t_step_stop
        s.weno_eps = self.weno_eps
        for i in range(self.num_fluids):
            s.gammas[i] = self.gammas[i]
            s.pi_infs[i] = self.pi_infs[i]
        for d in range(3):
            s.bc_beg[d] = self.bc_beg[d]
            s.bc_end[d] = self.bc_end[d]
        for i in range(NUM_FLUIDS_MAX):
            for r in range(2):
                s.Re[i][r] = self.Re[i][r] if self.Re and i < len(self.Re) else -1e6
        s.weno_Re_flux = 1 if self.weno_Re_flux else 0
        return s


def grid_buffers(cb_interior, b, bc_beg, bc_end):
    """cb_interior: s_cb(-1..s).  Returns (cb, ds, cc) with ghosts."""
    cbi = np.asarray(cb_interior, dtype=np.float64)
    s = len(cbi) - 2
    cb = np.zeros(s + 2 * b + 2)
    cb[b: b + s + 2] = cbi
    ds = np.zeros(s + 2 * b + 1)
    cc = np.zeros(s + 2 * b + 1)
    rc = lib().orc_grid_buffers(s, b, bc_beg, bc_end, _p(cb), _p(ds), _p(cc))
    if rc != 0:
        raise ValueError("unsupported boundary condition")
    return cb, ds, cc


def uniform_grid(case: Case, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0)):
    """Uniform grid as src/pre_process/m_grid.f90:59-64 builds it
    (cb(i-1) = beg + dx*i for i = 0..s, cb(s) = end), with ghosts."""
    out = []
    for d, s in enumerate(case.extents()):
        dx = (hi[d] - lo[d]) / float(s + 1)
        cbi = np.empty(s + 2)
        cbi[: s + 1] = lo[d] + dx * np.arange(0, s + 1, dtype=np.float64)
        cbi[s + 1] = hi[d]
        cb, ds, _ = grid_buffers(cbi, case.buff_size, case.bc_beg[d], case.bc_end[d])
        out.append((cb, ds))
    return out


def weno_coefficients(cb, s, b, weno_order):
    out = np.zeros((27, s + 2 * b + 1))
    lib().orc_weno_coefficients(s, b, weno_order, _p(np.ascontiguousarray(cb)), _p(out))
    return out


def _grid_args(grid):
    dp = ctypes.POINTER(ctypes.c_double)
    cbs = (dp * 3)()
    dss = (dp * 3)()
    for d in range(3):
        if d < len(grid):
            cbs[d] = _p(grid[d][0])
            dss[d] = _p(grid[d][1])
        else:
            cbs[d] = ctypes.cast(None, dp)
            dss[d] = ctypes.cast(None, dp)
    return cbs, dss


def cons_to_prim(case: Case, q_cons):
    q_prim = np.empty_like(q_cons)
    lib().orc_cons_to_prim(ctypes.byref(case.c()), _p(q_cons), _p(q_prim))
    return q_prim


def bc_fill(case: Case, q):
    rc = lib().orc_bc_fill(ctypes.byref(case.c()), _p(q))
    if rc != 0:
        raise ValueError("unsupported boundary condition")
    return q


def compute_rhs(case: Case, grid, q_cons, t_step=0, want_prim=False):
    rhs = np.zeros(case.interior_shape)
    q_prim = np.zeros(case.ghost_shape) if want_prim else None
    cbs, dss = _grid_args(grid)
    rc = lib().orc_compute_rhs(ctypes.byref(case.c()), cbs, dss, _p(q_cons), _p(q_prim), _p(rhs), t_step)
    if rc < 0:
        raise RuntimeError(f"orc_compute_rhs failed: {rc}")
    return (rhs, q_prim) if want_prim else rhs


def tvd_rk_step(case: Case, grid, dt, q1, q2=None, t_step=0):
    """Advance q1 (ghosted) in place by one TVD-RK step."""
    if q2 is None:
        q2 = np.zeros_like(q1)
    rhs = np.zeros(case.interior_shape)
    cbs, dss = _grid_args(grid)
    rc = lib().orc_tvd_rk_step(ctypes.byref(case.c()), cbs, dss, dt, _p(q1), _p(q2), _p(None), _p(rhs), t_step)
    if rc < 0:
        raise RuntimeError(f"orc_tvd_rk_step failed: {rc}")
    return q1


def icfl_max(case: Case, grid, q_prim, dt):
    out = ctypes.c_double(0.0)
    cbs, dss = _grid_args(grid)
    lib().orc_icfl_max(ctypes.byref(case.c()), dss, _p(q_prim), dt, ctypes.byref(out))
    return out.value


def stability(case: Case, grid, q_prim, dt):
    """(ICFL max, VCFL max, Rc min) of m_data_output.fpp:197-263."""
    out = np.zeros(3)
    cbs, dss = _grid_args(grid)
    if lib().orc_stability(ctypes.byref(case.c()), dss, _p(q_prim), dt, _p(out)) < 0:
        raise RuntimeError("orc_stability: viscous criteria are 1D and 2D only")
    return float(out[0]), float(out[1]), float(out[2])


def weno_cell(weno_order, weno_eps, coef27, stencil5):
    vl, vr = ctypes.c_double(), ctypes.c_double()
    cf = np.ascontiguousarray(coef27, dtype=np.float64)
    st = np.ascontiguousarray(stencil5, dtype=np.float64)
    lib().orc_weno_cell(weno_order, weno_eps, _p(cf), _p(st), ctypes.byref(vl), ctypes.byref(vr))
    return vl.value, vr.value


def set_learned(params):
    """Learned WENO5 weights for compute_rhs / tvd_rk_step: the 2128 MLP
    parameters (FP32), or None for the plain weights."""
    global _LEARNED
    if params is None:
        _LEARNED = None
        lib().orc_set_learned(ctypes.cast(None, ctypes.POINTER(ctypes.c_float)))
        return
    _LEARNED = np.ascontiguousarray(params, dtype=np.float32).copy()
    assert _LEARNED.size == 2128
    lib().orc_set_learned(_LEARNED.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))


_LEARNED = None


def weno5_learned_cell(weno_eps, coef27, stencil5, kind=0):
    """Learned reconstruction of one cell (set_learned first): (vL, vR)."""
    vl, vr = ctypes.c_double(), ctypes.c_double()
    cf = np.ascontiguousarray(coef27, dtype=np.float64)
    st = np.ascontiguousarray(stencil5, dtype=np.float64)
    lib().orc_weno5_learned_cell(weno_eps, _p(cf), _p(st), kind, ctypes.byref(vl), ctypes.byref(vr))
    return vl.value, vr.value


def hllc_face(case: Case, direction, qL, qR):
    flux = np.zeros(case.sys_size)
    us = ctypes.c_double()
    lib().orc_hllc_face(ctypes.byref(case.c()), direction,
                        _p(np.ascontiguousarray(qL, dtype=np.float64)),
                        _p(np.ascontiguousarray(qR, dtype=np.float64)), _p(flux), ctypes.byref(us))
    return flux, us.value


def prim_to_cons(case: Case, q_prim):
    """Primitive -> conservative (m_variables_conversion.fpp:385-454, the
    pre_process direction), vectorised over the trailing axes."""
    nf, D = case.num_fluids, case.num_dims
    q = np.array(q_prim, dtype=np.float64, copy=True)
    rho = np.zeros(q.shape[1:])
    gamma = np.zeros(q.shape[1:])
    pi_inf = np.zeros(q.shape[1:])
    for i in range(nf):
        rho = rho + q_prim[i]
        gamma = gamma + q_prim[nf + D + 1 + i] * case.gammas[i]
        pi_inf = pi_inf + q_prim[nf + D + 1 + i] * case.pi_infs[i]
    dyn = np.zeros(q.shape[1:])
    for i in range(nf, nf + D):
        q[i] = rho * q_prim[i]
        dyn = dyn + q[i] * q_prim[i] / 2.0
    q[nf + D] = gamma * q_prim[nf + D] + dyn + pi_inf
    return q
