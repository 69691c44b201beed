"""Host side of the MI355X RHS library: a Python mirror of MicroFC's
simulation modules over the C ABI of libmfc_rhs.so (include/mfc_rhs.h).

The names follow the reference's interfaces:
  * ``RhsModule.s_compute_rhs(q_cons_vf, q_prim_vf, rhs_vf, t_step)``
    -- src/simulation/m_rhs.fpp:405-410;
  * ``RhsModule.s_tvd_rk(t_step, q_cons_ts, q_prim_vf, rhs_vf, dt)``
    -- s_{1st,2nd,3rd}_order_tvd_rk, src/simulation/m_time_steppers.fpp:129-362;
  * ``populate_grid_buffers`` -- s_populate_grid_variables_buffers,
    src/simulation/m_start_up.fpp:517-655.
Fields are torch tensors on the GPU of shape (sys_size, nzg, nyg, nxg) (x
fastest, ghosted) or (sys_size, p+1, n+1, m+1) (rhs).  torch only provides
device memory; all arithmetic runs in the library's HIP kernels.  There is
no CPU fallback: a missing library raises.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmfc_rhs.so")
NUM_FLUIDS_MAX = 10
PREC_FP64 = 0
PREC_FP32_BULK = 1
PREC_FP32_LEARNED = 2
PREC_FP64_FAST = 3

EXPORTED = [
    "mfc_rhs_init", "mfc_rhs_set_grid", "mfc_compute_rhs", "mfc_tvd_rk_step",
    "mfc_cons_to_prim", "mfc_icfl_max", "mfc_stability_criteria", "mfc_cfl_dt", "mfc_cfl_dt_noalloc",
    "mfc_sizes",
    "mfc_weno_coefficients", "mfc_device_alloc", "mfc_device_free",
    "mfc_memcpy_h2d", "mfc_memcpy_d2h", "mfc_sync", "mfc_stream",
    "mfc_set_stream", "mfc_rhs_finalize", "mfc_last_error",
    "mfc_timing_enable", "mfc_timing_reset", "mfc_timing_read",
    "mfc_mlp_init", "mfc_mlp_get", "mfc_mlp_gradient", "mfc_mlp_train_step", "mfc_mlp_infer",
    "mfc_ml_restrict", "mfc_ml_samples", "mfc_ml_restrict3", "mfc_ml_samples3", "mfc_mlp_last_error",
]
KERNEL_IDS = {"bc_fill": 0, "cons_to_prim": 1, "sweep_x": 2, "sweep_y": 3, "sweep_z": 4, "icfl": 5,
              "rk1_update": 6, "viscous": 7, "sweep_xy": 8}


class MfcError(RuntimeError):
    pass


class CParams(ctypes.Structure):
    _fields_ = [
        ("m", ctypes.c_int), ("n", ctypes.c_int), ("p", ctypes.c_int),
        ("num_fluids", ctypes.c_int), ("weno_order", ctypes.c_int),
        ("time_stepper", ctypes.c_int), ("t_step_stop", ctypes.c_int),
        ("bc_x_beg", ctypes.c_int), ("bc_x_end", ctypes.c_int),
        ("bc_y_beg", ctypes.c_int), ("bc_y_end", ctypes.c_int),
        ("bc_z_beg", ctypes.c_int), ("bc_z_end", ctypes.c_int),
        ("weno_eps", ctypes.c_double),
        ("gammas", ctypes.c_double * NUM_FLUIDS_MAX),
        ("pi_infs", ctypes.c_double * NUM_FLUIDS_MAX),
        ("precision", ctypes.c_int),
        ("Re", (ctypes.c_double * 2) * NUM_FLUIDS_MAX),
        ("weno_Re_flux", ctypes.c_int),
    ]


_LIB = None


def load_library(path: str = LIB_PATH):
    """Loads libmfc_rhs.so and declares its signatures.  Raises when absent."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise MfcError(f"{path} is not built; run __graft_entry__.build() or python microfc_amd/build.py")
    L = ctypes.CDLL(path)
    pp = ctypes.POINTER(ctypes.c_void_p)
    dp = ctypes.POINTER(ctypes.c_double)
    L.mfc_rhs_init.argtypes = [ctypes.POINTER(CParams)]
    L.mfc_rhs_set_grid.argtypes = [ctypes.c_int, dp, dp]
    L.mfc_compute_rhs.argtypes = [pp, pp, pp, ctypes.c_int]
    L.mfc_tvd_rk_step.argtypes = [ctypes.c_int, ctypes.c_double, pp, pp, pp, pp]
    L.mfc_cons_to_prim.argtypes = [pp, pp]
    L.mfc_icfl_max.argtypes = [pp, ctypes.c_double, dp]
    L.mfc_cfl_dt.argtypes = [pp, ctypes.c_double, dp]
    L.mfc_cfl_dt_noalloc.argtypes = [pp, ctypes.c_double, dp]
    L.mfc_stability_criteria.argtypes = [pp, ctypes.c_double, dp]
    L.mfc_sizes.argtypes = [ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                            ctypes.POINTER(ctypes.c_long), ctypes.POINTER(ctypes.c_long)]
    L.mfc_weno_coefficients.argtypes = [ctypes.c_int, dp]
    L.mfc_sync.argtypes = []
    L.mfc_stream.restype = ctypes.c_void_p
    L.mfc_set_stream.argtypes = [ctypes.c_void_p]
    L.mfc_rhs_finalize.argtypes = []
    L.mfc_last_error.restype = ctypes.c_char_p
    L.mfc_timing_enable.argtypes = [ctypes.c_int]
    L.mfc_timing_reset.argtypes = []
    L.mfc_timing_read.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_char_p),
                                  ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_long)]
    _LIB = L
    return L


def _check(rc: int, what: str) -> int:
    if rc < 0:
        msg = load_library().mfc_last_error().decode()
        raise MfcError(f"{what} failed ({rc}): {msg}")
    return rc


@dataclass
class Params:
    """Case parameters with the meaning of the namelist &user_inputs
    (src/simulation/m_start_up.fpp:73-83); p and bc_z extend it to 3D."""
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
    precision: int = PREC_FP64
    # fluid_pp(i)%Re(1:2) per fluid; entries <= 0 leave the fluid out, None
    # (or no positive entry) is inviscid
    Re: list = None
    # viscous velocity gradients: WENO face values by the divergence theorem
    # (True) or finite differences (False)
    weno_Re_flux: bool = False

    @property
    def viscous(self) -> bool:
        return bool(self.Re) and any(r > 0 for pair in self.Re for r in pair)

    @property
    def num_dims(self) -> int:
        return 1 + (self.n > 0) + (self.p > 0)

    @property
    def sys_size(self) -> int:
        return 2 * self.num_fluids + self.num_dims + 1

    @property
    def buff_size(self) -> int:
        """Ghost layers per side, as the reference sizes them
        (src/simulation/m_global_parameters.fpp:355-359): 2*polyn + 2 when any
        fluid has a Reynolds number, else polyn + 2 (polyn = (weno_order-1)/2)."""
        polyn = (self.weno_order - 1) // 2
        return 2 * polyn + 2 if self.viscous else polyn + 2

    @property
    def ghost_shape(self):
        b = self.buff_size
        return (self.sys_size, self.p + 2 * b + 1 if self.p > 0 else 1,
                self.n + 2 * b + 1 if self.n > 0 else 1, self.m + 2 * b + 1)

    @property
    def interior_shape(self):
        return (self.sys_size, self.p + 1, self.n + 1, self.m + 1)

    def extents(self):
        return [self.m, self.n, self.p][: self.num_dims]

    def interior(self, q):
        """View of the interior cells of a ghosted field."""
        b = self.buff_size
        bj = b if self.n > 0 else 0
        bk = b if self.p > 0 else 0
        return q[:, bk:bk + self.p + 1, bj:bj + self.n + 1, b:b + self.m + 1]

    def to_c(self) -> CParams:
        c = CParams()
        c.m, c.n, c.p = self.m, self.n, self.p
        c.num_fluids, c.weno_order = self.num_fluids, self.weno_order
        c.time_stepper, c.t_step_stop = self.time_stepper, self.t_step_stop
        c.bc_x_beg, c.bc_x_end = self.bc_beg[0], self.bc_end[0]
        c.bc_y_beg, c.bc_y_end = self.bc_beg[1], self.bc_end[1]
        c.bc_z_beg, c.bc_z_end = self.bc_beg[2], self.bc_end[2]
        c.weno_eps = self.weno_eps
        for i in range(self.num_fluids):
            c.gammas[i] = self.gammas[i]
            c.pi_infs[i] = self.pi_infs[i]
        c.precision = self.precision
        for i in range(NUM_FLUIDS_MAX):
            for r in range(2):
                c.Re[i][r] = self.Re[i][r] if self.Re and i < len(self.Re) else -1e6
        c.weno_Re_flux = 1 if self.weno_Re_flux else 0
        return c


def populate_grid_buffers(cb_interior, b: int, bc_beg: int, bc_end: int):
    """s_populate_grid_variables_buffers (src/simulation/m_start_up.fpp:517-655)
    for one direction.  cb_interior holds s_cb(-1..s); returns the ghosted
    (s_cb(-1-b..s+b), ds(-b..s+b))."""
    cbi = np.asarray(cb_interior, dtype=np.float64)
    s = cbi.size - 2
    cb = np.zeros(s + 2 * b + 2)
    ds = np.zeros(s + 2 * b + 1)
    C = lambda i: i + 1 + b  # noqa: E731  index of s_cb(i)
    Dx = lambda i: i + b  # noqa: E731  index of ds(i)
    cb[C(-1):C(s) + 1] = cbi
    for i in range(0, s + 1):
        ds[Dx(i)] = cb[C(i)] - cb[C(i - 1)]
    for i in range(1, b + 1):
        if bc_beg <= -3:
            ds[Dx(-i)] = ds[Dx(0)]
        elif bc_beg == -2:
            ds[Dx(-i)] = ds[Dx(i - 1)]
        elif bc_beg == -1:
            ds[Dx(-i)] = ds[Dx(s - (i - 1))]
        else:
            raise ValueError(f"unsupported boundary condition {bc_beg}")
    for i in range(1, b + 1):
        cb[C(-1 - i)] = cb[C(-i)] - ds[Dx(-i)]
    for i in range(1, b + 1):
        if bc_end <= -3:
            ds[Dx(s + i)] = ds[Dx(s)]
        elif bc_end == -2:
            ds[Dx(s + i)] = ds[Dx(s - (i - 1))]
        elif bc_end == -1:
            ds[Dx(s + i)] = ds[Dx(i - 1)]
        else:
            raise ValueError(f"unsupported boundary condition {bc_end}")
    for i in range(1, b + 1):
        cb[C(s + i)] = cb[C(s + (i - 1))] + ds[Dx(s + i)]
    return cb, ds


def uniform_cell_boundaries(s: int, beg: float, end: float):
    """s_cb(-1..s) of a uniform grid as src/pre_process/m_grid.f90:59-64 makes it."""
    dx = (end - beg) / float(s + 1)
    cb = np.empty(s + 2)
    cb[: s + 1] = beg + dx * np.arange(0, s + 1, dtype=np.float64)
    cb[s + 1] = end
    return cb


class RhsModule:
    """One initialised case on the GPU (the library holds module state)."""

    def __init__(self, params: Params, grids=None, stream=None):
        import torch
        if not torch.cuda.is_available():
            raise MfcError("the HIP RHS library needs a GPU")
        self.torch = torch
        self.params = params
        self.lib = load_library()
        _check(self.lib.mfc_rhs_init(ctypes.byref(params.to_c())), "mfc_rhs_init")
        if stream is not None:
            _check(self.lib.mfc_set_stream(ctypes.c_void_p(stream)), "mfc_set_stream")
        self.grids = []
        if grids is None:
            grids = []
            for d, s in enumerate(params.extents()):
                cbi = uniform_cell_boundaries(s, 0.0, 1.0)
                grids.append(populate_grid_buffers(cbi, params.buff_size, params.bc_beg[d], params.bc_end[d]))
        for d, (cb, ds) in enumerate(grids):
            self.set_grid(d, cb, ds)

    # -- set-up ------------------------------------------------------------
    def set_grid(self, direction: int, cb, ds):
        cb = np.ascontiguousarray(cb, dtype=np.float64)
        ds = np.ascontiguousarray(ds, dtype=np.float64)
        dp = ctypes.POINTER(ctypes.c_double)
        _check(self.lib.mfc_rhs_set_grid(direction, cb.ctypes.data_as(dp), ds.ctypes.data_as(dp)),
               "mfc_rhs_set_grid")
        while len(self.grids) <= direction:
            self.grids.append(None)
        self.grids[direction] = (cb, ds)

    def new_field(self, ghosted=True):
        shape = self.params.ghost_shape if ghosted else self.params.interior_shape
        return self.torch.zeros(shape, dtype=self.torch.float64, device="cuda")

    def _ptrs(self, t):
        if t is None:
            return ctypes.cast(None, ctypes.POINTER(ctypes.c_void_p))
        assert t.is_cuda and t.dtype == self.torch.float64 and t.is_contiguous()
        S = self.params.sys_size
        assert t.shape[0] == S
        stride = t.stride(0) * 8
        arr = (ctypes.c_void_p * S)(*[t.data_ptr() + v * stride for v in range(S)])
        return arr

    def _enter(self):
        # work queued by torch on its stream must be visible to the library's
        self.torch.cuda.current_stream().synchronize()

    # -- the reference's interfaces ---------------------------------------
    def s_compute_rhs(self, q_cons_vf, q_prim_vf, rhs_vf, t_step=0):
        self._enter()
        return _check(self.lib.mfc_compute_rhs(self._ptrs(q_cons_vf), self._ptrs(q_prim_vf),
                                               self._ptrs(rhs_vf), t_step), "mfc_compute_rhs")

    def s_tvd_rk(self, t_step, q_cons_ts, q_prim_vf, rhs_vf, dt):
        """One step of the configured TVD-RK scheme; q_cons_ts = (q1, q2)."""
        self._enter()
        q1, q2 = q_cons_ts
        return _check(self.lib.mfc_tvd_rk_step(t_step, dt, self._ptrs(q1), self._ptrs(q2),
                                               self._ptrs(q_prim_vf), self._ptrs(rhs_vf)),
                      "mfc_tvd_rk_step")

    def s_convert_conservative_to_primitive_variables(self, q_cons_vf, q_prim_vf):
        self._enter()
        return _check(self.lib.mfc_cons_to_prim(self._ptrs(q_cons_vf), self._ptrs(q_prim_vf)),
                      "mfc_cons_to_prim")

    def icfl_max(self, q_prim_vf, dt):
        self._enter()
        out = ctypes.c_double()
        _check(self.lib.mfc_icfl_max(self._ptrs(q_prim_vf), dt, ctypes.byref(out)), "mfc_icfl_max")
        return out.value

    def stability_criteria(self, q_prim_vf, dt):
        """(ICFL max, VCFL max, Rc min) of s_write_run_time_information
        (src/simulation/m_data_output.fpp:197-263); the last two are 0 when
        the case is inviscid."""
        self._enter()
        out = (ctypes.c_double * 3)()
        _check(self.lib.mfc_stability_criteria(self._ptrs(q_prim_vf), dt, out), "mfc_stability_criteria")
        return out[0], out[1], out[2]

    def cfl_dt(self, q_cons_vf, cfl):
        self._enter()
        out = ctypes.c_double()
        _check(self.lib.mfc_cfl_dt(self._ptrs(q_cons_vf), cfl, ctypes.byref(out)), "mfc_cfl_dt")
        return out.value

    def weno_coefficients(self, direction):
        s = self.params.extents()[direction]
        out = np.zeros((27, s + 2 * self.params.buff_size + 1))
        _check(self.lib.mfc_weno_coefficients(direction, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
               "mfc_weno_coefficients")
        return out

    def sync(self):
        _check(self.lib.mfc_sync(), "mfc_sync")

    def stream(self):
        return self.lib.mfc_stream()

    def timing(self, enable=None, reset=False):
        """Per-kernel HIP-event timing: returns {name: (total_ms, launches)}."""
        if reset:
            _check(self.lib.mfc_timing_reset(), "mfc_timing_reset")
        if enable is not None:
            _check(self.lib.mfc_timing_enable(1 if enable else 0), "mfc_timing_enable")
        out = {}
        for name, k in KERNEL_IDS.items():
            nm = ctypes.c_char_p()
            ms = ctypes.c_double()
            cnt = ctypes.c_long()
            _check(self.lib.mfc_timing_read(k, ctypes.byref(nm), ctypes.byref(ms), ctypes.byref(cnt)),
                   "mfc_timing_read")
            out[name] = (ms.value, cnt.value)
        return out

    def finalize(self):
        _check(self.lib.mfc_rhs_finalize(), "mfc_rhs_finalize")
