"""Learned WENO5 weights: host side of the MLP kernels in libmfc_rhs.so.

* ``Mlp`` wraps the C ABI (mfc_mlp_*): parameter init/load/save, gradient,
  Adam steps and inference, all on the GPU in BF16 MFMA with FP32 accumulate.
* ``training_pairs`` makes training data with the plain solver itself: it
  runs a fine case, restricts each snapshot 2x (cell averages of the
  conservative variables), converts both grids to primitive variables and
  pairs every coarse face's WENO5 stencil with the fine-grid face value
  (mfc_ml_samples).
* ``train`` runs mini-batch Adam over such pairs.

The model (5 -> 32 -> 32 -> 3, ReLU; logits modulate the plain WENO5
weights) is described in include/mfc_rhs.h.  It has no reference
counterpart; it replaces the nonlinear weights of src/simulation/m_weno.fpp:508-528.
"""
from __future__ import annotations

import ctypes

import numpy as np

from .rhs import MfcError, Params, RhsModule, load_library

NPARAM = 2128
OFF = dict(W1=0, b1=512, W2=544, b2=1568, W3=1600, b3=2112)


def _lib():
    L = load_library()
    if not getattr(L, "_mlp_declared", False):
        fp = ctypes.POINTER(ctypes.c_float)
        dp = ctypes.POINTER(ctypes.c_double)
        L.mfc_mlp_init.argtypes = [fp]
        L.mfc_mlp_get.argtypes = [fp]
        L.mfc_mlp_gradient.argtypes = [ctypes.c_void_p, ctypes.c_long, fp, fp]
        L.mfc_mlp_train_step.argtypes = [ctypes.c_void_p, ctypes.c_long, ctypes.c_float, fp]
        L.mfc_mlp_infer.argtypes = [ctypes.c_void_p, ctypes.c_long, ctypes.c_void_p]
        L.mfc_ml_restrict.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        L.mfc_ml_samples.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                     ctypes.c_void_p, dp, ctypes.c_double, ctypes.c_void_p]
        L.mfc_ml_restrict3.argtypes = [ctypes.c_int] * 4 + [ctypes.c_void_p, ctypes.c_void_p]
        L.mfc_ml_samples3.argtypes = [ctypes.c_int] * 5 + [ctypes.c_void_p, ctypes.c_void_p, dp, ctypes.c_double,
                                                           ctypes.c_void_p]
        L.mfc_mlp_last_error.restype = ctypes.c_char_p
        L._mlp_declared = True
    return L


def _check(rc, what):
    if rc < 0:
        raise MfcError(f"{what} failed ({rc}): {_lib().mfc_mlp_last_error().decode()}")
    return rc


def init_params(seed: int = 0) -> np.ndarray:
    """He-initialised hidden layers and a zero output layer, so the untrained
    model reproduces the plain WENO5 weights exactly."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM, dtype=np.float32)
    W1 = np.zeros((16, 32), dtype=np.float32)
    W1[:5] = rng.normal(0.0, np.sqrt(2.0 / 5.0), (5, 32))
    p[OFF["W1"]:OFF["W1"] + 512] = W1.reshape(-1)
    p[OFF["W2"]:OFF["W2"] + 1024] = rng.normal(0.0, np.sqrt(2.0 / 32.0), 1024)
    return p


def _fptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class Mlp:
    def __init__(self, params=None, seed: int = 0):
        import torch
        if not torch.cuda.is_available():
            raise MfcError("the MLP kernels need a GPU")
        self.torch = torch
        self.load(init_params(seed) if params is None else params)

    def load(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        assert p.size == NPARAM
        _check(_lib().mfc_mlp_init(_fptr(p)), "mfc_mlp_init")

    def params(self) -> np.ndarray:
        out = np.zeros(NPARAM, dtype=np.float32)
        _check(_lib().mfc_mlp_get(_fptr(out)), "mfc_mlp_get")
        return out

    def _samples(self, S):
        assert S.is_cuda and S.dtype == self.torch.float32 and S.is_contiguous() and S.shape[0] >= 12
        self.torch.cuda.current_stream().synchronize()
        return ctypes.c_void_p(S.data_ptr()), S.shape[1]

    def gradient(self, S):
        ptr, n = self._samples(S)
        g = np.zeros(NPARAM, dtype=np.float32)
        loss = ctypes.c_float()
        _check(_lib().mfc_mlp_gradient(ptr, n, _fptr(g), ctypes.byref(loss)), "mfc_mlp_gradient")
        return g, loss.value

    def train_step(self, S, lr: float) -> float:
        ptr, n = self._samples(S)
        loss = ctypes.c_float()
        _check(_lib().mfc_mlp_train_step(ptr, n, lr, ctypes.byref(loss)), "mfc_mlp_train_step")
        return loss.value

    def infer(self, S):
        out = self.torch.zeros((4, S.shape[1]), dtype=self.torch.float32, device="cuda")
        ptr, n = self._samples(S)  # also orders the zero fill of out before the kernel
        _check(_lib().mfc_mlp_infer(ptr, n, ctypes.c_void_p(out.data_ptr())), "mfc_mlp_infer")
        return out

    def save(self, path: str):
        np.savez(path, params=self.params())

    @staticmethod
    def read(path: str) -> np.ndarray:
        with np.load(path, allow_pickle=False) as f:
            return np.asarray(f["params"], dtype=np.float32)


def _torch_done():
    """The library runs on its own streams: work torch queued on the caller's
    buffers (allocation fills, copies) must be complete before a call."""
    import torch
    torch.cuda.current_stream().synchronize()


def restrict(m: int, n: int, b: int, fine_var, coarse_var, p: int = 0):
    _torch_done()
    _check(_lib().mfc_ml_restrict3(m, n, p, b, ctypes.c_void_p(fine_var.data_ptr()),
                                   ctypes.c_void_p(coarse_var.data_ptr())), "mfc_ml_restrict3")


def samples(direction: int, m: int, n: int, b: int, coarse_var, fine_var, coef27, eps: float, p: int = 0):
    import torch
    ns = 2 * (m + 1) * (n + 1) * (p + 1)
    out = torch.zeros((13, ns), dtype=torch.float32, device="cuda")
    cf = np.ascontiguousarray(coef27, dtype=np.float64)
    _torch_done()
    _check(_lib().mfc_ml_samples3(direction, m, n, p, b, ctypes.c_void_p(coarse_var.data_ptr()),
                                  ctypes.c_void_p(fine_var.data_ptr()),
                                  cf.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), eps,
                                  ctypes.c_void_p(out.data_ptr())), "mfc_ml_samples3")
    return out


def coarse_params(fine: Params) -> Params:
    """The 2x coarser case of a fine one (same boundaries and fluids)."""
    return Params(m=(fine.m + 1) // 2 - 1, n=(fine.n + 1) // 2 - 1 if fine.n > 0 else 0,
                  p=(fine.p + 1) // 2 - 1 if fine.p > 0 else 0,
                  num_fluids=fine.num_fluids, weno_order=5, time_stepper=fine.time_stepper, weno_eps=fine.weno_eps,
                  gammas=list(fine.gammas), pi_infs=list(fine.pi_infs), bc_beg=tuple(fine.bc_beg),
                  bc_end=tuple(fine.bc_end), precision=0)


def keep(S):
    """Samples worth learning from: a stencil whose variation is not rounding
    noise (the sample kernel's validity flag) and a fine-grid target inside
    the hull of the three candidates.  The model only moves the weights of a
    convex combination, so a target outside the hull (about 5 % of the faces
    of a shear layer) carries an error no weights can remove; left in, these
    few faces dominate the mean squared loss and training trades accuracy on
    the others for them."""
    p, t = S[5:8], S[11]
    return (S[12] > 0.5) & (t >= p.amin(dim=0)) & (t <= p.amax(dim=0))


def training_pairs(fine_prm: Params, fine_grids, q_fine, nsnap: int, steps_between: int, cfl: float = 0.4,
                   max_per_snapshot: int | None = None, seed: int = 0):
    """Runs the plain FP64 solver on the fine case and returns the valid
    training samples (12 x N float32 on the GPU) of every snapshot, both
    directions and every variable."""
    import torch
    from . import ic
    assert fine_prm.weno_order == 5
    fine = RhsModule(fine_prm, fine_grids)
    q1 = torch.from_numpy(q_fine).cuda()
    q2 = fine.new_field()
    rhs = fine.new_field(ghosted=False)
    snaps = []
    t = 0
    for s in range(nsnap):
        dt = fine.cfl_dt(q1, cfl)
        for _ in range(steps_between):
            fine.s_tvd_rk(t, (q1, q2), None, rhs, dt)
            t += 1
        prim = fine.new_field()
        fine.s_compute_rhs(q1, prim, rhs, t)
        fine.sync()
        snaps.append((q1.clone(), prim))
    fine.finalize()

    cprm = coarse_params(fine_prm)
    cgrids = ic.uniform_grids(cprm)
    coarse = RhsModule(cprm, cgrids)
    coef = coarse.weno_coefficients(0)[:, cprm.buff_size]
    b = cprm.buff_size
    rng = np.random.default_rng(seed)
    out = []
    try:
        for qf, pf in snaps:
            qc = coarse.new_field()
            for v in range(cprm.sys_size):
                restrict(cprm.m, cprm.n, b, qf[v], qc[v], p=cprm.p)
            pc = coarse.new_field()
            rc = coarse.new_field(ghosted=False)
            coarse.s_compute_rhs(qc, pc, rc, 0)
            coarse.sync()
            for d in range(cprm.num_dims):
                for v in range(cprm.sys_size):
                    S = samples(d, cprm.m, cprm.n, b, pc[v], pf[v], coef, cprm.weno_eps, p=cprm.p)
                    S = S[:12, keep(S)]
                    if max_per_snapshot and S.shape[1] > max_per_snapshot:
                        idx = torch.from_numpy(rng.choice(S.shape[1], max_per_snapshot, replace=False)).cuda()
                        S = S[:, idx]
                    out.append(S)
    finally:
        coarse.finalize()
    return torch.cat(out, dim=1).contiguous()


def train(mlp: Mlp, data, epochs: int = 5, batch: int = 65536, lr: float = 1e-3, seed: int = 0, log=None):
    """Mini-batch Adam over the samples; returns the per-step losses."""
    import torch
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    n = data.shape[1]
    losses = []
    for ep in range(epochs):
        perm = torch.randperm(n, device="cuda", generator=g)
        for s in range(0, n, batch):
            idx = perm[s:s + batch]
            if idx.numel() < 16:
                continue
            B = data[:, idx].contiguous()
            losses.append(mlp.train_step(B, lr))
        if log:
            log(f"epoch {ep}: loss {losses[-1]:.4e}")
    return losses


def run(prm: Params, grids, q, t_end: float, cfl: float = 0.4, dt: float | None = None, progress=None,
        progress_every: int = 100):
    """Runs a case to t_end with TVD-RK3 and a fixed dt (from the CFL
    reduction at t = 0 unless given).  Returns the interior conservative
    field (numpy), the number of steps and the GPU seconds of the steps.
    progress(t), when given, is called after every progress_every steps
    have completed on the GPU (for long runs)."""
    import math

    import torch
    mod = RhsModule(prm, grids)
    try:
        q1 = torch.from_numpy(q.copy()).cuda()
        q2 = mod.new_field()
        rhs = mod.new_field(ghosted=False)
        if dt is None:
            dt = mod.cfl_dt(q1, cfl)
        nsteps = max(1, int(math.ceil(t_end / dt - 1e-12)))
        dt = t_end / nsteps
        mod.sync()
        stream = torch.cuda.ExternalStream(mod.stream())
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for t in range(nsteps):
            mod.s_tvd_rk(t, (q1, q2), None, rhs, dt)
            if progress is not None and (t + 1) % progress_every == 0:
                mod.sync()
                progress(t + 1)
        e1.record(stream)
        mod.sync()
        e1.synchronize()
        sec = e0.elapsed_time(e1) / 1e3
        out = prm.interior(q1.cpu().numpy())
    finally:
        mod.finalize()
    return out, nsteps, sec


def restrict_host(fine_interior):
    """2x cell averages of an interior field: (S, 1, 2n, 2m) -> (S, 1, n, m)
    in 2D, (S, 2p, 2n, 2m) -> (S, p, n, m) in 3D."""
    f = np.asarray(fine_interior)
    if f.shape[1] == 1:
        return 0.25 * (f[:, :, 0::2, 0::2] + f[:, :, 0::2, 1::2] + f[:, :, 1::2, 0::2] + f[:, :, 1::2, 1::2])
    return 0.125 * (f[:, 0::2, 0::2, 0::2] + f[:, 0::2, 0::2, 1::2] + f[:, 0::2, 1::2, 0::2]
                     + f[:, 0::2, 1::2, 1::2] + f[:, 1::2, 0::2, 0::2] + f[:, 1::2, 0::2, 1::2]
                     + f[:, 1::2, 1::2, 0::2] + f[:, 1::2, 1::2, 1::2])


def l1_error(a, b, var: int = 0):
    return float(np.mean(np.abs(a[var] - b[var])))
