"""The allocation-free CFL time step (mfc_cfl_dt_noalloc) against mfc_cfl_dt,
called through the C ABI: the same dt, bit for bit, on two calls in a row,
the second on a smaller grid (a new case; the fixed partials array keeps the
first call's contents)."""
import ctypes

import numpy as np
import pytest

from microfc_amd import RhsModule, ic
from microfc_amd.rhs import load_library

from helpers import perturbed, say


def _dt_pair(mod, q, cfl):
    lib = load_library()
    ptrs = mod._ptrs(q)
    mod._enter()
    a, b = ctypes.c_double(), ctypes.c_double()
    assert lib.mfc_cfl_dt(ptrs, cfl, ctypes.byref(a)) == 0, lib.mfc_last_error().decode()
    assert lib.mfc_cfl_dt_noalloc(ptrs, cfl, ctypes.byref(b)) == 0, lib.mfc_last_error().decode()
    return a.value, b.value


@pytest.mark.gpu
def test_cfl_dt_noalloc_equals_cfl_dt(gpu):
    torch = gpu
    for name, (prm, grids, q) in [("bubble2d_256", ic.shock_bubble_2d(N=256)),
                                  ("droplet3d_28", ic.shock_droplet_3d(N=28))]:
        say(f"[noalloc] {name}")
        q = perturbed(prm, q)
        mod = RhsModule(prm, grids)
        try:
            qd = torch.from_numpy(q.copy()).cuda()
            ref, got = _dt_pair(mod, qd, 0.4)
            assert np.isfinite(ref) and ref > 0
            assert np.float64(got).view(np.int64) == np.float64(ref).view(np.int64), (name, got, ref)
        finally:
            mod.finalize()


@pytest.mark.gpu
def test_cfl_dt_noalloc_rejects_uninitialised(gpu):
    lib = load_library()
    lib.mfc_rhs_finalize()
    out = ctypes.c_double()
    arr = (ctypes.c_void_p * 1)()
    assert lib.mfc_cfl_dt_noalloc(arr, 0.4, ctypes.byref(out)) == -4
