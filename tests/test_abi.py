"""CPU checks of the C ABI: the library loads, exports every function that
include/mfc_rhs.h declares, and rejects invalid cases before touching the
GPU (the input checks of src/simulation/m_start_up.fpp:117-290)."""
import ctypes
import os
import re

import pytest

from microfc_amd import rhs as R

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"\b(mfc_\w+)\s*\(", txt))
    return names


def test_header_declares_the_wrapper_list():
    assert declared_functions() == set(R.EXPORTED)


def test_library_exports_every_declared_symbol():
    lib = R.load_library()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def _init(**kw):
    base = dict(m=99, n=0, p=0, num_fluids=1, weno_order=5, time_stepper=3)
    base.update(kw)
    prm = R.Params(**base)
    lib = R.load_library()
    c = prm.to_c()
    rc = lib.mfc_rhs_init(ctypes.byref(c))
    return rc, lib.mfc_last_error().decode()


@pytest.mark.parametrize("kw,msg", [
    (dict(m=0), "m"),
    (dict(weno_order=4), "weno_order"),
    (dict(m=20), "m and weno_order"),
    (dict(weno_eps=1e-3), "weno_eps"),
    (dict(bc_beg=(-1, -3, -3)), "bc_x"),
    (dict(bc_beg=(-13, -3, -3), bc_end=(-13, -3, -3)), "bc_x"),
    (dict(num_fluids=0, gammas=[], pi_infs=[]), "Unsupported value of num_fluids"),
    (dict(num_fluids=5, gammas=[2.5] * 5, pi_infs=[0.0] * 5, precision=1), "FP32 forms are built for 1..4"),
    (dict(gammas=[-1.0]), "gamma"),
    (dict(pi_infs=[-1.0]), "pi_inf"),
    (dict(time_stepper=4), "time_stepper"),
    (dict(n=0, p=5), "p > 0"),
    (dict(weno_eps=1e-120), "weno_eps below 1e-100"),
    (dict(precision=2, weno_order=3), "weno_order = 5"),
    (dict(precision=7), "precision"),
    (dict(precision=1, weno_eps=1e-20), "weno_eps below 1e-18"),
    (dict(precision=3, weno_eps=1e-30), "weno_eps below 1e-25"),
])
def test_invalid_cases_are_rejected(kw, msg):
    rc, err = _init(**kw)
    assert rc == -1
    assert msg in err


def test_parity_matrix_cases_pass_the_input_checks():
    """Every case of the GPU parity matrix is one the library accepts: the
    input checks run before any device call, so on a machine without a GPU
    a valid case fails later (stream creation), never with rc -1."""
    from helpers import small_cases
    lib = R.load_library()
    for name, prm, _, _ in small_cases():
        c = prm.to_c()
        rc = lib.mfc_rhs_init(ctypes.byref(c))
        msg = lib.mfc_last_error().decode()
        if rc == 0:
            lib.mfc_rhs_finalize()
        assert rc != -1, (name, msg)
