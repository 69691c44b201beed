"""Mutation check of the oracle's pins: plausible slips planted in a copy of
oracle/mfc_oracle.c (compiled into a temporary directory, never into the
tree) must each make the pin files that cover that code fail, and the
unmodified source built the same way must pass them.  This is what shows
that the 3D checks pin the z code, and the viscous checks the weno_Re_flux
gradients, and do not merely restate them.
"""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "mfc_oracle.c")

PINS_3D = ("test_oracle_pinning_3d.py",)
PINS_WRF = ("test_oracle_viscous.py", "test_oracle_weno_re_flux.py")

# (name, text in mfc_oracle.c, replacement, pin files)
MUTANTS = [
    # the z ghost fill negates the y momentum instead of the z momentum
    ("z_fill_negates_rho_v",
     "(flip && v == g.momxb + 2) ? -val : val;", "(flip && v == g.momxb + 1) ? -val : val;", PINS_3D),
    # the z ghost fill negates no momentum at a reflective wall
    ("z_fill_no_negation",
     "(flip && v == g.momxb + 2) ? -val : val;", "(0 && v == g.momxb + 2) ? -val : val;", PINS_3D),
    # the z ghost fill covers only the interior x-y rows (corners left unset)
    ("z_fill_interior_rows_only",
     "        for (j = -g.b; j <= g.n + g.b; j++)\n            for (i = -g.b; i <= g.m + g.b; i++)\n"
     "                for (v = 0; v < g.S; v++) {\n                    double val = q[v * g.ng + gidx(&g, i, j, src)];",
     "        for (j = 0; j <= g.n; j++)\n            for (i = 0; i <= g.m; i++)\n"
     "                for (v = 0; v < g.S; v++) {\n                    double val = q[v * g.ng + gidx(&g, i, j, src)];", PINS_3D),
    # z pencils divide by the y cell widths
    ("z_divides_by_dy",
     "coef, qc, qp, ds[dir],", "coef, qc, qp, ds[dir == 2 ? 1 : dir],", PINS_3D),
    # the z flux divergence overwrites the x and y contributions
    ("z_assigns_rhs", "            if (dir == 0)\n", "            if (dir == 0 || dir == 2)\n", PINS_3D),
    # weno_Re_flux: the divergence theorem differences the face values the
    # wrong way round
    ("wrf_theorem_sign", "1.0 / ds[e][j + b] * (vR - vL);", "1.0 / ds[e][j + b] * (vL - vR);", PINS_WRF),
    # weno_Re_flux: the gradient of velocity component i is taken from
    # component D-1-i
    ("wrf_velocity_component", "stencil_of(c, g, vel + (long)i * g->ng, e, j, t, st);",
     "stencil_of(c, g, vel + (long)(D - 1 - i) * g->ng, e, j, t, st);", PINS_WRF),
    # weno_Re_flux: the face values of the x gradient are taken from the
    # gradient along the sweep direction instead
    ("wrf_wrong_gradient_plane", "stencil_of(c, g, G[e] + (long)i * g->ng, d, j, t, st);",
     "stencil_of(c, g, G[d] + (long)i * g->ng, d, j, t, st);", PINS_WRF),
    # weno_Re_flux: lower- and upper-face values of the gradients exchanged
    ("wrf_face_values_swapped", "orc_weno_cell(c->weno_order, c->weno_eps, cf, st, &L[at], &R[at]);",
     "orc_weno_cell(c->weno_order, c->weno_eps, cf, st, &R[at], &L[at]);", PINS_WRF),
]


def _build(tmp, name, text=None, repl=None):
    src = open(SRC).read()
    if text is not None:
        assert src.count(text) == 1, f"mutation site of {name} is not unique"
        src = src.replace(text, repl)
    d = os.path.join(tmp, name)
    os.makedirs(d)
    shutil.copy(os.path.join(ROOT, "oracle", "mfc_oracle.h"), d)
    with open(os.path.join(d, "mfc_oracle.c"), "w") as f:
        f.write(src)
    lib = os.path.join(d, "libmfc_oracle.so")
    subprocess.run(["gcc", "-O1", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-std=c99", "-shared", "-o", lib,
                    os.path.join(d, "mfc_oracle.c"), "-lm"], check=True, capture_output=True)
    return lib


def _run_pins(lib, pins):
    env = dict(os.environ, MFC_ORACLE_LIB=lib)
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                        *[os.path.join(ROOT, "tests", p) for p in pins]],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=300)
    return r.returncode, r.stdout[-600:]


@pytest.fixture(scope="module")
def tmp(tmp_path_factory):
    return str(tmp_path_factory.mktemp("oracle_mutants"))


def test_unmodified_oracle_passes_the_pins(tmp):
    rc, out = _run_pins(_build(tmp, "unmodified"), PINS_3D + PINS_WRF)
    assert rc == 0, out


@pytest.mark.parametrize("name,text,repl,pins", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_planted_slip_fails_the_pins(tmp, name, text, repl, pins):
    rc, out = _run_pins(_build(tmp, name, text, repl), pins)
    assert rc == 1, f"slip {name} was not caught (rc {rc}):\n{out}"
