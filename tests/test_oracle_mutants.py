"""Mutation check of the oracle's 3D pins: plausible slips planted in a copy of
oracle/mfc_oracle.c (compiled into a temporary directory, never into the
tree) must each make tests/test_oracle_pinning_3d.py fail, and the unmodified
source built the same way must pass it.  This is what shows that the 3D
checks pin the z code and do not merely restate it.
"""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "mfc_oracle.c")

# (name, text in mfc_oracle.c, replacement)
MUTANTS = [
    # the z ghost fill negates the y momentum instead of the z momentum
    ("z_fill_negates_rho_v",
     "(flip && v == g.momxb + 2) ? -val : val;", "(flip && v == g.momxb + 1) ? -val : val;"),
    # the z ghost fill negates no momentum at a reflective wall
    ("z_fill_no_negation",
     "(flip && v == g.momxb + 2) ? -val : val;", "(0 && v == g.momxb + 2) ? -val : val;"),
    # the z ghost fill covers only the interior x-y rows (corners left unset)
    ("z_fill_interior_rows_only",
     "        for (j = -g.b; j <= g.n + g.b; j++)\n            for (i = -g.b; i <= g.m + g.b; i++)\n"
     "                for (v = 0; v < g.S; v++) {\n                    double val = q[v * g.ng + gidx(&g, i, j, src)];",
     "        for (j = 0; j <= g.n; j++)\n            for (i = 0; i <= g.m; i++)\n"
     "                for (v = 0; v < g.S; v++) {\n                    double val = q[v * g.ng + gidx(&g, i, j, src)];"),
    # z pencils divide by the y cell widths
    ("z_divides_by_dy",
     "coef, qc, qp, ds[dir],", "coef, qc, qp, ds[dir == 2 ? 1 : dir],"),
    # the z flux divergence overwrites the x and y contributions
    ("z_assigns_rhs", "            if (dir == 0)\n", "            if (dir == 0 || dir == 2)\n"),
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


def _run_pins(lib):
    env = dict(os.environ, MFC_ORACLE_LIB=lib)
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                        os.path.join(ROOT, "tests", "test_oracle_pinning_3d.py")],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=300)
    return r.returncode, r.stdout[-600:]


@pytest.fixture(scope="module")
def tmp(tmp_path_factory):
    return str(tmp_path_factory.mktemp("oracle_mutants"))


def test_unmodified_oracle_passes_the_3d_pins(tmp):
    rc, out = _run_pins(_build(tmp, "unmodified"))
    assert rc == 0, out


@pytest.mark.parametrize("name,text,repl", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_planted_slip_fails_the_3d_pins(tmp, name, text, repl):
    rc, out = _run_pins(_build(tmp, name, text, repl))
    assert rc == 1, f"slip {name} was not caught (rc {rc}):\n{out}"
