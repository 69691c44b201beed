"""Initial conditions and grids from a case dictionary, as the reference's
pre-processor builds them (src/pre_process/), written in the files the
simulation reads.

A case is the dictionary a reference ``examples/*/case.py`` prints: the
&user_inputs keys of src/pre_process/m_start_up.fpp:87-95 (m, n, the
domain, grid stretching, num_fluids, fluid_pp, num_patches and
patch_icpp(i)%...).  ``PreProcess(case)`` lays it out:

* the grid (m_grid.f90:54-140): uniform cell boundaries, then the optional
  tanh-type stretching of each direction applied loops_x / loops_y times;
* the patches in order (m_initial_condition.fpp:37-105), each geometry
  restated from m_create_patches.fpp: 1 line segment, 2 circle, 3 rectangle,
  4 sweep line, 5 ellipse, 6 isentropic vortex, 7 2D analytical, 15 1D
  analytical, 18 varcircle; a cell takes a patch where its geometry covers
  it and the patch may overwrite the cell's current patch
  (alter_patch(patch_id_fp)), or, for smoothed patches, where the cell
  belongs to the smoothing patch; the patch's primitive state is blended
  with the cell's by the pseudo volume fraction eta
  (m_assign_patches.fpp:54-165);
* the primitive -> conservative conversion (m_variables_conversion.fpp:
  385-454).

Each cell's arithmetic follows the reference's statements in order; numpy
evaluates them elementwise over the cells a patch covers.  ``write`` stores
the result as the serial files p_all/p0/0/ or, with parallel_io, the
restart_data/lustre_* files.
"""
from __future__ import annotations

import math

import numpy as np

from . import case as C
from .ic import ghosted, prim_to_cons
from .rhs import Params, populate_grid_buffers

DFLT_REAL = -1e6
DFLT_INT = -100


def _logical(v) -> bool:
    if isinstance(v, str):
        return v.strip().strip("'").upper().startswith("T")
    return bool(v)


class PreProcess:
    """One case laid out on its grid (2D when n > 0, else 1D)."""

    def __init__(self, case: dict):
        self.case = dict(case)
        g = self.case.get
        self.m = int(g("m"))
        self.n = int(g("n", 0))
        self.num_fluids = int(g("num_fluids"))
        self.num_patches = int(g("num_patches"))
        self.parallel_io = _logical(g("parallel_io", "F"))
        nf = self.num_fluids
        self.gammas = [float(g(f"fluid_pp({i + 1})%gamma")) for i in range(nf)]
        self.pi_infs = [float(g(f"fluid_pp({i + 1})%pi_inf")) for i in range(nf)]
        self.D = 1 + (self.n > 0)
        self.S = 2 * nf + self.D + 1
        self.E = nf + self.D
        self.adv = self.E + 1
        self.x_cb, self.x_cc, self.dx = self._grid("x", self.m)
        if self.n > 0:
            self.y_cb, self.y_cc, self.dy = self._grid("y", self.n)
        self.prim = None
        self.patch_id_fp = None

    # -- grid: m_grid.f90:54-140 --------------------------------------------
    def _grid(self, d: str, s: int):
        g = self.case.get
        beg, end = float(g(f"{d}_domain%beg")), float(g(f"{d}_domain%end"))
        dx = (end - beg) / float(s + 1)
        i = np.arange(0, s + 1, dtype=np.float64)
        cc = beg + 5e-1 * dx * (2.0 * i + 1.0)
        cb = np.empty(s + 2)  # x_cb(-1:s)
        cb[:s + 1] = beg + dx * i
        cb[s + 1] = end
        if _logical(g(f"stretch_{d}", "F")):
            a = float(g(f"a_{d}"))
            xa, xb = float(g(f"{d}_a")), float(g(f"{d}_b"))
            loops = int(g(f"loops_{d}", 1))
            length = abs(cb[s + 1] - cb[0])
            cb = cb / length
            xa, xb = xa / length, xb / length
            for _ in range(loops):
                cb = cb / a * (a + np.log(np.cosh(a * (cb - xa))) + np.log(np.cosh(a * (cb - xb)))
                               - 2.0 * math.log(math.cosh(a * (xb - xa) / 2.0)))
            cb = cb * length
            cc = (cb[1:] + cb[:-1]) / 2.0
            dx = float(np.min(cb[1:] - cb[:-1]))
        return cb, cc, dx

    # -- patches ------------------------------------------------------------
    def _patch(self, i: int, key: str, default=None):
        return self.case.get(f"patch_icpp({i})%{key}", default)

    def _alter(self, i: int, owner):
        """alter_patch(patch_id_fp) of patch i for each cell's current owner
        (alter_patch(0) is true by default, the others false)."""
        allowed = np.zeros(self.num_patches + 1, dtype=bool)
        allowed[0] = True
        for j in range(self.num_patches + 1):
            v = self._patch(i, f"alter_patch({j})")
            if v is not None:
                allowed[j] = _logical(v)
        return allowed[owner]

    def _assign(self, i: int, mask, eta):
        """s_assign_patch_species_primitive_variables for the cells in mask,
        eta a scalar or an array over the whole grid.  A patch value the case
        does not set is dflt_real, as in the reference's defaults
        (src/pre_process/m_global_parameters.fpp:185-215)."""
        nf, E, adv = self.num_fluids, self.E, self.adv
        eta = np.broadcast_to(np.asarray(eta, dtype=np.float64), mask.shape)[mask]
        orig = self.prim[:, mask].copy()
        new = self.prim[:, mask]
        for f in range(nf):
            new[f] = eta * float(self._patch(i, f"alpha_rho({f + 1})", DFLT_REAL)) + (1.0 - eta) * orig[f]
        for f in range(nf):
            new[adv + f] = eta * float(self._patch(i, f"alpha({f + 1})", DFLT_REAL)) + (1.0 - eta) * orig[adv + f]
        for d in range(self.D):
            new[nf + d] = eta * float(self._patch(i, f"vel({d + 1})", DFLT_REAL)) + (1.0 - eta) * orig[nf + d]
        new[E] = eta * float(self._patch(i, "pres", DFLT_REAL)) + (1.0 - eta) * orig[E]
        self.prim[:, mask] = new
        owner = self.patch_id_fp[mask]
        owner[1.0 - eta < 1e-16] = i
        self.patch_id_fp[mask] = owner

    def _smoothing(self, i: int):
        smoothen = _logical(self._patch(i, "smoothen", "F"))
        sid = int(self._patch(i, "smooth_patch_id", i))
        coeff = float(self._patch(i, "smooth_coeff", DFLT_REAL))
        return smoothen, sid, coeff

    def _line_segment(self, i, analytical=False):
        xc, lx = float(self._patch(i, "x_centroid")), float(self._patch(i, "length_x"))
        beg, end = xc - 0.5 * lx, xc + 0.5 * lx
        X = self.x_cc[None, :]
        mask = (beg <= X) & (end >= X) & self._alter(i, self.patch_id_fp)
        self._assign(i, mask, 1.0)
        if analytical:
            xb = np.broadcast_to(self.x_cb[1:][None, :], mask.shape)[mask]  # x_cb(i)
            self.prim[self.E][mask] = self.prim[self.E][mask] * \
                (1.0 + 0.2 * np.exp(-1.0 * ((xb - xc) ** 2.0) / (2.0 * 0.005)))

    def _circle(self, i):
        xc, yc = float(self._patch(i, "x_centroid")), float(self._patch(i, "y_centroid"))
        r = float(self._patch(i, "radius"))
        smoothen, sid, coeff = self._smoothing(i)
        X, Y = self.x_cc[None, :], self.y_cc[:, None]
        d2 = (X - xc) * (X - xc) + (Y - yc) * (Y - yc)
        eta = 1.0
        if smoothen:
            eta = np.tanh(coeff / min(self.dx, self.dy) * (np.sqrt(d2) - r)) * (-0.5) + 0.5
        mask = ((d2 <= r * r) & self._alter(i, self.patch_id_fp)) | (self.patch_id_fp == sid)
        self._assign(i, mask, eta)

    def _varcircle(self, i):
        xc, yc = float(self._patch(i, "x_centroid")), float(self._patch(i, "y_centroid"))
        r, th = float(self._patch(i, "radius")), float(self._patch(i, "epsilon"))
        X, Y = self.x_cc[None, :], self.y_cc[:, None]
        myr = np.sqrt((X - xc) * (X - xc) + (Y - yc) * (Y - yc))
        mask = (myr <= r + th / 2.0) & (myr >= r - th / 2.0) & self._alter(i, self.patch_id_fp)
        self._assign(i, mask, 1.0)

    def _ellipse(self, i):
        xc, yc = float(self._patch(i, "x_centroid")), float(self._patch(i, "y_centroid"))
        a, b = float(self._patch(i, "radii(1)")), float(self._patch(i, "radii(2)"))
        smoothen, sid, coeff = self._smoothing(i)
        X, Y = self.x_cc[None, :], self.y_cc[:, None]
        ex, ey = (X - xc) / a, (Y - yc) / b
        e2 = ex * ex + ey * ey
        eta = 1.0
        if smoothen:
            eta = np.tanh(coeff / min(self.dx, self.dy) * (np.sqrt(e2) - 1.0)) * (-0.5) + 0.5
        mask = ((e2 <= 1.0) & self._alter(i, self.patch_id_fp)) | (self.patch_id_fp == sid)
        self._assign(i, mask, eta)

    def _rectangle(self, i, analytical=False):
        xc, yc = float(self._patch(i, "x_centroid")), float(self._patch(i, "y_centroid"))
        lx, ly = float(self._patch(i, "length_x")), float(self._patch(i, "length_y"))
        xb, xe = xc - 0.5 * lx, xc + 0.5 * lx
        yb, ye = yc - 0.5 * ly, yc + 0.5 * ly
        X, Y = self.x_cc[None, :], self.y_cc[:, None]
        mask = (xb <= X) & (xe >= X) & (yb <= Y) & (ye >= Y) & self._alter(i, self.patch_id_fp)
        self._assign(i, mask, 1.0)
        if analytical:
            XB = np.broadcast_to(self.x_cb[1:][None, :], mask.shape)[mask]
            YB = np.broadcast_to(self.y_cb[1:][:, None], mask.shape)[mask]
            self.prim[self.E][mask] = self.prim[self.E][mask] * \
                (1.0 + 0.2 * np.exp(-1.0 * ((XB - xc) ** 2.0 + (YB - yc) ** 2.0) / (2.0 * 0.005)))

    def _sweep_line(self, i):
        xc, yc = float(self._patch(i, "x_centroid")), float(self._patch(i, "y_centroid"))
        a, b = float(self._patch(i, "normal(1)")), float(self._patch(i, "normal(2)"))
        c = -a * xc - b * yc
        smoothen, sid, coeff = self._smoothing(i)
        X, Y = self.x_cc[None, :], self.y_cc[:, None]
        lin = a * X + b * Y + c
        eta = 1.0
        if smoothen:
            eta = 5e-1 + 5e-1 * np.tanh(coeff / min(self.dx, self.dy) * lin / math.sqrt(a * a + b * b))
        mask = ((lin >= 0.0) & self._alter(i, self.patch_id_fp)) | (self.patch_id_fp == sid)
        self._assign(i, mask, eta)

    def _isentropic_vortex(self, i):
        xc, yc = float(self._patch(i, "x_centroid")), float(self._patch(i, "y_centroid"))
        r = float(self._patch(i, "radius"))
        X, Y = self.x_cc[None, :], self.y_cc[:, None]
        mask = ((X - xc) * (X - xc) + (Y - yc) * (Y - yc) <= r * r) & self._alter(i, self.patch_id_fp)
        self._assign(i, mask, 1.0)

    def initial_condition(self):
        """s_generate_initial_condition: the primitive field (S, n+1, m+1)
        after every patch, and the conservative field from it."""
        shape = (self.n + 1, self.m + 1)
        self.prim = np.full((self.S,) + shape, DFLT_REAL)
        self.patch_id_fp = np.zeros(shape, dtype=np.int64)
        for i in range(1, self.num_patches + 1):
            geo = int(self._patch(i, "geometry", DFLT_INT))
            if self.n > 0:
                {2: self._circle, 3: self._rectangle, 4: self._sweep_line, 5: self._ellipse,
                 6: self._isentropic_vortex, 7: lambda k: self._rectangle(k, analytical=True),
                 18: self._varcircle}.get(geo, lambda k: None)(i)
            else:
                {1: self._line_segment, 15: lambda k: self._line_segment(k, analytical=True)}.get(
                    geo, lambda k: None)(i)
        return self.prim, prim_to_cons(self.params(), self.prim[:, None])[:, 0]

    # -- outputs ------------------------------------------------------------
    def params(self) -> Params:
        """The simulation's Params of this case (boundary conditions from
        the case; the viscous keys the simulation would read are included)."""
        g = self.case.get
        bcb = tuple(int(g(f"bc_{d}%beg", -3)) for d in "xyz")
        bce = tuple(int(g(f"bc_{d}%end", -3)) for d in "xyz")
        Re = []
        for f in range(self.num_fluids):
            Re.append(tuple(float(g(f"fluid_pp({f + 1})%Re({r + 1})", -1.0)) for r in range(2)))
        prm = Params(m=self.m, n=self.n, num_fluids=self.num_fluids, weno_order=int(g("weno_order", 5)),
                     gammas=self.gammas, pi_infs=self.pi_infs, bc_beg=bcb, bc_end=bce,
                     Re=Re if any(r > 0 for pair in Re for r in pair) else None,
                     weno_Re_flux=_logical(g("weno_Re_flux", "F")))
        return prm

    def grids(self):
        """Ghosted (s_cb, ds) per direction for the simulation's Params."""
        prm = self.params()
        out = [populate_grid_buffers(self.x_cb, prm.buff_size, prm.bc_beg[0], prm.bc_end[0])]
        if self.n > 0:
            out.append(populate_grid_buffers(self.y_cb, prm.buff_size, prm.bc_beg[1], prm.bc_end[1]))
        return out

    def state(self):
        """(Params, grids, ghosted conservative field) for the simulation."""
        prm = self.params()
        if self.prim is None:
            self.initial_condition()
        cons = prim_to_cons(prm, self.prim[:, None])
        return prm, self.grids(), ghosted(prm, cons)

    def write(self, case_dir: str, t_step: int = 0):
        """The files src/pre_process/m_data_output.f90 writes: serial
        p_all/p0/<t_step>/ or, with parallel_io, restart_data/lustre_*."""
        prm, grids, q = self.state()
        if self.parallel_io:
            C.write_parallel_grid(case_dir, prm, grids)
            C.write_parallel(case_dir, t_step, prm, q)
        else:
            C.write_serial(case_dir, t_step, prm, grids, q)
        return prm, grids, q
