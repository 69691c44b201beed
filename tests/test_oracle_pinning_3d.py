"""Pins the oracle's 3D generalisation (the reference is 1D/2D only, so the z
direction of oracle/mfc_oracle.c has no reference text to follow; it is the
dimension-split extension of m_rhs.fpp:469-656 and :686-908).

Each check is chosen so that a plausible slip in the z code fails it:

  ghost fill         explicit index expectations for every BC code in z,
                     composed after the x and y fills, with the z-momentum
                     (and only it) negated across a reflective z wall;
  z-uniform field    a field that does not vary along z, with w = 0, gives
                     the 2D RHS in every z plane and no z-momentum change,
                     for each z BC code (non-zero v at reflective y walls, so
                     a z fill that negates the wrong momentum shows);
  axis permutation   relabelling the axes (cyclic, and x <-> z), with the
                     momenta, grids, extents and BC codes permuted alike,
                     permutes the RHS, on grids stretched differently in each
                     direction with non-zero u, v, w at every wall: the z
                     sweep must use the z grid, the z momentum and the z BCs
                     exactly as the x and y sweeps use theirs;
  conservation       with periodic boundaries on a stretched grid the volume
                     integral of the mass, momentum and energy RHS vanishes
                     to round-off (the z flux difference must be divided by
                     the z cell width of the same cell).
"""
import numpy as np
import pytest

from microfc_amd import ic
from oracle import oracle as O

from helpers import smooth_prim_3d, stretched

GW, PIW = ic.WATER
GA, PIA = ic.AIR


def two_fluid(m, n, p, bcb, bce, weno_order=5):
    return O.Case(m=m, n=n, p=p, num_fluids=2, weno_order=weno_order, gammas=[GW, GA], pi_infs=[PIW, PIA],
                  bc_beg=tuple(bcb), bc_end=tuple(bce))


def scaled_err(a, b, mom):
    """Per-variable max |a - b| over a scale: the largest momentum component
    for the momenta (indices in mom), max |b_v| otherwise."""
    ms = max(np.max(np.abs(b[v])) for v in mom)
    out = []
    for v in range(b.shape[0]):
        den = ms if v in mom else np.max(np.abs(b[v]))
        out.append(np.max(np.abs(a[v] - b[v])) / den if den > 0 else np.max(np.abs(a[v] - b[v])))
    return np.array(out)


# --------------------------------------------------------------------------
# ghost fill


def _src(bcb, bce, s, i):
    """(source index, sign) of the fill along one direction."""
    if 0 <= i <= s:
        return i, 1
    if i < 0:
        return (0, 1) if bcb <= -3 else ((-i - 1, -1) if bcb == -2 else (s + 1 + i, 1))
    return (s, 1) if bce <= -3 else ((2 * s + 1 - i, -1) if bce == -2 else (i - s - 1, 1))


@pytest.mark.parametrize("bcs", [
    ((-1, -1), (-2, -2), (-3, -4)),
    ((-2, -5), (-3, -3), (-2, -2)),
    ((-3, -6), (-2, -4), (-1, -1)),
    ((-4, -2), (-1, -1), (-2, -3)),
])
def test_bc_fill_index_semantics_3d(bcs):
    c = two_fluid(9, 8, 7, [x[0] for x in bcs], [x[1] for x in bcs])
    b = c.buff_size
    q = np.zeros(c.ghost_shape)
    K, J, I = np.meshgrid(np.arange(c.p + 1), np.arange(c.n + 1), np.arange(c.m + 1), indexing="ij")

    def base(v, i, j, k):
        return 100000.0 * v + i + 32.0 * j + 1024.0 * k + 1.0

    for v in range(c.sys_size):
        q[v, b:b + c.p + 1, b:b + c.n + 1, b:b + c.m + 1] = base(v, I, J, K)
    O.bc_fill(c, q)
    momx = c.num_fluids
    got_ghosts = 0
    for kk in range(-b, c.p + b + 1):
        sk, fz = _src(*bcs[2], c.p, kk)
        for jj in range(-b, c.n + b + 1):
            sj, fy = _src(*bcs[1], c.n, jj)
            for ii in range(-b, c.m + b + 1):
                if 0 <= ii <= c.m and 0 <= jj <= c.n and 0 <= kk <= c.p:
                    continue
                si, fx = _src(*bcs[0], c.m, ii)
                for v in range(c.sys_size):
                    e = base(v, si, sj, sk)
                    if (v == momx and fx < 0) or (v == momx + 1 and fy < 0) or (v == momx + 2 and fz < 0):
                        e = -e
                    assert q[v, kk + b, jj + b, ii + b] == e, (v, ii, jj, kk)
                got_ghosts += 1
    assert got_ghosts == np.prod(c.ghost_shape[1:]) - (c.m + 1) * (c.n + 1) * (c.p + 1)


# --------------------------------------------------------------------------
# a z-uniform 3D field is the 2D field


@pytest.mark.parametrize("zbc", [(-1, -1), (-2, -2), (-3, -3), (-4, -4), (-2, -5)])
def test_z_uniform_3d_equals_2d(zbc):
    m, n, p = 23, 19, 7  # 8 z cells: exact z cell boundaries
    bcb, bce = (-3, -2, zbc[0]), (-4, -2, zbc[1])
    c3 = two_fluid(m, n, p, bcb, bce)
    c2 = two_fluid(m, n, 0, bcb[:2] + (-3,), bce[:2] + (-3,))
    q3 = smooth_prim_3d(c3, seed=4, w_zero=True)
    b = c3.buff_size
    # make the 3D field z-uniform: every plane is the k = 0 plane
    q3[:, b:b + p + 1] = q3[:, b:b + 1]
    # v must be non-zero at the reflective y walls for this check to see a
    # z fill that negates the y momentum
    prim = O.cons_to_prim(c3, q3)
    assert np.min(np.abs(prim[3, b:b + p + 1, b, b:b + m + 1])) > 1.0
    assert np.min(np.abs(prim[3, b:b + p + 1, b + n, b:b + m + 1])) > 1.0
    q2 = np.zeros(c2.ghost_shape)
    sel = [0, 1, 2, 3, 5, 6, 7]  # the 3D variables without rho*w
    q2[:, 0] = q3[sel, b]
    # x and y stretched; z uniform with exact cell boundaries, so every z face
    # of a z-uniform field has the same bits (on a stretched z grid the
    # per-cell weights would differ in the last bit)
    g2 = stretched(c2, (0.4, 0.3), (0.3, 1.1))
    g3 = g2 + O.uniform_grid(c3)[2:]
    r2 = O.compute_rhs(c2, g2, q2)
    r3 = O.compute_rhs(c3, g3, q3)
    tol = 1e-13 if -4 in zbc else 0.0
    for k in range(p + 1):
        err = scaled_err(r3[sel, k], r2[:, 0], mom=(2, 3))
        assert err.max() <= tol, (k, err)
        # the z-momentum equation sees only round-off
        assert np.max(np.abs(r3[4, k])) <= tol * np.max(np.abs(r2[2:4])), k


# --------------------------------------------------------------------------
# axis permutation


def permute_case(c, perm):
    """perm[d] = the old direction that becomes new direction d."""
    ext = [c.m, c.n, c.p]
    return two_fluid(ext[perm[0]], ext[perm[1]], ext[perm[2]], [c.bc_beg[perm[d]] for d in range(3)],
                     [c.bc_end[perm[d]] for d in range(3)], weno_order=c.weno_order)


def permute_field(c, q, perm):
    """Field arrays are (S, z, y, x); variable order [a1r1, a2r2, ru, rv, rw, E, a1, a2]."""
    nf = c.num_fluids
    old_axis = {0: 3, 1: 2, 2: 1}  # direction -> array axis
    # new array axes (z', y', x') = old axes of directions perm[2], perm[1], perm[0]
    out = np.transpose(q, (0, old_axis[perm[2]], old_axis[perm[1]], old_axis[perm[0]]))
    order = list(range(nf)) + [nf + perm[d] for d in range(3)] + list(range(nf + 3, c.sys_size))
    return np.ascontiguousarray(out[order])


@pytest.mark.parametrize("perm", [(1, 2, 0), (2, 0, 1), (2, 1, 0), (0, 2, 1)])
@pytest.mark.parametrize("weno_order", [5, 3])
def test_axis_permutation_3d(perm, weno_order):
    # distinct extents, grids and BC codes per direction; non-zero u, v, w at every wall
    c = two_fluid(17, 14, 11, (-2, -1, -4), (-3, -1, -2), weno_order=weno_order)
    q = smooth_prim_3d(c, seed=7)
    b = c.buff_size
    prim = O.cons_to_prim(c, q)
    for d in range(3):
        assert np.min(np.abs(prim[2 + d, b:-b, b:-b, b:-b])) > 0.5
    g = stretched(c, (0.5, 0.3, 0.4), (0.2, 1.3, 2.9))
    r = O.compute_rhs(c, g, q)
    cp = permute_case(c, perm)
    qp = permute_field(c, q, perm)
    gp = [g[perm[d]] for d in range(3)]
    rp = O.compute_rhs(cp, gp, qp)
    want = permute_field(c, r, perm)
    # the directions' contributions are summed in another order: round-off only
    err = scaled_err(rp, want, mom=(2, 3, 4))
    assert err.max() <= 1e-12, err


# --------------------------------------------------------------------------
# conservation


@pytest.mark.parametrize("weno_order", [5, 3])
def test_periodic_conservation_3d_stretched(weno_order):
    c = two_fluid(15, 13, 11, (-1, -1, -1), (-1, -1, -1), weno_order=weno_order)
    q = smooth_prim_3d(c, seed=2)
    g = stretched(c, (0.5, 0.3, 0.4), (0.0, 0.0, 0.0))
    r = O.compute_rhs(c, g, q)
    b = c.buff_size
    dx, dy, dz = (g[d][1][b:b + s + 1] for d, s in enumerate(c.extents()))
    vol = dz[:, None, None] * dy[None, :, None] * dx[None, None, :]
    for v in range(6):  # two partial densities, three momenta, energy
        tot = np.sum(r[v] * vol)
        mag = np.sum(np.abs(r[v]) * vol)
        assert abs(tot) <= 1e-13 * mag, (v, tot, mag)

