"""Pins the learned-weight reconstruction of the C oracle
(orc_weno5_learned_cell, used by its compute_rhs while weights are set):

* logits forced by the last bias select single candidates, which must be the
  textbook third-order polynomials of each face (this fixes the stencil
  orientation of both faces, independently of the code under test);
* equal logits give the plain WENO5 value, and so do any logits where the
  plain weights are ideal (the gate on their distance from ideal closes) or
  where the stencil's variation is rounding noise;
* the per-face fallback keeps the plain value where the learned one breaks a
  bound the plain one keeps;
* the C forward pass agrees with the independent numpy MLP oracle;
* a zero output layer leaves the whole right-hand side at its plain value.
"""
import numpy as np
import pytest

from oracle import mlp as Q
from oracle import oracle as O

from test_mlp_oracle import random_params

B5 = 4


def _coef(s=40):
    cb, _ = O.uniform_grid(O.Case(m=s))[0]
    return O.weno_coefficients(cb, s, B5, 5)[:, B5 + 10]


@pytest.fixture
def learned():
    yield O.set_learned
    O.set_learned(None)


def _bias_only(b3):
    p = np.zeros(Q.NPARAM, dtype=np.float32)
    p[Q.OFF["b3"]:Q.OFF["b3"] + 3] = b3
    return p


def test_forced_logits_pick_textbook_candidates(learned):
    cf = _coef()
    v = np.array([0.3, -0.2, 0.7, 1.1, 0.4])
    # face j+1/2 from cells (j, j+1, j+2), (j-1, j, j+1), (j-2, j-1, j)
    right = [(2 * v[2] + 5 * v[3] - v[4]) / 6, (-v[1] + 5 * v[2] + 2 * v[3]) / 6,
             (2 * v[0] - 7 * v[1] + 11 * v[2]) / 6]
    # face j-1/2, the mirror images
    left = [(2 * v[2] + 5 * v[1] - v[0]) / 6, (-v[3] + 5 * v[2] + 2 * v[1]) / 6,
            (2 * v[4] - 7 * v[3] + 11 * v[2]) / 6]
    for k in range(3):
        b3 = np.zeros(3)
        b3[k] = 80.0
        learned(_bias_only(b3))
        vl, vr = O.weno5_learned_cell(1e-16, cf, v)
        assert vr == pytest.approx(right[k], rel=1e-12), k
        assert vl == pytest.approx(left[k], rel=1e-12), k


def test_gate_closes_at_ideal_weights(learned):
    """A huge weno_eps swamps the smoothness indicators, so the plain weights
    are the ideal ones to ~1e-10 and the gate (their distance from ideal over
    0.03) is ~1e-8: logits forced to one candidate then leave the plain
    WENO5 value, the linear fifth-order one, although the candidates differ.
    With the gate open (rough data, weno_eps 1e-16) the same logits select
    single candidates (test_forced_logits_pick_textbook_candidates)."""
    cf = _coef()
    v = np.array([0.3, -0.2, 0.7, 1.1, 0.4])
    linear_r = (2 * v[0] - 13 * v[1] + 47 * v[2] + 27 * v[3] - 3 * v[4]) / 60
    for b3 in ([80.0, 0.0, -80.0], [0.0, 0.0, 80.0]):
        learned(_bias_only(b3))
        pl, pr = O.weno_cell(5, 1e10, cf, v)
        ll, lr = O.weno5_learned_cell(1e10, cf, v)
        assert pr == pytest.approx(linear_r, rel=1e-8)
        assert lr == pytest.approx(pr, rel=1e-6) and ll == pytest.approx(pl, rel=1e-6)


def test_rounding_noise_keeps_plain_weno5(learned):
    """A stencil whose variation is below 1e-6 of its magnitude (the rule
    that admits training pairs) keeps the plain weights whatever the logits,
    while the same shape at a resolvable amplitude does not."""
    cf = _coef()
    learned(_bias_only([80.0, 0.0, -80.0]))
    shape = np.array([0.3, -0.2, 0.7, 1.1, 0.4])
    for amp, plain_expected in ((1e-8, True), (1e-2, False)):
        v = 1.0 + amp * shape
        pl, pr = O.weno_cell(5, 1e-30, cf, v)
        ll, lr = O.weno5_learned_cell(1e-30, cf, v)
        same = lr == pytest.approx(pr, rel=1e-15, abs=0) and ll == pytest.approx(pl, rel=1e-15, abs=0)
        assert same == plain_expected, amp


def test_equal_logits_give_plain_weno5(learned):
    cf = _coef()
    rng = np.random.default_rng(3)
    for c in (0.0, 2.5, -7.0):
        learned(_bias_only([c, c, c]))
        for _ in range(50):
            v = rng.normal(0, 1, 5)
            pl, pr = O.weno_cell(5, 1e-16, cf, v)
            ll, lr = O.weno5_learned_cell(1e-16, cf, v)
            assert ll == pytest.approx(pl, rel=1e-13, abs=1e-15)
            assert lr == pytest.approx(pr, rel=1e-13, abs=1e-15)


def test_fallback_keeps_bounds(learned):
    cf = _coef()
    # the cell-(j-2..j) candidate of face j+1/2 undershoots below zero, the
    # plain value stays positive
    v = np.array([0.0, 0.5, 0.2, 0.05, 0.0])
    learned(_bias_only([0.0, 0.0, 30.0]))
    _, pr = O.weno_cell(5, 1e-16, cf, v)
    _, free = O.weno5_learned_cell(1e-16, cf, v, kind=0)
    _, kept = O.weno5_learned_cell(1e-16, cf, v, kind=1)
    assert pr > 0 and free < 0
    assert free == pytest.approx((2 * v[0] - 7 * v[1] + 11 * v[2]) / 6, rel=1e-9)
    assert kept == pr
    # upper bound of volume fractions: mirror the profile about 1/2
    w = 1.0 - v
    _, pr = O.weno_cell(5, 1e-16, cf, w)
    _, free = O.weno5_learned_cell(1e-16, cf, w, kind=1)
    _, kept = O.weno5_learned_cell(1e-16, cf, w, kind=3)
    assert pr < 1 and free > 1 and kept == pr


def test_c_forward_matches_numpy_oracle(learned):
    cf = _coef()
    params = random_params(11, scale=0.6, last=0.8).astype(np.float32)
    learned(params)
    rng = np.random.default_rng(4)
    n = 200
    V = rng.normal(0, 1, (n, 5))
    S = np.zeros((12, n))
    want = np.zeros(n)
    for i, v in enumerate(V):
        s = np.max(np.abs(np.diff(v)))
        p = [v[2] + cf[6] * (v[4] - v[3]) + cf[7] * (v[3] - v[2]),
             v[2] + cf[8] * (v[3] - v[2]) + cf[9] * (v[2] - v[1]),
             v[2] + cf[10] * (v[2] - v[1]) + cf[11] * (v[1] - v[0])]
        d = [v[4] - v[3], v[3] - v[2], v[2] - v[1], v[1] - v[0]]
        b = [cf[18] * d[0] ** 2 + cf[19] * d[0] * d[1] + cf[20] * d[1] ** 2 + 1e-16,
             cf[21] * d[1] ** 2 + cf[22] * d[1] * d[2] + cf[23] * d[2] ** 2 + 1e-16,
             cf[24] * d[2] ** 2 + cf[25] * d[2] * d[3] + cf[26] * d[3] ** 2 + 1e-16]
        a = np.array([cf[15 + k] / b[k] ** 2 for k in range(3)])
        v32 = v.astype(np.float32)
        s32 = np.max(np.abs(np.diff(v32)))
        S[0:5, i] = (v32 - v32[2]) * (np.float32(1.0) / s32)  # FP32 features
        S[5:8, i] = (np.array(p) - v[2]) / s
        S[8:11, i] = a / a.sum()
        want[i] = O.weno5_learned_cell(1e-16, cf, v)[1]
    f = Q.forward(params.astype(np.float64), S, bf16_points=True)
    got = V[:, 2] + np.max(np.abs(np.diff(V, axis=1)), axis=1) * f["y"]
    np.testing.assert_allclose(want, got, rtol=1e-10, atol=1e-12)
    # and the network does move the values away from plain WENO5
    plain = np.array([O.weno_cell(5, 1e-16, cf, v)[1] for v in V])
    assert np.max(np.abs(want - plain)) > 1e-2


def test_zero_output_layer_keeps_plain_rhs(learned):
    from microfc_amd import ic
    from helpers import scaled_linf, to_oracle
    prm, grids, q = ic.shock_bubble_2d(N=32)
    case = to_oracle(prm)
    r0 = O.compute_rhs(case, grids, q)
    p = random_params(2).astype(np.float32)
    p[Q.OFF["W3"]:Q.OFF["W3"] + 512] = 0.0
    p[Q.OFF["b3"]:Q.OFF["b3"] + 16] = 0.0
    learned(p)
    r1 = O.compute_rhs(case, grids, q)
    assert scaled_linf(prm, r1, r0).max() < 1e-12
