"""Pins the MLP oracle (oracle/mlp.py): its exact-arithmetic gradient
against central finite differences, the zero-last-layer identity (learned
weights equal the plain ones), and BF16 rounding against known values."""
import numpy as np
import pytest

from oracle import mlp as Q


def random_params(seed, scale=0.5, last=0.3):
    rng = np.random.default_rng(seed)
    W1 = np.zeros((16, 32))
    W1[:5] = rng.normal(0, scale, (5, 32))
    W2 = rng.normal(0, scale, (32, 32))
    W3 = np.zeros((32, 16))
    W3[:, :3] = rng.normal(0, last, (32, 3))
    b3 = np.zeros(16)
    b3[:3] = rng.normal(0, 0.1, 3)
    return Q.pack(W1, rng.normal(0, 0.1, 32), W2, rng.normal(0, 0.1, 32), W3, b3)


def random_samples(seed, n):
    rng = np.random.default_rng(seed)
    S = np.zeros((12, n))
    S[0:5] = rng.uniform(-1, 1, (5, n))
    S[2] = 0.0
    S[5:8] = rng.uniform(-1, 1, (3, n))
    w = rng.uniform(0.05, 1, (3, n))
    S[8:11] = w / w.sum(axis=0)
    S[11] = rng.uniform(-1, 1, n)
    return S


def test_bf16_rounding():
    assert Q.bf16(1.0) == 1.0
    ulp = 2.0 ** -7                                # bf16 spacing at 1.0
    assert Q.bf16(1.0 + 0.5 * ulp) == 1.0          # tie -> even (down)
    assert Q.bf16(1.0 + 1.5 * ulp) == 1.0 + 2 * ulp  # tie -> even (up)
    assert Q.bf16(1.0 + 0.75 * ulp) == 1.0 + ulp
    assert Q.bf16(-3.140625) == -3.140625


def test_exact_gradient_matches_finite_differences():
    P = random_params(1)
    S = random_samples(2, 64)
    g, _ = Q.gradient(P, S, bf16_points=False)
    rng = np.random.default_rng(3)
    idx = list(rng.choice(np.arange(0, 5 * 32), 6)) + list(Q.OFF["b1"] + rng.choice(32, 3)) + \
        list(Q.OFF["W2"] + rng.choice(1024, 6)) + list(Q.OFF["b2"] + rng.choice(32, 3)) + \
        [Q.OFF["W3"] + 16 * r + c for r, c in zip(rng.choice(32, 6), rng.choice(3, 6))] + \
        [Q.OFF["b3"] + 0, Q.OFF["b3"] + 2]
    h = 1e-6
    for i in idx:
        Pp, Pm = P.copy(), P.copy()
        Pp[i] += h
        Pm[i] -= h
        fd = (Q.forward(Pp, S, False)["loss"] - Q.forward(Pm, S, False)["loss"]) / (2 * h)
        assert g[i] == pytest.approx(fd, rel=1e-5, abs=1e-9), i


def test_zero_last_layer_gives_plain_weights():
    P = random_params(4, last=0.0)
    P[Q.OFF["b3"]:Q.OFF["b3"] + 16] = 0.0
    S = random_samples(5, 50)
    f = Q.forward(P, S, bf16_points=True)
    np.testing.assert_allclose(f["w"], S[8:11].T, rtol=1e-14)
    np.testing.assert_allclose(f["y"], (S[5:8] * S[8:11]).sum(axis=0), rtol=1e-13)


def test_padded_parameters_get_zero_gradient():
    P = random_params(6)
    S = random_samples(7, 100)
    g, _ = Q.gradient(P, S, bf16_points=True)
    W1, b1, W2, b2, W3, b3 = Q.unpack(g)
    assert np.all(W1[5:] == 0) and np.all(W3[:, 3:] == 0) and np.all(b3[3:] == 0)


def test_keep_selects_valid_in_hull_targets():
    """Training pairs are kept when the stencil's variation is real (flag in
    row 12) and the fine-grid target lies inside the candidates' hull."""
    import torch
    from microfc_amd import learned
    S = torch.zeros((13, 5))
    S[5:8] = torch.tensor([[-1.0] * 5, [0.0] * 5, [1.0] * 5])
    S[11] = torch.tensor([0.5, -1.0, 1.5, 0.0, 0.2])
    S[12] = torch.tensor([1.0, 1.0, 1.0, 1.0, 0.0])
    assert learned.keep(S).tolist() == [True, True, False, True, False]
