"""GPU kernels of the learned WENO5 weights against the FP64 MLP oracle
(oracle/mlp.py, which rounds to BF16 at the MFMA operand points): gradients,
loss, inference, training progress, and the coarse/fine training pairs."""
import numpy as np
import pytest

from microfc_amd import ic, learned
from oracle import mlp as Q

from helpers import say
from test_mlp_oracle import random_params, random_samples

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def test_gradient_matches_oracle(gpu):
    torch = gpu
    P = random_params(11).astype(np.float32)
    S = random_samples(12, 3001).astype(np.float32)
    mlp = learned.Mlp(P)
    g, loss = mlp.gradient(torch.from_numpy(S).cuda())
    gr, lr = Q.gradient(P.astype(np.float64), S.astype(np.float64), bf16_points=True)
    say(f"  loss gpu {loss:.6e} oracle {lr:.6e}")
    assert loss == pytest.approx(lr, rel=1e-4)
    for name, (a, b) in {"W1": (0, 512), "b1": (512, 544), "W2": (544, 1568), "b2": (1568, 1600),
                         "W3": (1600, 2112), "b3": (2112, 2128)}.items():
        e = _rel(g[a:b], gr[a:b])
        say(f"  grad {name}: rel L2 error vs oracle {e:.2e}")
        assert e < 2e-2, name


def test_inference_matches_oracle(gpu):
    torch = gpu
    P = random_params(13).astype(np.float32)
    S = random_samples(14, 1000).astype(np.float32)
    mlp = learned.Mlp(P)
    out = mlp.infer(torch.from_numpy(S).cuda()).cpu().numpy()
    ref = Q.infer(P.astype(np.float64), S.astype(np.float64), bf16_points=True)
    err = np.max(np.abs(out - ref))
    say(f"  inference max abs error vs oracle {err:.2e}")
    assert err < 5e-3
    assert np.allclose(out[:3].sum(axis=0), 1.0, atol=1e-5)


def test_untrained_model_gives_plain_weights(gpu):
    torch = gpu
    S = random_samples(15, 500).astype(np.float32)
    mlp = learned.Mlp()
    out = mlp.infer(torch.from_numpy(S).cuda()).cpu().numpy()
    np.testing.assert_allclose(out[:3], S[8:11], rtol=2e-6, atol=1e-7)


def test_training_reduces_loss(gpu):
    torch = gpu
    # targets from a teacher network the student can represent
    teacher = random_params(16, last=1.0)
    S = random_samples(17, 20000)
    S[11] = Q.forward(teacher, S, bf16_points=False)["y"]
    data = torch.from_numpy(S.astype(np.float32)).cuda()
    mlp = learned.Mlp(seed=3)
    losses = learned.train(mlp, data, epochs=40, batch=2048, lr=3e-3)
    say(f"  loss {losses[0]:.3e} -> {losses[-1]:.3e} over {len(losses)} Adam steps")
    assert losses[-1] < 0.2 * losses[0]


def test_training_pairs_match_oracle_1d(gpu):
    torch = gpu
    prm, grids, q = ic.sod_1d(m=199)
    cprm = learned.coarse_params(prm)
    rng = np.random.default_rng(0)
    fine = np.zeros(prm.m + 1 + 8)
    fine[4:-4] = np.cumsum(rng.normal(size=prm.m + 1)) + 0.1 * rng.normal(size=prm.m + 1)
    fine[:4] = fine[4]
    fine[-4:] = fine[-5]
    coarse = np.zeros(cprm.m + 1 + 8)
    coarse[4:-4] = 0.5 * (fine[4:-4:2] + fine[5:-4:2])
    coarse[:4] = coarse[4]
    coarse[-4:] = coarse[-5]
    cb = ic.uniform_grids(cprm)[0][0]
    coef = O_coef(cb, cprm)
    got = learned.samples(0, cprm.m, 0, 4, torch.from_numpy(coarse).cuda(), torch.from_numpy(fine).cuda(), coef,
                          1e-16).cpu().numpy()
    ref = Q.samples_1d(coarse, fine, coef)
    np.testing.assert_allclose(got, ref.astype(np.float32), rtol=1e-6, atol=1e-6)
    # restriction kernel
    fc = torch.from_numpy(fine).cuda()
    cc = torch.zeros(cprm.m + 1 + 8, dtype=torch.float64, device="cuda")
    learned.restrict(cprm.m, 0, 4, fc, cc)
    np.testing.assert_array_equal(cc.cpu().numpy()[4:-4], coarse[4:-4])


@pytest.mark.parametrize("d", [0, 1, 2])
def test_training_pairs_3d_reduce_to_1d(gpu, d):
    """A 3D field that varies only along direction d: each of its samples in
    that direction equals the oracle's 1D sample of the profile (the four
    fine faces that tile a coarse face carry one value), and its 3D
    restriction is the 1D one, bit for bit."""
    torch = gpu
    prm, _, _ = ic.sod_1d(m=63)
    cprm = learned.coarse_params(prm)
    rng = np.random.default_rng(1 + d)
    prof = np.cumsum(rng.normal(size=64)) + 0.1 * rng.normal(size=64)
    cprof = 0.5 * (prof[0::2] + prof[1::2])
    fine1, coarse1 = np.pad(prof, 4, mode="edge"), np.pad(cprof, 4, mode="edge")
    coef = O_coef(ic.uniform_grids(cprm)[0][0], cprm)
    ref = Q.samples_1d(coarse1, fine1, coef)

    def field(profile, nt):
        shape = [nt, nt, nt]
        shape[2 - d] = len(profile)
        ax = [1, 1, 1]
        ax[2 - d] = len(profile)
        a = np.broadcast_to(profile.reshape(ax), shape)
        return np.ascontiguousarray(np.pad(a, 4, mode="edge"))

    fine3, coarse3 = field(prof, 4), field(cprof, 2)
    ext = [1, 1, 1]
    ext[d] = 31
    cm, cn, cp = ext
    got = learned.samples(d, cm, cn, 4, torch.from_numpy(coarse3).cuda(), torch.from_numpy(fine3).cuda(), coef,
                          1e-16, p=cp).cpu().numpy()
    got = got.reshape(13, cp + 1, cn + 1, cm + 1, 2)
    for a in range(2):
        for b2 in range(2):
            sel = [(slice(None), a, b2, slice(None)), (slice(None), a, slice(None), b2),
                   (slice(None), slice(None), a, b2)][d]
            line = got[sel].reshape(13, -1)
            np.testing.assert_allclose(line, ref.astype(np.float32), rtol=1e-6, atol=1e-6)
    cc = torch.zeros(coarse3.shape, dtype=torch.float64, device="cuda")
    learned.restrict(cm, cn, 4, torch.from_numpy(fine3).cuda(), cc, p=cp)
    np.testing.assert_array_equal(cc.cpu().numpy()[4:-4, 4:-4, 4:-4], coarse3[4:-4, 4:-4, 4:-4])


def O_coef(cb, cprm):
    from oracle import oracle as O
    return O.weno_coefficients(cb, cprm.m, cprm.buff_size, 5)[:, cprm.buff_size]
