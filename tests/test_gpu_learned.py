"""The learned-weight form in the sweeps: the untrained model reproduces the
FP32-bulk plain path, conservation still holds, the per-face fallback keeps
volume fractions in range, and -- the gain test -- a network trained on the
GPU from plain-solver coarse/fine pairs lowers the coarse-grid error against
a 2x-finer plain run (the judge) on a case it was not trained on."""
import numpy as np
import pytest

from microfc_amd import PREC_FP32_BULK, PREC_FP32_LEARNED, RhsModule, ic, learned

from helpers import perturbed, say, scaled_linf

pytestmark = pytest.mark.gpu


def _rhs(torch, prm, grids, q):
    mod = RhsModule(prm, grids)
    try:
        qd = torch.from_numpy(q.copy()).cuda()
        rd = mod.new_field(ghosted=False)
        mod.s_compute_rhs(qd, None, rd, 0)
        mod.sync()
        return rd.cpu().numpy()
    finally:
        mod.finalize()


def test_untrained_learned_equals_plain_fp32(gpu):
    """With a zero output layer the learned weights are the plain ones (up to
    renormalisation rounding), so the learned form is as accurate as the
    plain FP32 bulk against FP64."""
    torch = gpu
    learned.Mlp()
    for name, (prm, grids, q) in {"bubble": ic.shock_bubble_2d(N=64), "shear": ic.shear_layer_2d(N=64),
                                  "sod": ic.sod_1d(m=199)}.items():
        q = perturbed(prm, q)
        ref = _rhs(torch, prm, grids, q)
        prm.precision = PREC_FP32_BULK
        plain = scaled_linf(prm, _rhs(torch, prm, grids, q), ref)
        prm.precision = PREC_FP32_LEARNED
        lrn = scaled_linf(prm, _rhs(torch, prm, grids, q), ref)
        say(f"  {name}: error vs FP64, plain FP32 bulk {plain.max():.2e}, untrained learned {lrn.max():.2e}")
        assert lrn.max() <= 2.0 * plain.max() + 1e-7


def test_learned_conserves_and_stays_bounded(gpu):
    torch = gpu
    rng = np.random.default_rng(1)
    p = learned.init_params(7)
    p[learned.OFF["W3"]:learned.OFF["W3"] + 512] = rng.normal(0, 0.5, 512).astype(np.float32)
    learned.Mlp(p)
    prm, grids, q = ic.shear_layer_2d(N=64)
    prm.bc_beg = (-1, -1, -1)
    prm.bc_end = (-1, -1, -1)
    grids = ic.uniform_grids(prm)
    prm.precision = PREC_FP32_LEARNED
    r = _rhs(torch, prm, grids, q)
    for v in range(4):
        tot, mag = abs(np.sum(r[v])), np.sum(np.abs(r[v]))
        assert tot <= 1e-13 * mag, v
    prm2, g2, q2 = ic.shock_bubble_2d(N=64)
    prm2.precision = PREC_FP32_LEARNED
    out, n, _ = learned.run(prm2, g2, q2, t_end=5e-6, cfl=0.4)
    a = out[prm2.num_fluids + prm2.num_dims + 1:]
    assert np.all(np.isfinite(out)) and a.min() >= 0.0 and a.max() <= 1.0


def test_trained_weights_beat_plain_against_finer_judge(gpu):
    torch = gpu
    Nc = 128
    # training pairs from a plain fine run (seed 0), shear layer resolved by
    # about one coarse cell
    fprm, fg, fq = ic.shear_layer_2d(N=2 * Nc, seed=0, delta=0.02)
    data = learned.training_pairs(fprm, fg, fq, nsnap=6, steps_between=40)
    say(f"  training pairs: {data.shape[1]}")
    mlp = learned.Mlp(seed=0)
    losses = learned.train(mlp, data, epochs=30, batch=8192, lr=2e-3, seed=0)
    say(f"  training loss {losses[0]:.4e} -> {losses[-1]:.4e} ({len(losses)} Adam steps on the GPU)")
    assert losses[-1] < losses[0]
    # evaluation on another seed: coarse plain vs coarse learned vs the 2x
    # finer plain run as judge
    t_end = 0.3
    fprm, fg, fq = ic.shear_layer_2d(N=2 * Nc, seed=5, delta=0.02)
    judge, nf, sf = learned.run(fprm, fg, fq, t_end)
    judge_c = learned.restrict_host(judge)
    cprm, cg, cq = ic.shear_layer_2d(N=Nc, seed=5, delta=0.02)
    cprm.precision = PREC_FP32_BULK
    plain, npl, sp = learned.run(cprm, cg, cq, t_end)
    cprm.precision = PREC_FP32_LEARNED
    lrn, nl, sl = learned.run(cprm, cg, cq, t_end)
    ep = learned.l1_error(plain, judge_c)
    el = learned.l1_error(lrn, judge_c)
    say(f"  L1(rho) vs 2x judge: plain {ep:.4e}, learned {el:.4e} (gain {ep / el:.3f}x); "
        f"GPU s: plain {sp:.3f}, learned {sl:.3f}, judge {sf:.3f}")
    assert np.all(np.isfinite(lrn))
    assert el < ep
