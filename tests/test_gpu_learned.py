"""The learned-weight form in the sweeps: with a network far from zero its
right-hand side matches the C oracle's learned right-hand side (FP64 with
BF16 at the matrix-core operands) to the FP32-bulk tolerance, the untrained
model reproduces the FP32-bulk plain path, conservation still holds, the per-face fallback keeps
volume fractions in range, and -- the gain test -- a network trained on the
GPU from plain-solver coarse/fine pairs lowers the coarse-grid error against
a 2x-finer plain run (the reference solution) on a case it was not trained on."""
import numpy as np
import pytest

from microfc_amd import PREC_FP32_BULK, PREC_FP32_LEARNED, RhsModule, ic, learned

from helpers import perturbed, say, scaled_linf, stretched_grids, to_oracle
from oracle import oracle as O
from test_mlp_oracle import random_params

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


def _cases():
    out = {"bubble": ic.shock_bubble_2d(N=64), "shear": ic.shear_layer_2d(N=64), "sod": ic.sod_1d(m=199),
           "droplet3d": ic.shock_droplet_3d(N=28), "cloud3d": ic.bubble_cloud_3d(N=32, nbubbles=16)}
    # non-uniform grids: per-cell coefficients in the learned sweeps
    for name, (prm, _, q) in (("bubble_stretched", ic.shock_bubble_2d(N=48)),
                              ("droplet3d_stretched", ic.shock_droplet_3d(N=28))):
        out[name] = (prm, stretched_grids(prm), q)
    return out


def test_learned_rhs_matches_oracle(gpu):
    """Per variable, the GPU learned RHS is within the FP32-bulk tolerance of
    the oracle's learned RHS, while the network moves that RHS far more than
    this away from the plain one (so a misplaced stencil, candidate or weight
    would not pass)."""
    torch = gpu
    params = random_params(11, scale=0.6, last=0.8).astype(np.float32)
    learned.Mlp(params)
    try:
        for name, (prm, grids, q) in _cases().items():
            q = perturbed(prm, q)
            case = to_oracle(prm)
            O.set_learned(None)
            plain = O.compute_rhs(case, grids, q)
            O.set_learned(params)
            ref = O.compute_rhs(case, grids, q)
            prm.precision = PREC_FP32_LEARNED
            got = _rhs(torch, prm, grids, q)
            err = scaled_linf(prm, got, ref)
            eff = scaled_linf(prm, plain, ref)
            say(f"  {name}: learned GPU vs learned oracle {err.max():.2e}; learned vs plain oracle {eff.max():.2e}")
            assert np.all(np.isfinite(got))
            assert err.max() <= 2e-3
            assert eff.max() >= 20.0 * err.max()
    finally:
        O.set_learned(None)


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


def test_trained_weights_beat_plain_against_finer_run(gpu):
    torch = gpu
    Nc = 128
    # training pairs from a plain fine run (seed 0), shear layer resolved by
    # about one coarse cell
    parts = []
    for seed in (0, 1, 2):
        fprm, fg, fq = ic.shear_layer_2d(N=2 * Nc, seed=seed, delta=0.02)
        parts.append(learned.training_pairs(fprm, fg, fq, nsnap=12, steps_between=20))
    data = torch.cat(parts, dim=1).contiguous()
    say(f"  training pairs: {data.shape[1]} (three 256^2 plain runs)")
    mlp = learned.Mlp(seed=0)
    _, loss0 = mlp.gradient(data)
    losses = learned.train(mlp, data, epochs=90, batch=8192, lr=3e-3, seed=0)
    _, loss1 = mlp.gradient(data)
    say(f"  loss over all pairs {loss0:.4e} -> {loss1:.4e} ({len(losses)} Adam steps on the GPU)")
    assert loss1 < 0.7 * loss0
    # evaluation on two other seeds: coarse plain vs coarse learned, both
    # against the 2x finer plain FP64 run (the reference solution),
    # restricted to the coarse grid
    names = ("rho", "rho*u", "rho*v", "E")
    for seed in (5, 7):
        t_end = 0.3
        fprm, fg, fq = ic.shear_layer_2d(N=2 * Nc, seed=seed, delta=0.02)
        ref_f, nf, sf = learned.run(fprm, fg, fq, t_end)
        ref_c = learned.restrict_host(ref_f)
        cprm, cg, cq = ic.shear_layer_2d(N=Nc, seed=seed, delta=0.02)
        res = {}
        for prec in (0, PREC_FP32_BULK, PREC_FP32_LEARNED):
            cprm.precision = prec
            out, nsteps, sec = learned.run(cprm, cg, cq, t_end)
            assert np.all(np.isfinite(out))
            res[prec] = (np.array([learned.l1_error(out, ref_c, v) for v in range(4)]), sec)
        e64, e32, el = res[0][0], res[PREC_FP32_BULK][0], res[PREC_FP32_LEARNED][0]
        say(f"  seed {seed}, t = {t_end}, {nsteps} steps on {Nc}^2; L1 vs the 2x finer run per variable {names}:")
        say(f"    plain FP64   {' '.join(f'{x:.3e}' for x in e64)}  GPU s {res[0][1]:.3f}")
        say(f"    plain FP32   {' '.join(f'{x:.3e}' for x in e32)}  GPU s {res[PREC_FP32_BULK][1]:.3f}")
        say(f"    learned      {' '.join(f'{x:.3e}' for x in el)}  GPU s {res[PREC_FP32_LEARNED][1]:.3f}")
        say(f"    gain over plain FP32 {' '.join(f'{x:.3f}x' for x in e32 / el)}; finer run GPU s {sf:.3f}")
        # every variable the network reconstructs is closer to the finer
        # run than plain FP32 (measured gains: density 1.10-1.15x,
        # streamwise momentum 1.03-1.04x, transverse momentum 1.9-2.3x,
        # energy 1.25-1.29x)
        for v in range(4):
            assert el[v] < e32[v], (seed, names[v])
    # matched-error view: everything against a 4x finer plain FP64 run,
    # compared on the coarse grid
    seed, t_end = 5, 0.3
    jprm, jg, jq = ic.shear_layer_2d(N=4 * Nc, seed=seed, delta=0.02)
    ref4, _, s4 = learned.run(jprm, jg, jq, t_end)
    ref = learned.restrict_host(learned.restrict_host(ref4))
    fprm, fg, fq = ic.shear_layer_2d(N=2 * Nc, seed=seed, delta=0.02)
    fine, _, s2 = learned.run(fprm, fg, fq, t_end)
    cprm, cg, cq = ic.shear_layer_2d(N=Nc, seed=seed, delta=0.02)
    cprm.precision = PREC_FP32_LEARNED
    lrn, _, sl = learned.run(cprm, cg, cq, t_end)
    cprm.precision = PREC_FP32_BULK
    pl, _, sp = learned.run(cprm, cg, cq, t_end)
    rows = [("plain FP32 128^2", pl, sp), ("learned 128^2", lrn, sl),
            ("plain FP64 256^2", learned.restrict_host(fine), s2)]
    say(f"  against the 512^2 plain FP64 run (GPU s {s4:.3f}), L1 on the 128^2 grid {names}:")
    for nm, x, sec in rows:
        say(f"    {nm:18s} {' '.join(f'{learned.l1_error(x, ref, v):.3e}' for v in range(4))}  GPU s {sec:.3f}")
