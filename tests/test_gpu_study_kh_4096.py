"""The Kelvin-Helmholtz shear layer at its benchmark sizes: learned WENO5
weights on a 4096^2 grid, trained on 2x-restricted pairs from an 8192^2
plain FP64 run, compared with the plain method at 4096^2 against the
8192^2 plain FP64 run (the reference solution) restricted to 4096^2.

A study rather than a check: a few GPU-minutes, so it runs only with
MFC_STUDIES=1 in the environment."""
import os

import numpy as np
import pytest

from helpers import say
from microfc_amd import PREC_FP32_BULK, PREC_FP32_LEARNED, ic, learned

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(os.environ.get("MFC_STUDIES") != "1",
                                 reason="study of a few GPU-minutes: set MFC_STUDIES=1")]

NAMES = ("rho", "rho*u", "rho*v", "E")


def test_kh_4096_learned_against_8192_plain(gpu):
    torch = gpu
    Nc, t_end = 4096, 0.05
    # pairs from a plain FP64 8192^2 run of another seed, six snapshots
    fprm, fg, fq = ic.shear_layer_2d(N=2 * Nc, seed=0)
    data = learned.training_pairs(fprm, fg, fq, nsnap=6, steps_between=150, max_per_snapshot=200_000)
    del fq
    mlp = learned.Mlp(seed=0)
    _, loss0 = mlp.gradient(data)
    steps = learned.train(mlp, data, epochs=30, batch=8192, lr=3e-3, seed=0)
    _, loss1 = mlp.gradient(data)
    say(f"[kh 4096] {data.shape[1]} pairs from 8192^2 (seed 0); loss {loss0:.3e} -> {loss1:.3e} "
        f"({len(steps)} Adam steps)")
    assert loss1 < loss0
    del data
    torch.cuda.empty_cache()

    seed = 5
    jprm, jg, jq = ic.shear_layer_2d(N=2 * Nc, seed=seed)
    fine, fsteps, fsec = learned.run(jprm, jg, jq, t_end)
    del jq
    assert np.all(np.isfinite(fine))
    ref = learned.restrict_host(fine)
    del fine
    say(f"  reference: plain FP64 8192^2, {fsteps} steps to t = {t_end}, GPU s {fsec:.2f}")
    cprm, cg, cq = ic.shear_layer_2d(N=Nc, seed=seed)
    rows = {}
    for prec, name in ((0, "plain FP64"), (PREC_FP32_BULK, "plain FP32"), (PREC_FP32_LEARNED, "learned")):
        cprm.precision = prec
        out, nsteps, sec = learned.run(cprm, cg, cq, t_end)
        assert np.all(np.isfinite(out)), name
        rows[name] = (np.array([learned.l1_error(out, ref, v) for v in range(4)]), sec)
        say(f"  {name:11s} 4096^2, {nsteps} steps, GPU s {sec:.2f}; L1 vs reference {NAMES}: "
            f"{' '.join(f'{x:.3e}' for x in rows[name][0])}")
    gain = rows["plain FP32"][0] / rows["learned"][0]
    say(f"  learned gain over plain FP32 per variable: {' '.join(f'{x:.3f}x' for x in gain)}")
