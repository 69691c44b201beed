"""The 3D shock-bubble cloud at its benchmark size: 768^3 cells (453 M),
64 air bubbles in water, 300 TVD-RK3 steps at CFL 0.3, in the three forms;
the FP32 bulk and the learned weights compared with plain FP64 on the same
grid, and all three with a 2x refinement of a 384^3 sub-box.  The network
is trained as in test_gpu_cloud.py (3D pairs from a 64^3 plain FP64 run of
another cloud).

The sub-box, x in [0, 0.5], y in [346/768, +0.5], z in [92/768, +0.5], is
aligned with the 768^3 cells.  Its refinement is a plain FP64 run of the
same cloud on the sub-box alone (768^3 cells of 1/1536, 600 steps of dt/2)
with extrapolation boundaries, restricted 2x2x2.  Within the time window
the boundaries see only states the shock has not changed, or the planar
shock itself, except near the y and z faces, where a bubble hit at the end
of the window is cut; the comparison leaves out a band of 38 coarse cells
(about 0.05) along those faces.

A study rather than a check: about twelve GPU-minutes and 150 GB of host
memory, so it runs only with MFC_STUDIES=1 in the environment."""
import os

import numpy as np
import pytest

from helpers import say
from microfc_amd import PREC_FP32_BULK, PREC_FP32_LEARNED, ic, learned

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(os.environ.get("MFC_STUDIES") != "1",
                                 reason="study of about ten GPU-minutes: set MFC_STUDIES=1")]

NAMES = ("a1rho1", "a2rho2", "rho*u", "rho*v", "rho*w", "E", "a1", "a2")


def _rel_l1(a, b):
    """Per-variable L1 of a - b over the L1 of b (momenta over the summed
    momentum magnitude), one variable at a time."""
    mom = sum(float(np.sum(np.abs(b[v]))) for v in (2, 3, 4))
    out = []
    for v in range(a.shape[0]):
        den = mom if v in (2, 3, 4) else float(np.sum(np.abs(b[v])))
        out.append(float(np.sum(np.abs(a[v] - b[v]))) / den)
    return np.array(out)


Y0, Z0, HALF = 346, 92, 384  # sub-box origin and size in 768^3 cells
BAND = 38


def test_cloud_768_forms_against_fp64_and_refined_subbox(gpu):
    torch = gpu
    fprm, fg, fq = ic.bubble_cloud_3d(N=64, seed=1)
    data = learned.training_pairs(fprm, fg, fq, nsnap=4, steps_between=8, cfl=0.3)
    mlp = learned.Mlp(seed=0)
    learned.train(mlp, data, epochs=30, batch=8192, lr=3e-3, seed=0)
    del data, fq
    torch.cuda.empty_cache()
    say("[cloud 768] network trained on 64^3 pairs (seed 1)")

    N, nsteps, cfl = 768, 300, 0.3
    say(f"[cloud 768] writing the {N}^3 initial condition")
    prm, grids, q = ic.bubble_cloud_3d(N=N)
    say(f"[cloud 768] initial condition {prm.ghost_shape}")
    # one dt for all forms: the FP64 CFL reduction at t = 0
    from microfc_amd import RhsModule
    mod = RhsModule(prm, grids)
    try:
        dt = mod.cfl_dt(torch.from_numpy(q).cuda(), cfl)
    finally:
        mod.finalize()
    torch.cuda.empty_cache()
    t_end = nsteps * dt
    res = {}
    for prec, name in ((0, "plain FP64"), (PREC_FP32_BULK, "FP32 bulk"), (PREC_FP32_LEARNED, "learned")):
        prm.precision = prec
        say(f"  {name}: running")
        out, n, sec = learned.run(prm, grids, q, t_end, dt=dt, progress=lambda t: say(f"    step {t}"),
                                  progress_every=50)
        torch.cuda.empty_cache()
        assert n == nsteps
        finite = bool(np.all(np.isfinite(out)))
        say(f"  {name:11s} {N}^3, {n} steps to t = {t_end:.3e}, GPU s {sec:.2f} "
            f"({N ** 3 * n / sec / 1e9:.2f} G cell-updates/s), finite {finite}")
        assert finite, name
        res[prec] = out
    e64 = res[0]
    for prec, name in ((PREC_FP32_BULK, "FP32 bulk"), (PREC_FP32_LEARNED, "learned")):
        err = _rel_l1(res[prec], e64)
        say(f"  {name:11s} rel L1 vs FP64 {N}^3 {NAMES}: {np.array2string(err, precision=2)}")
    assert _rel_l1(res[PREC_FP32_BULK], e64).max() <= 1e-4
    # the sub-box of each coarse run, then the refined sub-box run
    sub = {k: np.ascontiguousarray(v[:, Z0:Z0 + HALF, Y0:Y0 + HALF, 0:HALF]) for k, v in res.items()}
    del res, e64, q, out
    lo = (0.0, Y0 / N, Z0 / N)
    hi = (HALF / N, (Y0 + HALF) / N, (Z0 + HALF) / N)
    say(f"[cloud 768] writing the refined sub-box {lo} .. {hi} at {2 * HALF}^3")
    fprm, fgrids, fq = ic.bubble_cloud_3d(N=2 * HALF, lo=lo, hi=hi)
    say("  refined sub-box plain FP64: running")
    fine, fn, fsec = learned.run(fprm, fgrids, fq, t_end, dt=dt / 2, progress=lambda t: say(f"    step {t}"),
                                 progress_every=100)
    del fq
    assert fn == 2 * nsteps and bool(np.all(np.isfinite(fine)))
    say(f"  refined sub-box {2 * HALF}^3 (dx 1/{2 * N}), {fn} steps, GPU s {fsec:.2f}")
    judge = learned.restrict_host(fine)
    del fine
    inner = np.s_[:, BAND:-BAND, BAND:-BAND, :]
    ref = judge[inner]
    for prec, name in ((0, "plain FP64"), (PREC_FP32_BULK, "FP32 bulk"), (PREC_FP32_LEARNED, "learned")):
        err = _rel_l1(sub[prec][inner], ref)
        say(f"  {name:11s} rel L1 vs the refined sub-box {NAMES}: {np.array2string(err, precision=3)}")
