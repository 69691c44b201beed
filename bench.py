"""Throughput benchmark of the MI355X RHS + TVD-RK3 path (one GPU).

Workload (default): the 2D shock-bubble, two fluids (water + air bubble),
2048 x 2048 cells, WENO5 + HLLC + RK3, plain FP64, dt from the CFL 0.4
reduction at t = 0 and then held fixed (the reference's dt is an input).
A step is one full TVD-RK3 time step (three RHS evaluations with the stage
updates) over the whole grid; inputs are resident in HBM before timing.

Prints one JSON line with the throughput (cell-updates/s per RK3 step), the
roofline of the dominant kernel (HIP events on the library's stream) and the
CPU baseline (the FP64 oracle, one core, one RK3 step of the same grid).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6

# Measured HBM bytes (read + written) per launch of the dominant kernel, from
# rocprofv3 --pmc FETCH_SIZE and --pmc WRITE_SIZE runs of this script's
# default command, scaled by a 2 GiB device copy (FETCH_SIZE x2.0, WRITE_SIZE
# x1.0), and its VALU-busy fraction from SQ_ACTIVE_INST_VALU / GRBM_GUI_ACTIVE
# (profiles/hbm_traffic_shock_bubble_2d_2048_fp64.md,
# profiles/sweep_kernels_shock_bubble_2d_2048_fp64.md).  Other workloads
# report null.
PROFILED = {("shock_bubble_2d", 2048, "fp64", "sweep_y"): {"traffic": 678.5e6 + 236.6e6, "valu_busy": 0.79}}


def sweep_bytes_per_step(prm):
    """Algorithmic HBM bytes per RK3 step of each kernel, per cell: the
    stencil input is read once (S doubles), rhs is written by the first
    sweep, read and written by middle ones, read by the last, which reads the
    RK base (stages 2 and 3) and writes the stage output.  Viscous runs use
    the viscous kernels in place of the sweeps: per direction the face kernel
    reads the stencil input and writes a face record (HLLC flux, contact
    velocity, viscous flux: S+2+D doubles), the finish reads it and the
    volume fractions, and reads/writes rhs like a sweep (or, in the last
    direction, reads the RK base and stage input and writes the output)."""
    S, D, NF = prm.sys_size, prm.num_dims, prm.num_fluids
    out = {}
    names = ["sweep_x", "sweep_y", "sweep_z"][:D]
    if prm.viscous:
        total = 0
        for d in range(D):
            first, last = d == 0, d == D - 1
            per = S + 2 * (S + 2 + D) + NF + (0 if first else S) + (3 * S if last else S)
            total += 3 * 8 * per
        out["viscous"] = total
        return out
    for d, nm in enumerate(names):
        first, last = d == 0, d == D - 1
        per_stage = []
        for stage in range(3):
            r = S  # stencil input
            w = 0
            if not first:
                r += S  # rhs partial
            if last:
                if stage > 0:
                    r += S  # RK base q1
                w += S  # stage output
            else:
                w += S  # rhs
            per_stage.append(8 * (r + w))
        out[nm] = sum(per_stage)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--case", default="shock_bubble_2d",
                    choices=["shock_bubble_2d", "shock_droplet_3d", "shear_layer_2d", "bubble_cloud_3d", "sod_1d",
                             "viscous_shear_2d"])
    ap.add_argument("--N", type=int, default=0, help="cells per direction (default: the case's size)")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32_bulk", "fp32_learned"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-aux", action="store_true",
                    help="skip the auxiliary-kernel block (cons->prim, CFL reduction, ghost fill, RK1 update)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a 2 GiB device-to-device copy (a kernel of known traffic for counter scaling)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU: the multi-GPU halo exchange is designed but "
                                   "not built, so --gpus must be 1"}))
        sys.exit(2)

    import numpy as np
    import torch

    from microfc_amd import PREC_FP32_BULK, PREC_FP32_LEARNED, PREC_FP64, RhsModule, ic

    prec = {"fp64": PREC_FP64, "fp32_bulk": PREC_FP32_BULK, "fp32_learned": PREC_FP32_LEARNED}[args.precision]
    if prec == PREC_FP32_LEARNED:
        # throughput does not depend on the weights; a network with a non-zero
        # output layer exercises the softmax and the per-face fallback
        from microfc_amd import learned
        p = learned.init_params(0)
        p[learned.OFF["W3"]:learned.OFF["W3"] + 512] = np.random.default_rng(1).normal(0, 0.1, 512)
        learned.Mlp(p)
    if args.case == "shock_bubble_2d":
        N = args.N or 2048
        prm, grids, q = ic.shock_bubble_2d(N=N, precision=prec)
        workload = f"2D shock-bubble water/air {N}x{N}, WENO5+HLLC+RK3"
    elif args.case == "shock_droplet_3d":
        N = args.N or 256
        prm, grids, q = ic.shock_droplet_3d(N=N, precision=prec)
        workload = f"3D shock-droplet water/air {N}^3, WENO5+HLLC+RK3"
    elif args.case == "bubble_cloud_3d":
        N = args.N or 768
        prm, grids, q = ic.bubble_cloud_3d(N=N, precision=prec)
        workload = f"3D shock-bubble cloud water/air {N}^3 (64 air bubbles), WENO5+HLLC+RK3"
    elif args.case == "viscous_shear_2d":
        N = args.N or 2048
        prm, grids, q = ic.viscous_shear_2d(N=N)
        prm.precision = prec
        workload = (f"2D viscous shear layer (the reference's 2D_viscous case) {N}x{N}, two fluids, Re 1e-4, "
                    "WENO5+HLLC+viscous+RK3")
    elif args.case == "shear_layer_2d":
        N = args.N or 4096
        prm, grids, q = ic.shear_layer_2d(N=N, precision=prec)
        workload = f"2D shear layer (Kelvin-Helmholtz) {N}x{N}, one gas, WENO5+HLLC+RK3"
    else:
        N = args.N or 1000
        prm, grids, q = ic.sod_1d(m=N - 1)
        prm.precision = prec
        workload = f"1D Sod {N} cells, WENO5+HLLC+RK3"
    ncell = (prm.m + 1) * (prm.n + 1) * (prm.p + 1)

    if args.copy_calibration:
        src = torch.ones(1 << 28, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    mod = RhsModule(prm, grids)
    q1 = torch.from_numpy(q).cuda()
    q2 = mod.new_field()
    rhs = mod.new_field(ghosted=False)
    cfl = 0.4
    dt = mod.cfl_dt(q1, cfl)
    if prm.viscous:
        dt = min(dt, ic.viscous_dt(prm, grids, q))
    for t in range(args.warmup):
        mod.s_tvd_rk(t, (q1, q2), None, rhs, dt)
    mod.sync()
    torch.cuda.synchronize()

    mod.timing(enable=True, reset=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for t in range(args.steps):
        mod.s_tvd_rk(args.warmup + t, (q1, q2), None, rhs, dt)
    mod.sync()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    timing = mod.timing(enable=False)
    sec = t1 - t0
    ms_per_step = 1e3 * sec / args.steps
    value = ncell * args.steps / sec

    assert torch.isfinite(q1).all().item(), "non-finite state after the timed steps"

    # roofline of the dominant kernel
    per_step = sweep_bytes_per_step(prm)
    dom = max(per_step, key=lambda k: timing[k][0])
    dom_ms, dom_n = timing[dom]
    dom_bytes_per_launch = per_step[dom] * ncell / (timing[dom][1] / args.steps)
    dom_avg_s = dom_ms / 1e3 / dom_n
    achieved = dom_bytes_per_launch / dom_avg_s / 1e9
    prof = PROFILED.get((args.case, N, args.precision, dom), {})
    S, D = prm.sys_size, prm.num_dims
    step_bytes = 64 * S * ncell
    f_alg = 3 * (D * (85 * S + 180 + 4 * S + 10) + 20 + 4 * S)
    out = {
        "metric": "cell-updates/s per RK3 step",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {PREC_FP64: "f64", PREC_FP32_BULK: "f32 bulk / f64 accumulate",
                  PREC_FP32_LEARNED: "f32 bulk + bf16 MFMA weights / f64 accumulate"}[prec],
        "data": "synthetic initial condition of the named case (generated, not read)",
        "config": {"workload": workload, "cells": ncell, "sys_size": S, "cfl": cfl, "dt": dt,
                   "precision": args.precision, "time_stepper": "TVD-RK3"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": prof.get("traffic"),
                     "valu_busy_profiled": prof.get("valu_busy"),
                     "algorithmic_bytes_per_launch": dom_bytes_per_launch,
                     "avg_launch_ms": dom_avg_s * 1e3},
        "step_hbm": {"algorithmic_bytes_per_cell_step": 64 * S, "achieved_GBps": step_bytes * args.steps / sec / 1e9,
                     "frac": step_bytes * args.steps / sec / 1e9 / HBM_PEAK_GBS},
        "fp64_valu": {"nominal_flops_per_cell_step": f_alg, "achieved_TFLOPs": f_alg * value / 1e12,
                      "peak_TFLOPs": FP64_PEAK_TFLOPS, "frac": f_alg * value / 1e12 / FP64_PEAK_TFLOPS},
        "kernels_ms_per_step": {k: v[0] / args.steps for k, v in timing.items() if v[1]},
    }
    # every kernel of the step against HBM peak on its algorithmic bytes
    kr = {}
    for k, bps in per_step.items():
        ms, n = timing[k]
        if n:
            per_launch = bps * ncell / (n / args.steps)
            gbs = per_launch / (ms / 1e3 / n) / 1e9
            kr[k] = {"avg_launch_ms": ms / n, "algorithmic_bytes_per_launch": per_launch, "achieved_GBps": gbs,
                     "frac": gbs / HBM_PEAK_GBS, "launches_per_step": n / args.steps}
    out["kernels_roofline"] = kr
    if not args.no_aux:
        out["aux_kernels"] = aux_kernels(mod, prm, grids, q1, dt, args, RhsModule, torch)
    else:
        mod.finalize()
    del q1, q2, rhs
    torch.cuda.empty_cache()

    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args, prm, grids, q, np)
    print(json.dumps(out))


def aux_kernels(mod, prm, grids, q1, dt, args, RhsModule, torch, calls=5):
    """The path's other kernels on the same state, outside the timed steps:
    average launch time (HIP events on the library's stream), algorithmic
    HBM bytes per launch and the achieved fraction of HBM peak.  Finalizes
    mod (the RK1 update needs a case initialised for time_stepper 1)."""
    import copy
    S = prm.sys_size
    ng = 1
    for e in prm.ghost_shape[1:]:
        ng *= e
    nint = (prm.m + 1) * (prm.n + 1) * (prm.p + 1)
    prim = mod.new_field()
    rhs = mod.new_field(ghosted=False)
    mod.sync()
    mod.timing(enable=True, reset=True)
    for _ in range(calls):
        mod.s_compute_rhs(q1, prim, rhs, 0)  # ghost fill, sweeps and the q_prim_vf output
        mod.icfl_max(prim, dt)
    mod.sync()
    t = mod.timing(enable=False)
    mod.finalize()
    prm1 = copy.deepcopy(prm)
    prm1.time_stepper = 1
    mod1 = RhsModule(prm1, grids)
    try:
        q2 = mod1.new_field()
        rhs1 = mod1.new_field(ghosted=False)
        mod1.sync()
        mod1.timing(enable=True, reset=True)
        for k in range(calls):
            mod1.s_tvd_rk(k, (q1, q2), None, rhs1, dt)
        mod1.sync()
        t1 = mod1.timing(enable=False)
    finally:
        mod1.finalize()
    # algorithmic bytes per launch: cons->prim reads and writes every ghosted
    # cell; the CFL pass reads the interior primitives; the ghost fill reads
    # one source and writes one ghost value per ghost cell; RK1 reads q and
    # rhs and writes q on the interior
    spec = {"cons_to_prim": (t, 16 * S * ng), "icfl": (t, 8 * S * nint), "bc_fill": (t, 16 * S * (ng - nint)),
            "rk1_update": (t1, 24 * S * nint)}
    out = {}
    for name, (tm, nbytes) in spec.items():
        ms, n = tm[name]
        if not n:
            continue
        avg = ms / n
        out[name] = {"avg_launch_ms": avg, "algorithmic_bytes": nbytes, "achieved_GBps": nbytes / avg / 1e6,
                     "frac": nbytes / avg / 1e6 / HBM_PEAK_GBS, "launches": n}
    return out


def cpu_baseline(args, prm, grids, q, np):
    """The FP64 oracle on the host: one TVD-RK3 step of the same grid (a
    bounded sample of the workload), same dt rule, on every host thread the
    run may use (OpenMP over pencils, OMP_NUM_THREADS) and on one core."""
    import ctypes

    from oracle import oracle as O
    oc = O.Case(m=prm.m, n=prm.n, p=prm.p, num_fluids=prm.num_fluids, weno_order=prm.weno_order,
                time_stepper=prm.time_stepper, weno_eps=prm.weno_eps, gammas=list(prm.gammas),
                pi_infs=list(prm.pi_infs), bc_beg=prm.bc_beg, bc_end=prm.bc_end,
                Re=[tuple(r) for r in prm.Re] if prm.Re else None)
    ncell = (prm.m + 1) * (prm.n + 1) * (prm.p + 1)
    if ncell > 20_000_000:
        return {"value": None, "unit": "cell-updates/s", "cores": None, "kind": "port",
                "sample": "skipped: grid too large for a bounded CPU sample"}
    L = O.lib()
    L.orc_set_threads.argtypes = [ctypes.c_int]
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)

    def one_step(n):
        L.orc_set_threads(n)
        q1 = q.copy()
        q2 = np.zeros_like(q1)
        t0 = time.perf_counter()
        O.tvd_rk_step(oc, grids, 1e-9, q1, q2, t_step=0)
        return time.perf_counter() - t0

    sec_all = one_step(threads)
    sec_one = one_step(1)
    return {"value": ncell / sec_all, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "sample": f"one TVD-RK3 step of the same {ncell}-cell grid on {threads} host threads "
                      f"({sec_all:.2f} s) and on one core ({sec_one:.1f} s)",
            "single_core": {"value": ncell / sec_one, "cores": 1}}


if __name__ == "__main__":
    main()
