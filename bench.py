"""Throughput benchmark of the MI355X RHS + TVD-RK3 path (one GPU).

Workload (default): the 3D shock-droplet, two fluids (a water droplet of
radius 0.1 in air hit by a Mach 2.4 air shock), 512^3 cells, WENO5 + HLLC +
TVD-RK3 in FP64 -- the contracted FP64 form (FMA contraction, reciprocal-
multiply division; within 1e-12 relative of the FP64 oracle per step), with
the plain, bit-exact FP64 form timed beside it.  dt comes from the CFL 0.4
reduction at t = 0 and is then held fixed (the reference's dt is an input).
A step is one full TVD-RK3 time step (three RHS evaluations with the stage
updates) over the whole grid; inputs are resident in HBM before timing.

Prints one JSON line: throughput (cell-updates/s per RK3 step), the timed
form's relative L1 distance from the plain FP64 form after the same steps
(the "matched L1 error vs FP64" of the metric; any form but the plain one),
the roofline
of the dominant kernel on the compulsory FP64 state traffic (64*S bytes per
cell per step, a third of it per stage and launch; launch times from HIP
events on the library's stream), every kernel's figures, and the CPU
baseline (the FP64 oracle on the host's cores, one RK3 step of a slab of the
same case).
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
# rocprofv3 --pmc FETCH_SIZE and --pmc WRITE_SIZE runs of this script with the
# same case and form, scaled by a 2 GiB device copy (FETCH_SIZE x2.0,
# WRITE_SIZE x1.0) and averaged over the kernel's launches in the timed steps,
# and its VALU-busy fraction (SQ_ACTIVE_INST_VALU * 4 / (1024 CU-SIMDs *
# GRBM_GUI_ACTIVE / 8)), with the profile summary they come from.  Other
# workloads report null.
_P512 = "profiles/sweep_kernels_shock_droplet_3d_512_fp64_fast.md"
_P2048 = "profiles/sweep_kernels_shock_bubble_2d_2048_fp64_fast.md"
PROFILED = {
    ("shock_droplet_3d", 512, "fp64_fast", "sweep_xy"): {
        "traffic": 18.89e9 + 8.60e9, "valu_busy": 0.80,
        "limiter": "VALU issue (FP64), not HBM (27.5 GB at 3.1 TB/s)", "profile": _P512},
    ("shock_droplet_3d", 512, "fp64_fast", "sweep_z"): {
        "traffic": 13.86e9 + 8.59e9, "valu_busy": 0.69,
        "limiter": "VALU issue in RK stage 1 (80 % busy); with 6 RK-base variables in stages 2 and 3, 24.8 GB at "
                   "4.2 TB/s", "profile": _P512},
    ("shock_bubble_2d", 2048, "fp64_fast", "sweep_xy"): {
        "traffic": 453.1e6 + 236.0e6, "valu_busy": 0.67,
        "limiter": "VALU issue and latency (63-74 % busy, four waves per SIMD), not HBM", "profile": _P2048},
}

DEFAULT_PRECISION = "fp64_fast"
# RK-base variables the contracted form's 3D z sweep adds (the y sweep, or
# the fused x and y sweep, adds the rest; the library's split)
BASE_ON_Z = 4
BASE_ON_Z_FUSED = 6
FORMS = {"fp64": "plain FP64: the reference's evaluation order, no FMA contraction, IEEE division (bit-exact "
                 "against the FP64 oracle)",
         "fp64_fast": "contracted FP64: FMA contraction, reciprocal-multiply division, WENO weights normalised "
                      "by one reciprocal (within 1e-12 relative of the FP64 oracle per step)",
         "fp32_bulk": "FP32 WENO/HLLC, FP64 cons->prim, flux differences and RK stages",
         "fp32_learned": "FP32 bulk with learned WENO5 weights (BF16 MFMA)"}


def sweep_order(prm):
    """The library's sweep order: x, y, z as the reference adds them; the
    contracted FP64, FP32-bulk and learned forms of an inviscid case run them
    in reverse, so that the RK stage's output is written by the x sweep."""
    names = ["sweep_x", "sweep_y", "sweep_z"][:prm.num_dims]
    rev = prm.precision in (1, 2, 3) and not prm.viscous
    return names[::-1] if rev else names


def sweep_bytes_per_step(prm, fused=False):
    """Algorithmic HBM bytes per RK3 step of each kernel, per cell: the
    stencil input is read once (S doubles), rhs is written by the first
    sweep, read and written by middle ones, read by the last, which writes
    the stage output; the RK base (stages 2 and 3) is read by the last sweep,
    or, in the contracted FP64, FP32-bulk and learned forms of a 2D/3D case,
    which spread the stage over the sweeps, by the y sweep (in 3D its first
    BASE_ON_Z variables by the z sweep).  Viscous runs use the
    viscous kernels in place of the sweeps: per direction the face kernel
    reads the stencil input and writes a face record (HLLC flux, contact
    velocity, viscous flux: S+2+D doubles), the finish reads it and the
    volume fractions, and reads/writes rhs like a sweep (or, in the last
    direction, reads the RK base and stage input and writes the output).
    With weno_Re_flux the velocity-gradient planes and their face averages
    add their passes.  fused: the fused x and y sweep (sweep_xy) reads the
    stencil input once; in the contracted and FP32-bulk forms it comes after
    the z sweep in 3D, reads the z sweep's rhs and its part of the RK base
    and writes the stage output; in the plain form it comes first, writing
    rhs in 3D (the z sweep then carries the RK stage) and the stage output
    in 2D."""
    S, D, NF = prm.sys_size, prm.num_dims, prm.num_fluids
    out = {}
    names = sweep_order(prm)
    fold = D > 1 and prm.precision in (1, 2, 3)
    if prm.viscous:
        total = 0
        for d in range(D):
            first, last = d == 0, d == D - 1
            per = S + 2 * (S + 2 + D) + NF + (0 if first else S) + (3 * S if last else S)
            if prm.weno_Re_flux:
                # the gradient pass of direction d reads the densities and
                # momenta and writes D gradient planes; the face-average pass
                # reads the D*D planes and writes D*D face values, which the
                # face kernel reads
                per += (NF + 2 * D) + 3 * D * D
            total += 3 * 8 * per
        out["viscous"] = total
        return out
    if fused and prm.precision == 0:
        if D == 3:
            out["sweep_xy"] = 8 * 3 * 2 * S
            out["sweep_z"] = 8 * (3 * 3 * S + 2 * S)
        else:
            out["sweep_xy"] = 8 * (3 * 2 * S + 2 * S)
        return out
    if fused:
        zb = BASE_ON_Z_FUSED if D == 3 else 0
        if D == 3:
            out["sweep_z"] = 8 * (3 * 2 * S + 2 * zb)
        out["sweep_xy"] = 8 * (3 * (S + (S if D == 3 else 0) + S) + 2 * (S - zb))
        return out
    for d, nm in enumerate(names):
        first, last = d == 0, d == D - 1
        per_stage = []
        for stage in range(3):
            r = S  # stencil input
            w = 0
            if not first:
                r += S  # rhs partial
            if stage > 0 and fold:
                # RK base q1, split between z and y in the contracted 3D form
                r += {"sweep_z": BASE_ON_Z, "sweep_y": S - BASE_ON_Z}.get(nm, 0) if D == 3 else (
                    S if nm == "sweep_y" else 0)
            elif stage > 0 and last:
                r += S  # RK base q1
            if last:
                w += S  # stage output
            else:
                w += S  # rhs
            per_stage.append(8 * (r + w))
        out[nm] = sum(per_stage)
    return out


def make_case(args, prec, ic):
    """(params, grids, q, N, workload name) of the benchmark case."""
    if args.case == "shock_bubble_2d":
        N = args.N or 2048
        prm, grids, q = ic.shock_bubble_2d(N=N, precision=prec)
        workload = f"2D shock-bubble water/air {N}x{N}, WENO5+HLLC+RK3"
    elif args.case == "shock_droplet_3d":
        N = args.N or 512
        prm, grids, q = ic.shock_droplet_3d(N=N, precision=prec)
        workload = f"3D shock-droplet water/air {N}^3 (two fluids, Mach 2.4 air shock), WENO5+HLLC+RK3"
    elif args.case == "bubble_cloud_3d":
        N = args.N or 768
        prm, grids, q = ic.bubble_cloud_3d(N=N, precision=prec)
        workload = f"3D shock-bubble cloud water/air {N}^3 (64 air bubbles), WENO5+HLLC+RK3"
    elif args.case == "viscous_shear_2d":
        N = args.N or 2048
        prm, grids, q = ic.viscous_shear_2d(N=N, weno_Re_flux=not args.fd_gradients)
        prm.precision = prec
        grad = "finite-difference" if args.fd_gradients else "WENO-reconstructed (weno_Re_flux)"
        workload = (f"2D viscous shear layer (the reference's 2D_viscous case) {N}x{N}, two fluids, Re 1e-4, "
                    f"WENO5+HLLC+viscous ({grad} velocity gradients)+RK3")
    elif args.case == "shear_layer_2d":
        N = args.N or 4096
        prm, grids, q = ic.shear_layer_2d(N=N, precision=prec)
        workload = f"2D shear layer (Kelvin-Helmholtz) {N}x{N}, one gas, WENO5+HLLC+RK3"
    else:
        N = args.N or 1000
        prm, grids, q = ic.sod_1d(m=N - 1)
        prm.precision = prec
        workload = f"1D Sod {N} cells, WENO5+HLLC+RK3"
    return prm, grids, q, N, workload


def time_steps(RhsModule, torch, ic, prm, grids, q, args, dt=None):
    """Warm-up, then exactly args.steps timed TVD-RK3 steps with the state
    resident in HBM, bracketed by device synchronisations; per-kernel HIP
    event timing on the library's stream.  Returns (seconds, timing, dt,
    module) with the module still initialised."""
    mod = RhsModule(prm, grids)
    q1 = torch.from_numpy(q).cuda()
    q2 = mod.new_field()
    rhs = mod.new_field(ghosted=False)
    if dt is None:
        dt = mod.cfl_dt(q1, 0.4)
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
    sec = time.perf_counter() - t0
    timing = mod.timing(enable=False)
    assert torch.isfinite(q1).all().item(), "non-finite state after the timed steps"
    return sec, timing, dt, mod, (q1, q2, rhs)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--case", default="shock_droplet_3d",
                    choices=["shock_bubble_2d", "shock_droplet_3d", "shear_layer_2d", "bubble_cloud_3d", "sod_1d",
                             "viscous_shear_2d"])
    ap.add_argument("--N", type=int, default=0, help="cells per direction (default: the case's size)")
    ap.add_argument("--precision", default=None, choices=["fp64", "fp64_fast", "fp32_bulk", "fp32_learned"],
                    help=f"arithmetic form (default {DEFAULT_PRECISION}; fp64 for the viscous case, whose "
                         "terms run in the plain form only)")
    ap.add_argument("--fd-gradients", action="store_true",
                    help="viscous case: finite-difference velocity gradients instead of the example's "
                         "weno_Re_flux")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-aux", action="store_true",
                    help="skip the auxiliary-kernel block (cons->prim, CFL reduction, ghost fill, RK1 update)")
    ap.add_argument("--no-plain", action="store_true",
                    help="skip timing the plain (bit-exact) FP64 form beside the timed one")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a 2 GiB device-to-device copy (a kernel of known traffic for counter scaling)")
    args = ap.parse_args()
    if args.precision is None:
        args.precision = "fp64" if args.case == "viscous_shear_2d" else DEFAULT_PRECISION

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU: the multi-GPU halo exchange is designed but "
                                   "not built, so --gpus must be 1"}))
        sys.exit(2)

    import numpy as np
    import torch

    from microfc_amd import PREC_FP32_BULK, PREC_FP32_LEARNED, PREC_FP64, PREC_FP64_FAST, RhsModule, ic

    prec = {"fp64": PREC_FP64, "fp64_fast": PREC_FP64_FAST, "fp32_bulk": PREC_FP32_BULK,
            "fp32_learned": PREC_FP32_LEARNED}[args.precision]
    if prec == PREC_FP32_LEARNED:
        # throughput does not depend on the weights; a network with a non-zero
        # output layer exercises the softmax and the per-face fallback
        from microfc_amd import learned
        p = learned.init_params(0)
        p[learned.OFF["W3"]:learned.OFF["W3"] + 512] = np.random.default_rng(1).normal(0, 0.1, 512)
        learned.Mlp(p)
    prm, grids, q, N, workload = make_case(args, prec, ic)
    ncell = (prm.m + 1) * (prm.n + 1) * (prm.p + 1)
    print(f"[bench] {workload}, {args.precision}: {args.warmup} warm-up + {args.steps} timed steps", file=sys.stderr,
          flush=True)

    if args.copy_calibration:
        src = torch.ones(1 << 28, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    sec, timing, dt, mod, (q1, q2, rhs) = time_steps(RhsModule, torch, ic, prm, grids, q, args)
    compare = prec != PREC_FP64 and not args.no_plain
    final = prm.interior(q1).clone() if compare else None  # the state after the warm-up and timed steps
    ms_per_step = 1e3 * sec / args.steps
    value = ncell * args.steps / sec

    S, D = prm.sys_size, prm.num_dims
    # Algorithmic HBM bytes: 64*S per cell per RK3 step (the compulsory FP64
    # state traffic of the three stages: 16*S, 24*S, 24*S), i.e. 64*S/3 per
    # cell per stage.  Each kernel launch of a stage covers every cell once,
    # so a launch is charged one stage's bytes of all cells.
    step_bytes = 64 * S * ncell
    stage_bytes = step_bytes / 3
    per_step = sweep_bytes_per_step(prm, fused=bool(timing.get("sweep_xy", (0, 0))[1]))
    launched = {k: v for k, v in timing.items() if v[1] and k in per_step}
    dom = max(launched, key=lambda k: launched[k][0])
    dom_ms, dom_n = launched[dom]
    dom_avg_s = dom_ms / 1e3 / dom_n
    achieved = stage_bytes / dom_avg_s / 1e9
    prof = PROFILED.get((args.case, N, args.precision, dom), {})
    f_alg = 3 * (D * (85 * S + 180 + 4 * S + 10) + 20 + 4 * S)
    out = {
        "metric": "cell-updates/s per RK3 step at matched L1 error vs FP64; HBM GB/s % of peak",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {PREC_FP64: "f64", PREC_FP64_FAST: "f64", PREC_FP32_BULK: "f32 bulk / f64 accumulate",
                  PREC_FP32_LEARNED: "f32 bulk + bf16 MFMA weights / f64 accumulate"}[prec],
        "data": "synthetic initial condition of the named case (generated, not read)",
        "config": {"workload": workload, "cells": ncell, "sys_size": S, "cfl": 0.4, "dt": dt,
                   "precision": args.precision, "form": FORMS[args.precision], "time_stepper": "TVD-RK3"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": prof.get("traffic"),
                     "algorithmic_bytes_per_launch": stage_bytes, "avg_launch_ms": dom_avg_s * 1e3,
                     "limiter": prof.get("limiter"), "valu_busy_profiled": prof.get("valu_busy"),
                     "profile": prof.get("profile")},
        # the step's compulsory state traffic, and the bytes this design's
        # kernels must move (the sweep_bytes_per_step model: rhs partials, RK
        # base, and in viscous runs the face records and finish passes)
        "step_hbm": {"algorithmic_bytes_per_cell_step": 64 * S, "achieved_GBps": step_bytes * args.steps / sec / 1e9,
                     "frac": step_bytes * args.steps / sec / 1e9 / HBM_PEAK_GBS,
                     "design_bytes_per_cell_step": sum(per_step.values()),
                     "design_GBps": sum(per_step.values()) * ncell * args.steps / sec / 1e9},
        "kernels_ms_per_step": {k: v[0] / args.steps for k, v in timing.items() if v[1]},
    }
    if not prm.viscous:
        out["fp64_valu"] = {"nominal_flops_per_cell_step": f_alg, "achieved_TFLOPs": f_alg * value / 1e12,
                            "peak_TFLOPs": FP64_PEAK_TFLOPS, "frac": f_alg * value / 1e12 / FP64_PEAK_TFLOPS}
    # every kernel of the step: one stage's algorithmic bytes per launch, and
    # the bytes the fused design must move (stencil input, rhs partials, RK
    # base and output) beside it
    kr = {}
    for k, bps in per_step.items():
        ms, n = timing[k]
        if n:
            per_launch = bps * ncell / (n / args.steps)
            kr[k] = {"avg_launch_ms": ms / n, "launches_per_step": n / args.steps,
                     "algorithmic_bytes_per_launch": stage_bytes,
                     "achieved_GBps": stage_bytes / (ms / 1e3 / n) / 1e9,
                     "frac": stage_bytes / (ms / 1e3 / n) / 1e9 / HBM_PEAK_GBS,
                     "design_bytes_per_launch": per_launch,
                     "design_GBps": per_launch / (ms / 1e3 / n) / 1e9}
    out["kernels_roofline"] = kr
    if not args.no_aux:
        out["aux_kernels"] = aux_kernels(mod, prm, grids, q1, dt, args, RhsModule, torch)
    else:
        mod.finalize()
    del q1, q2, rhs
    torch.cuda.empty_cache()

    if compare:
        # the plain FP64 form (bit-exact against the oracle) on the same case,
        # the same dt and the same number of steps: its time, and the timed
        # form's L1 distance from it after those steps
        prm.precision = PREC_FP64
        psec, ptiming, _, pmod, bufs = time_steps(RhsModule, torch, ic, prm, grids, q, args, dt=dt)
        pmod.finalize()
        plain = prm.interior(bufs[0])
        # momentum components share one norm (a transverse momentum that
        # is zero by symmetry would otherwise divide rounding by rounding)
        nf = prm.num_fluids
        mom = sum(float(plain[v].abs().sum()) for v in range(nf, nf + D))
        rel = [float((final[v] - plain[v]).abs().sum()) / (mom if nf <= v < nf + D else float(plain[v].abs().sum()))
               for v in range(S)]
        del bufs, plain, final
        torch.cuda.empty_cache()
        prm.precision = prec
        out["plain_fp64"] = {"value": ncell * args.steps / psec, "ms_per_step": 1e3 * psec / args.steps,
                             "kernels_ms_per_step": {k: v[0] / args.steps for k, v in ptiming.items() if v[1]},
                             "form": FORMS["fp64"]}
        out["matched_error"] = {"vs": "plain FP64 form, same initial state, dt and steps",
                                "steps": args.warmup + args.steps, "rel_l1_max": max(rel),
                                "rel_l1_per_variable": rel,
                                "speedup_over_plain_fp64": psec / sec}

    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(prm, grids, q, np, ic)
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


def _cpu_step_seconds(O, L, prm, grids, q, np, threads):
    oc = O.Case(m=prm.m, n=prm.n, p=prm.p, num_fluids=prm.num_fluids, weno_order=prm.weno_order,
                time_stepper=prm.time_stepper, weno_eps=prm.weno_eps, gammas=list(prm.gammas),
                pi_infs=list(prm.pi_infs), bc_beg=prm.bc_beg, bc_end=prm.bc_end,
                Re=[tuple(r) for r in prm.Re] if prm.Re else None)
    L.orc_set_threads(threads)
    q1 = q.copy()
    q2 = np.zeros_like(q1)
    t0 = time.perf_counter()
    O.tvd_rk_step(oc, grids, 1e-9, q1, q2, t_step=0)
    return time.perf_counter() - t0


def _sample(prm, grids, q, ic, cells):
    """The benchmark case itself when it has at most `cells` cells, else a
    slab of it through the domain centre (whole extent in x and y, as many
    z (3D) or y (2D) layers as fit), cut faces extrapolated."""
    ncell = (prm.m + 1) * (prm.n + 1) * (prm.p + 1)
    if ncell <= cells:
        return prm, grids, q, "the whole grid"
    D = prm.num_dims
    ext = [prm.m + 1, prm.n + 1, prm.p + 1][:D]
    layer = 1
    for e in ext[:-1]:
        layer *= e
    k = max(5 * prm.weno_order, cells // layer)
    lo = [0] * D
    shape = list(ext)
    lo[-1] = (ext[-1] - k) // 2
    shape[-1] = k
    sub = ic.sub_box(prm, grids, q, lo, shape)
    axis = "zy"[3 - D] if D > 1 else "x"
    return sub + (f"a {'x'.join(map(str, shape))} slab of it ({axis} = {lo[-1]}..{lo[-1] + k - 1})",)


def cpu_baseline(prm, grids, q, np, ic):
    """The FP64 oracle on the host: one TVD-RK3 step (same dt rule) of a
    bounded sample of the same case, on every host thread the run may use
    (OpenMP over pencils, OMP_NUM_THREADS) and on one core."""
    import ctypes

    from oracle import oracle as O
    L = O.lib()
    L.orc_set_threads.argtypes = [ctypes.c_int]
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    pa, ga, qa, what_a = _sample(prm, grids, q, ic, 4_400_000)
    sec_all = _cpu_step_seconds(O, L, pa, ga, qa, np, threads)
    pb, gb, qb, what_b = _sample(prm, grids, q, ic, 600_000)
    sec_one = _cpu_step_seconds(O, L, pb, gb, qb, np, 1)
    na = (pa.m + 1) * (pa.n + 1) * (pa.p + 1)
    nb = (pb.m + 1) * (pb.n + 1) * (pb.p + 1)
    return {"value": na / sec_all, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "sample": f"one TVD-RK3 step of {what_a} ({na} cells) on {threads} host threads ({sec_all:.2f} s); "
                      f"one core: {what_b} ({nb} cells, {sec_one:.2f} s)",
            "single_core": {"value": nb / sec_one, "cores": 1}}


if __name__ == "__main__":
    main()
