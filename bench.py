"""Benchmark: whole SIMPLE iterations on one MI355X (the headline), with the pressure corrector alone, the
pressure predictor's kernels, serving and the CPU baseline beside them.

Headline workload (one step = one SIMPLE iteration of simpleFoam with kEpsilon on the 500x250x160
graded (ratio 20), 10 %-jittered channel 4 x 1 x 0.5, 20,000,000 cells): UEqn (upwind convection with
linearUpwind's deferred correction, corrected laplacian(nuEff, U), divDevReff's explicit stress, relax 0.7),
the momentum predictor (PBiCGStab per component), the pressure predictor (rAU, HbyA, phiHbyA, adjustPhi),
nNonOrthogonalCorrectors + 1 pressure solves (corrected laplacian, AMG-PCG to 1e-6 normalised residual),
phi, p relaxed by 0.3, U corrected, then the k-epsilon correct() with nutk/epsilon wall functions. The
mixed path runs the pressure V-cycle in FP16 with per-grid power-of-two scaling (--precision fp32: FP32)
inside FP64 flexible CG; the FP64 path runs the same iterations with an FP64 cycle. `value` is the mixed
path's seconds per iteration (lower is better); the FP64 time, the ratio, iterations, residuals and the
mixed/FP64 field differences are reported beside it.

Inputs are synthetic and seeded: perturbed plug flow U = (10,0,0) + uniform(-1,1) per component (inlet
fixedValue (10,0,0), k 0.375, epsilon 14.855; outlet zeroGradient with p fixedValue 0; walls no-slip with
wall functions), p = 0, phi = interpolate(U).Sf; the timed iterations start from the seeded fields.

Opt-in legs, each a key of the same JSON line: --c5 (the same iteration on the 800x400x375 channel,
120 M cells, FP32 against FP64, with peak HBM), --batch N (N small solves at once, batched into one
launch against the per-grid kernels), --gamg-step 1 (OpenFOAM's GAMG-PCG on the device), --profile-simple
(only the mixed SIMPLE iterations, for a profiler).
"""
import argparse
import collections
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

METRIC = "s/SIMPLE-iter @20M cells, FP64 vs mixed AMG at equal resid; solves/s + p99"
PEAK_HBM_GBS = 8000.0
NTAGS = 144
GRID_DOWN, GRID_UP = 16, 80


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", default="500x250x160")
    ap.add_argument("--precision", default="fp16", choices=["fp32", "fp16"],
                    help="the mixed path's V-cycle (the other one is timed beside it)")
    ap.add_argument("--no-other", action="store_true", help="skip timing the other mixed precision")
    ap.add_argument("--tolerance", type=float, default=1e-6)
    ap.add_argument("--cpu-sample", default="160x80x50", help="channel the CPU oracle is timed on")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--check-every", type=int, default=4)
    ap.add_argument("--graph", type=int, default=0, help="1: replay captured groups of PCG iterations")
    ap.add_argument("--scale", action="store_true",
                    help="scale each grid's prolonged correction by (r.e)/(e.Ae) (brae's corrScaling); off by default: "
                         "with V(1,1) on two-pass levels the SIMPLE iteration takes the same PCG iterations without it, "
                         "one SpMV per grid cheaper")
    # V(2,1) on levels of two composed pairwise passes: the fastest measured cycle for the FP64 path and
    # within 1 % of the fastest for the FP16 one on the 20 M channel
    ap.add_argument("--npre", type=int, default=1)
    ap.add_argument("--npost", type=int, default=1)
    ap.add_argument("--merge-levels", type=int, default=2, help="pairwise passes per AMG level")
    ap.add_argument("--non-orth", type=int, default=1, help="nNonOrthogonalCorrectors (solves per step - 1)")
    ap.add_argument("--serving", type=int, default=512, help="requests of the serving stream (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=8.0, help="Poisson arrival rate, requests/s")
    ap.add_argument("--serve-slots", type=int, default=8)
    ap.add_argument("--serve-resident", type=float, default=0.0,
                    help="GB of the serving budget that keep systems between requests (the stream repeats no box)")
    ap.add_argument("--predictor", type=int, default=1,
                    help="1: also time the pressure predictor's glue (rAU, HbyA, constrainHbyA, phiHbyA, adjustPhi)")
    ap.add_argument("--simple", type=int, default=1, help="1: time whole SIMPLE iterations (the headline)")
    ap.add_argument("--nu", type=float, default=1e-5, help="laminar viscosity of the SIMPLE iterations")
    ap.add_argument("--turbulence", type=int, default=1, help="1: k-epsilon with wall functions; 0: laminar")
    ap.add_argument("--tol-u", type=float, default=1e-5, help="momentum solve tolerance")
    ap.add_argument("--reltol-u", type=float, default=0.1, help="momentum solve relative tolerance")
    ap.add_argument("--batch", type=int, default=0,
                    help="also time this many small solves at once, batched and on the per-grid kernels")
    ap.add_argument("--c5", action="store_true",
                    help="also time the largest single-GPU case: the 800x400x375 channel (120 M cells), turbulent SIMPLE "
                         "iterations with the FP32 and the FP64 cycle, and the peak HBM in use (adds a few minutes)")
    ap.add_argument("--profile-simple", action="store_true",
                    help="only the mixed path's SIMPLE iterations (for a profiler run): no FP64 or pressure-corrector "
                         "legs, no serving or CPU baseline; prints one short JSON line")
    ap.add_argument("--gamg-step", type=int, default=0,
                    help="1: also time the step with OpenFOAM's GAMG-PCG on the device (the CPU baseline's solver)")
    return ap.parse_args()


def simple_leg(args, ch, ctx, torch, mfv, prec_mixed, other, on_step=None):
    from brae_amd import channel as C
    """Whole SIMPLE iterations on the channel (the headline): UEqn (upwind convection with linearUpwind's
    deferred correction, corrected laplacian(nuEff, U), the explicit stress, relax 0.7), the momentum
    predictor (PBiCGStab per component), the pressure predictor, nNonOrthogonalCorrectors + 1 pressure
    solves (AMG-PCG to --tolerance) with the corrected laplacian, phi, p relaxed by 0.3 and U corrected.
    The mixed path and the FP64 path run the same iterations from the same seeded fields; only the
    pressure V-cycle's precision differs (and the outer CG is flexible when the cycle is not FP64)."""
    uT, uV, pT, pV = C.channel_bcs(ch, torch)
    geom = mfv.Geometry(*(mfv.ptr(t) for t in (ch.Sf, ch.Cf, ch.magSf, ch.C, ch.V, ch.w, ch.dc, ch.ndc, ch.corr,
                                                ch.bDelta)))
    U_ctl = dict(tolerance=args.tol_u, relTol=args.reltol_u, maxIter=1000,
                 preconditioner=mfv.PRECOND_DIAGONAL, checkEvery=args.check_every)

    def controls(prec):
        return mfv.simple_controls(
            U=U_ctl, p=dict(tolerance=args.tolerance, relTol=0.0, precision=prec, nPreSweeps=args.npre,
                            nPostSweeps=args.npost, flexible=int(prec != mfv.PREC_FP64 or args.scale),
                            scaleCorrection=int(args.scale), checkEvery=args.check_every,
                            captureVcycle=args.graph),
            mergeLevels=args.merge_levels, nu=args.nu, relaxU=0.7, relaxP=0.3, linearUpwind=1,
            correctedLaplacian=1, nNonOrthCorr=args.non_orth, pRefCell=-1)

    turb = None
    if args.turbulence:
        turb = C.channel_kepsilon(ch, torch, mfv, nu=args.nu, tol=args.tol_u, relTol=args.reltol_u,
                                  checkEvery=args.check_every, log=log)
    runs = [("mixed", prec_mixed), ("fp64", mfv.PREC_FP64)] + ([] if args.no_other else [("other", other)])
    if args.profile_simple:
        runs = runs[:1]
    out = {}
    for name, prec in runs:
        t0 = time.time()
        sim = mfv.Simple(ch.mesh, geom, uT, uV.ravel(), pT, pV, controls(prec))
        ke = None
        if turb is not None:
            ke = mfv.KEpsilon(ch.mesh, geom, turb["isWall"], turb["y"], turb["type"], turb["kV"], turb["type"],
                              turb["eV"], turb["controls"])
        log("SIMPLE %s: hierarchy built (%.1fs)" % (name, time.time() - t0))

        def fresh():
            f = C.simple_fields(ch, torch, uT, uV)
            if turb is not None:
                f.update({k: v.clone() for k, v in turb["fields0"].items()})
            torch.cuda.synchronize()
            return f

        def iterate(f):
            a = (f["U"], f["Ub"], f["p"], f["pb"], f["phi"], f["phiB"])
            if ke is None:
                return sim.step(*a), None
            return mfv.simple_step_kepsilon(sim, ke, *a, f["k"], f["kb"], f["eps"], f["epsb"], f["nut"], f["nutb"])

        f = fresh()
        for _ in range(args.warmup):
            iterate(f)
        # the timed iterations start again from the seeded fields
        del f
        f = fresh()
        t = time.perf_counter()
        res = [iterate(f) for _ in range(args.steps)]
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t) / args.steps
        if on_step:
            on_step()
        # one more iteration with events around its phases (outside the timed ones)
        ctx.profile(True)
        for tag in mfv.PROF_SIMPLE:
            ctx.profile_read(tag)
        iterate(f)
        phases = {mfv.PROF_SIMPLE[tag]: round(ctx.profile_read(tag)[1], 3) for tag in mfv.PROF_SIMPLE}
        ctx.profile(False)
        phases = {k: v for k, v in phases.items() if v > 0}
        out[name] = dict(dt=dt, res=[r for r, _ in res], keres=[kr for _, kr in res], p=f["p"].clone(),
                         U=torch.stack(f["U"]).clone(), phases_ms=phases,
                         k=None if turb is None else f["k"].clone(), eps=None if turb is None else f["eps"].clone())
        rr = out[name]["res"]
        log("SIMPLE %s: %.4f s/iter; p iterations %s; U iterations %s; final p residuals %s; phases (ms) %s" % (
            name, dt, [list(r.pIters[:args.non_orth + 1]) for r in rr], [list(r.UIters) for r in rr],
            ["%.2e" % r.pFinal for r in rr], phases))
        if turb is not None:
            log("  k-epsilon: eps/k iterations %s" % [(kr.epsIters, kr.kIters) for kr in out[name]["keres"]])
        sim.close()
        if ke is not None:
            ke.close()
        del f
        torch.cuda.synchronize()
    if args.profile_simple:
        m = out["mixed"]
        return dict(dt_mixed=m["dt"], p_iterations=[list(r.pIters[:args.non_orth + 1]) for r in m["res"]],
                    final_p_residual=max(r.pFinal for r in m["res"]), phases_ms=m["phases_ms"])
    m, d = out["mixed"], out["fp64"]
    line = dict(
        dt_mixed=m["dt"], dt_fp64=d["dt"],
        p_iterations={k: [list(r.pIters[:args.non_orth + 1]) for r in v["res"]] for k, v in out.items()},
        U_iterations={k: [list(r.UIters) for r in v["res"]] for k, v in out.items()},
        final_p_residual={k: max(r.pFinal for r in v["res"]) for k, v in out.items()},
        p_rel_l2_vs_fp64={k: float((v["p"] - d["p"]).norm() / d["p"].norm()) for k, v in out.items() if k != "fp64"},
        U_rel_l2_vs_fp64={k: float((v["U"] - d["U"]).norm() / d["U"].norm()) for k, v in out.items() if k != "fp64"},
        phases_ms={k: v["phases_ms"] for k, v in out.items()})
    if turb is not None:
        line["ke_iterations"] = {k: [(kr.epsIters, kr.kIters) for kr in v["keres"]] for k, v in out.items()}
        line["k_rel_l2_vs_fp64"] = {k: float((v["k"] - d["k"]).norm() / d["k"].norm())
                                    for k, v in out.items() if k != "fp64"}
        line["eps_rel_l2_vs_fp64"] = {k: float((v["eps"] - d["eps"]).norm() / d["eps"].norm())
                                      for k, v in out.items() if k != "fp64"}
        line["wall_cells"] = out["mixed"]["keres"][0].nWallCells
    if "other" in out:
        line["dt_other"] = out["other"]["dt"]
    return line


def gpu_run(args):
    import torch
    from brae_amd import channel as C
    from brae_amd import hexmesh, mfv

    nx, ny, nz = (int(v) for v in args.size.split("x"))
    # one stream (default priority) for torch's own work and the library's, so the two stay in order
    stream = torch.cuda.Stream(device="cuda:0")
    torch.cuda.set_stream(stream)
    ctx = mfv.Context(0, stream=stream.cuda_stream)
    ch = C.Channel(nx, ny, nz, torch, mfv, hexmesh, ctx, log=log)
    prec_mixed = mfv.PREC_FP32 if args.precision == "fp32" else mfv.PREC_FP16
    other = mfv.PREC_FP16 if prec_mixed == mfv.PREC_FP32 else mfv.PREC_FP32
    simple = simple_leg(args, ch, ctx, torch, mfv, prec_mixed, other) if args.simple or args.profile_simple else None
    if args.profile_simple:
        ch.mesh.close()
        ctx.close()
        calibration_kernels(torch, ch.nC)
        return dict(profile_simple=simple)
    t0 = time.time()
    w_host = ch.magSfI.cpu().numpy()
    amg = mfv.Amg(ch.mesh, w_host, mfv.AGG_BRAE, 64, mergeLevels=args.merge_levels)
    sizes = [amg.grid_size(g) for g in range(amg.nGrids)]
    log("AMG hierarchy: %d grids, coarsest %d cells (%.1fs host build)" % (amg.nGrids, sizes[-1][0], time.time() - t0))
    P = mfv.ptr
    A = mfv.ldu(ch.diagC, ch.upper)

    def step(prec, gamg=None):
        """one SIMPLE pressure corrector (brae_amd.channel.pressure_corrector)"""
        return C.pressure_corrector(ch, ctx, mfv, amg, prec, tolerance=args.tolerance, nonOrth=args.non_orth,
                                    checkEvery=args.check_every, nPre=args.npre, nPost=args.npost,
                                    scale=args.scale, graph=args.graph, gamg=gamg)

    def timed(prec, profile):
        for _ in range(args.warmup):
            step(prec)
        torch.cuda.synchronize()
        if profile:
            ctx.profile(True)
            for tag in range(NTAGS):
                ctx.profile_read(tag)
        torch.cuda.synchronize()
        t = time.perf_counter()
        perfs = [q for _ in range(args.steps) for q in step(prec)]
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t) / args.steps
        prof = None
        if profile:
            prof = {tag: ctx.profile_read(tag) for tag in range(NTAGS)}
            ctx.profile(False)
        return dt, perfs, prof

    res = {}
    runs = [("mixed", prec_mixed), ("fp64", mfv.PREC_FP64)] + ([] if args.no_other else [("other", other)])
    for name, prec in runs:
        dt, perfs, _ = timed(prec, profile=False)
        res[name] = dict(dt=dt, perfs=perfs, p=ch.p.clone())
        log("%s (%s): %.4f s/step, iterations %s, final residual %.3e" % (
            name, {64: "fp64", 32: "fp32", 16: "fp16"}[prec], dt, [q.nIterations for q in perfs],
            perfs[-1].finalResidual))
    # per-kernel timing: one more step of the mixed path with events around the tagged launches (kept out
    # of the timed steps above)
    # (checkEvery 1: the stop flag is polled after every iteration, so no launch is queued past
    # convergence and every tagged launch is a working one)
    saved = args.steps, args.warmup, args.check_every
    args.steps, args.warmup, args.check_every = 1, 0, 1
    _, prof_perfs, prof = timed(prec_mixed, profile=True)
    args.steps, args.warmup, args.check_every = saved
    res["mixed"]["prof"] = prof
    res["mixed"]["prof_perfs"] = prof_perfs
    pm, p64 = res["mixed"]["p"], res["fp64"]["p"]
    rel_l2 = float((pm - p64).norm() / p64.norm())
    other_out = None
    if "other" in res:
        po = res["other"]
        other_out = {"v_cycle": {16: "fp16", 32: "fp32"}[other], "s_per_iter": round(po["dt"], 5),
                     "speedup_vs_fp64": round(res["fp64"]["dt"] / po["dt"], 3),
                     "pcg_iterations": [q.nIterations for q in po["perfs"]],
                     "final_residual": po["perfs"][-1].finalResidual,
                     "p_rel_l2_vs_fp64": float((po["p"] - p64).norm() / p64.norm())}

    # roofline of the dominant kernel of the mixed step (largest summed event time)
    N, F, Nc = ch.nC, ch.nI, sizes[1][0]
    # widths in the mixed cycle: vectors (FP16 cycle: FP16; FP32 cycle: FP32), coefficients, the diagonal
    # (FP32 in both)
    zb = 4 if prec_mixed == mfv.PREC_FP32 else 2
    vb = 4 if prec_mixed == mfv.PREC_FP32 else 2
    dw = 4
    iters = sum(q.nIterations for q in res["mixed"]["prof_perfs"])
    # the fine grid read as diagonals streams no owner/neighbour indices: the index term is credited
    # only to the sliced-row layout, which reads explicit columns or pattern ids
    idx = 0 if amg.fine_stencil() else 8 * F
    names = {
        mfv.PROF_PCG_SPMV64: (("pcgDirDiaK: p = z + beta*p, q = A.p (FP64 diagonals), p.q" if amg.fine_stencil() else
                               "pcgDirAmulK: p = z + beta*p, q = A.p (FP64), p.q"),
                              8 * (N + F) + idx + zb * N + 24 * N, iters),
        # reads p, q, psi, r; writes psi and the next r (into the other of two alternating residual
        # buffers: the flexible beta's r_old needs no copy); the next cycle's b and x; the diagonal
        mfv.PROF_PCG_UPDATE: ("pcgUpdateK: psi, r update, sumMag(r), next cycle entry",
                              8 * (4 * N + 2 * N) + 2 * zb * N + dw * N, iters),
        mfv.PROF_PCG_ZR: ("pcgZrK: z.r, z.r_old", zb * N + 16 * N, iters),
        mfv.PROF_AMG_RESIDUAL0: ("residualK grid 0 (cycle precision)", vb * F + dw * N + 3 * zb * N + idx, iters),
        mfv.PROF_AMG_CORR_AMUL0: ("corrAmulK grid 0: A e, r.e, e.Ae", vb * F + dw * N + 3 * zb * N + idx, iters),
        mfv.PROF_AMG_CORR_SMOOTH0: (("corrSmoothK grid 0", 5 * zb * N + dw * N, iters) if args.scale else
                                    ("prolongSmoothK grid 0: x + e, one sweep", vb * F + dw * N + 4 * zb * N + idx,
                                     iters)),
        mfv.PROF_AMG_RESTRICT0: ("restrictK grid 0 -> 1", zb * N + 4 * N + 4 * Nc + 2 * zb * Nc + dw * Nc, iters),
    }
    prof = res["mixed"]["prof"]
    dom = max(names, key=lambda t: prof[t][1])
    label, alg_bytes, _ = names[dom]
    # the profiled step polls the stop flag after every iteration, so every tagged launch did work;
    # a mismatch is reported (noop_launches) rather than averaged in silently
    noop = prof[dom][0] - iters
    if noop:
        log("warning: %d tagged launches of %s past convergence in the profiled step" % (noop, label))
    avg_ms = prof[dom][1] / max(1, prof[dom][0])
    achieved = alg_bytes / (avg_ms * 1e-3) / 1e9
    per_kernel = {names[t][0]: {"launches": prof[t][0], "avg_us": round(prof[t][1] / max(1, prof[t][0]) * 1e3, 2),
                                "alg_bytes": int(names[t][1]),
                                "alg_GBps": round(names[t][1] / (prof[t][1] / max(1, prof[t][0]) * 1e-3) / 1e9, 1)}
                  for t in names if prof[t][1] > 0}
    free_b, total_b = torch.cuda.mem_get_info()
    hbm_used_gib = (total_b - free_b) / 2 ** 30
    copy_gbs = copy_peak(torch, N)
    triad_gbs = triad_peak(torch, N)
    vc_n, vc_ms = prof[mfv.PROF_VCYCLE]
    grids = []
    for gi in range(amg.nGrids):
        dn, up = prof[GRID_DOWN + gi][1], prof[GRID_UP + gi][1]
        grids.append({"grid": gi, "cells": sizes[gi][0], "down_us": round(dn / max(1, iters) * 1e3, 1),
                      "up_us": round(up / max(1, iters) * 1e3, 1)})
        log("  grid %2d %10d cells: down %8.1f us  up %8.1f us" % (gi, sizes[gi][0], dn / max(1, iters) * 1e3,
                                                                   up / max(1, iters) * 1e3))
    out = dict(
        dt_mixed=res["mixed"]["dt"], dt_fp64=res["fp64"]["dt"],
        iters_mixed=[q.nIterations for q in res["mixed"]["perfs"]],
        iters_fp64=[q.nIterations for q in res["fp64"]["perfs"]],
        final_mixed=res["mixed"]["perfs"][-1].finalResidual, final_fp64=res["fp64"]["perfs"][-1].finalResidual,
        p_rel_l2_mixed_vs_fp64=rel_l2, n_grids=amg.nGrids, coarsest=sizes[-1][0],
        roofline=dict(bound="hbm", achieved=round(achieved, 1), peak=PEAK_HBM_GBS, unit="GB/s",
                      frac=round(achieved / PEAK_HBM_GBS, 4), traffic=measured_traffic(label), kernel=label,
                      alg_bytes_per_launch=int(alg_bytes), avg_launch_us=round(avg_ms * 1e3, 2),
                      launches=prof[dom][0], noop_launches=noop,
                      copy_peak_GBps=round(copy_gbs, 1), triad_GBps=round(triad_gbs, 1)),
        per_kernel=per_kernel, vcycle_avg_us=round(vc_ms / max(1, iters) * 1e3, 1), grids=grids,
        hbm_used_gib=round(hbm_used_gib, 2), other=other_out, simple=simple,
    )
    if args.predictor:
        out["predictor"] = predictor_run(ch, ctx, patches=ch.patches)
    if args.gamg_step:
        # the CPU baseline's own algorithm on the device: OpenFOAM's faceAreaPair hierarchy and GAMG-PCG
        Sf = ch.Sf[:3 * ch.nI].cpu().numpy()
        wts = np.zeros(ch.nI)
        mfv.lib().mfv_faceareapair_weights(ch.nI, Sf.ctypes.data, w_host.ctypes.data, wts.ctypes.data)
        t0 = time.time()
        gamg = mfv.Amg(ch.mesh, wts, mfv.AGG_FACEAREAPAIR, 10)
        build_s = time.time() - t0
        step(None, gamg)   # first use: the coarse grids' addressing and sweep schedules
        torch.cuda.synchronize()
        t = time.perf_counter()
        gp = step(None, gamg)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t
        log("openfoam gamg-pcg on the device: %.4f s/step, iterations %s (%d grids, %.1fs host build)" % (
            dt, [q.nIterations for q in gp], gamg.nGrids, build_s))
        out["gamg_step"] = dict(s_per_step=round(dt, 4), pcg_iterations=[q.nIterations for q in gp],
                                final_residual=gp[-1].finalResidual, grids=gamg.nGrids,
                                solver="OpenFOAM PCG + GAMG preconditioner (faceAreaPair, GaussSeidel, "
                                       "V(0,2,2), nVcycles 2), FP64, tree reductions")
        gamg.close()
    amg.close()
    ch.mesh.close()
    ctx.close()
    return out


def c5_run(args):
    """The largest single-GPU configuration: turbulent SIMPLE iterations on the 800x400x375 channel (120 M
    cells) with the FP32 cycle (the configuration's mixed precision) and with the FP64 cycle, as the headline's
    iterations (one untimed, then --steps timed, from the seeded fields), and the peak HBM in use."""
    import copy
    import torch
    from brae_amd import channel as C
    from brae_amd import hexmesh, mfv
    a = copy.copy(args)
    a.size, a.no_other, a.profile_simple = "800x400x375", True, False
    a.steps = min(args.steps, 2)
    import gc
    gc.collect()
    torch.cuda.synchronize()
    torch.cuda.empty_cache()   # the 20 M legs' cached blocks back to the device for the library's allocations
    free_b, total_b = torch.cuda.mem_get_info()
    before = (total_b - free_b) / 2 ** 30
    stream = torch.cuda.Stream(device="cuda:0")
    torch.cuda.set_stream(stream)
    ctx = mfv.Context(0, stream=stream.cuda_stream)
    t0 = time.time()
    ch = C.Channel(800, 400, 375, torch, mfv, hexmesh, ctx, log=log, corrector=False)
    peak = [0.0]

    def used():
        free_b, total_b = torch.cuda.mem_get_info()
        peak[0] = max(peak[0], (total_b - free_b) / 2 ** 30)
    used()
    r = simple_leg(a, ch, ctx, torch, mfv, mfv.PREC_FP32, mfv.PREC_FP16, on_step=used)
    ch.mesh.close()
    ctx.close()
    out = dict(workload="turbulent SIMPLE iteration (as the headline) on the 800x400x375 channel, 120,000,000 cells",
               fp32_s_per_iter=round(r["dt_mixed"], 4), fp64_s_per_iter=round(r["dt_fp64"], 4),
               speedup_fp32_vs_fp64=round(r["dt_fp64"] / r["dt_mixed"], 3), p_iterations=r["p_iterations"],
               final_p_residual=r["final_p_residual"], p_rel_l2_fp32_vs_fp64=r["p_rel_l2_vs_fp64"].get("mixed"),
               phases_ms=r["phases_ms"], peak_hbm_gib=round(peak[0] - before, 1),
               hbm_in_use_before_gib=round(before, 1), wall_s=round(time.time() - t0, 1))
    log("c5: FP32 %.3f s/iter, FP64 %.3f s/iter, peak HBM %.1f GiB" % (r["dt_mixed"], r["dt_fp64"], peak[0] - before))
    return out


def predictor_run(ch, ctx, patches, reps=5):
    """The SIMPLE pressure predictor's glue (UEqn.relax() on the momentum matrix, linearUpwind's
    deferred correction and divDevReff's explicit stress, then pEqn.cu
    pressurePredictor: A()'s folded diagonal and rAU,
    HbyA = rAU*H() per component, constrainHbyA, phiHbyA with its boundary faces, adjustPhi) on the
    channel, with a synthetic momentum matrix: upwind convection of the channel's flux made diagonally
    dominant, seeded internalCoeffs/boundaryCoeffs and sources, a seeded velocity. Inlet and walls fix U
    (constrainHbyA takes U there), the outlet is adjustable. Per kernel: HIP events on the library's
    stream, algorithmic bytes as for the pressure path (each array once, int32 owner/neighbour credited)."""
    import torch
    from brae_amd import mfv

    P, N, F, B = mfv.ptr, ch.nC, ch.nI, ch.nB
    g = torch.Generator(device="cuda:0")
    g.manual_seed(11)

    def rnd(n, lo=-1.0, hi=1.0):
        return torch.rand(n, dtype=torch.float64, device="cuda:0", generator=g) * (hi - lo) + lo

    def z(n):
        return torch.zeros(n, dtype=torch.float64, device="cuda:0")

    mD, mU, mL = z(N), z(F), z(F)
    ctx.call("mfv_div_upwind_coeffs", ch.mesh.h, P(ch.phiHbyA), P(mD), P(mU), P(mL))
    mD += 6.0 * mU.abs().max()
    ic = [rnd(B, 0.1, 1.0) for _ in range(3)]
    bc = [rnd(B) for _ in range(3)]
    src = [rnd(N) * ch.V for _ in range(3)]
    U = [rnd(N) for _ in range(3)]
    Ub = [rnd(B) for _ in range(3)]
    takeU = torch.ones(B, dtype=torch.int32, device="cuda:0")
    adjustable = torch.zeros(B, dtype=torch.int32, device="cuda:0")
    for name, start, size in patches:
        if name == "outlet":
            takeU[start - F:start - F + size] = 0
            adjustable[start - F:start - F + size] = 1
    A = mfv.ldu(mD, mU, mL)
    D, rAU, phi, phiB, R, delta, icMag, icMin = z(N), z(N), z(F), z(B), z(N), z(N), z(B), z(B)
    gU = [rnd(3 * N) for _ in range(3)]
    gradU9, nuEff, sigmaB = rnd(9 * N), rnd(N, 1e-3, 1e-1), rnd(9 * B)
    divSig = [z(N) for _ in range(3)]
    corrU = [z(N) for _ in range(3)]
    bd = [z(B) for _ in range(3)]
    HbyA = [z(N) for _ in range(3)]
    Hb = [z(B) for _ in range(3)]
    stream = torch.cuda.current_stream()
    events = collections.defaultdict(list)

    def timed(name, fn):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        fn()
        e1.record(stream)
        events[name].append((e0, e1))

    mesh = ch.mesh.h
    massCorr = []
    for _ in range(reps + 1):
        # UEqn.relax(): cmptMax(cmptMag(iC)), cmptMin(iC), then the relaxed diagonal and the three sources
        timed("cmpt3K (max|iC|, min iC)", lambda: (
            ctx.call("mfv_cmpt_max_mag3", B, *(P(a) for a in ic), P(icMag)),
            ctx.call("mfv_cmpt_min3", B, *(P(a) for a in ic), P(icMin))))
        timed("relaxK (diagonal, 3 sources)",
              lambda: ctx.call("mfv_relax_diag", mesh, ctypes.byref(A), P(ic[0]), 0.7, None, P(icMag), P(icMin),
                               P(R), P(delta), 3, mfv.ptr3(U), mfv.ptr3(src)))
        # linearUpwind's deferred correction of div(phi, U), the three components in one pass
        timed("linearUpwindCorrK (3 components)",
              lambda: ctx.call("mfv_linear_upwind_corr", mesh, P(ch.phiHbyA), P(ch.Cf), P(ch.C), 3, mfv.ptr3(gU),
                               mfv.ptr3(corrU)))
        # divDevReff's explicit stress: nuEff*dev2(T(grad U)), then V*div of it with the boundary faces' stress
        # (formed where each cell's stress is read: one pass, the stress never stored)
        timed("tensorDivK (V div nuEff dev2(T(grad U)), fused)",
              lambda: ctx.call("mfv_div_dev_reff_explicit", mesh, P(ch.w), P(ch.Sf), None, P(gradU9), P(nuEff),
                               P(sigmaB), *(P(t) for t in divSig)))
        timed("foldBoundaryDiagK", lambda: ctx.call("mfv_fold_boundary_diag", mesh, P(R), *(P(a) for a in ic), P(D)))
        timed("reciprocalVK", lambda: ctx.call("mfv_reciprocal_v", mesh, P(D), P(ch.V), P(rAU)))
        for k in range(3):
            timed("boundaryDiagTermK", lambda: ctx.call("mfv_boundary_diag_term", mesh, *(P(a) for a in ic), k,
                                                        P(bd[k])))
        # HbyA of the three components in one pass over the momentum matrix
        timed("matrixH3K (HbyA, 3 components)",
              lambda: ctx.call("mfv_matrix_h3", mesh, ctypes.byref(A), mfv.ptr3(U), mfv.ptr3(src), mfv.ptr3(bd),
                               mfv.ptr3(bc), P(ch.V), P(rAU), None, mfv.ptr3(HbyA)))
        for k in range(3):
            timed("constrainHbyAK", lambda: ctx.call("mfv_constrain_hbya", mesh, P(takeU), P(Ub[k]), P(HbyA[k]),
                                                     P(Hb[k])))
        timed("vectorFluxK", lambda: ctx.call("mfv_vector_flux", mesh, P(ch.w), P(ch.Sf), *(P(h) for h in HbyA),
                                              P(phi)))
        timed("boundaryFluxK", lambda: ctx.call("mfv_boundary_flux", mesh, P(ch.Sf), *(P(h) for h in Hb), None,
                                                P(phiB)))
        mc = ctypes.c_double()
        timed("adjustPhi (sums, decision, scale)",
              lambda: ctx.call("mfv_adjust_phi", B, P(adjustable), P(phiB), F, P(phi), ctypes.byref(mc), None))
        massCorr.append(mc.value)
    torch.cuda.synchronize()
    alg = {"cmpt3K (max|iC|, min iC)": 2 * 32 * B,
           "relaxK (diagonal, 3 sources)": 16 * F + 4 * F + 8 * N + 16 * N + 3 * 24 * N + 16 * B,
           "linearUpwindCorrK (3 components)": 8 * F + 24 * F + 8 * F + 24 * N + 72 * N + 24 * N,
           "tensorDivK (V div nuEff dev2(T(grad U)), fused)": 8 * F + 24 * F + 8 * F + 80 * N + 24 * N + 72 * B + 24 * B,
           "foldBoundaryDiagK": 16 * N + 4 * N + 24 * B + 4 * B, "reciprocalVK": 24 * N,
           "boundaryDiagTermK": 32 * B,
           "matrixH3K (HbyA, 3 components)": 16 * F + 8 * F + 3 * 24 * N + 16 * N + 3 * 16 * B,
           "constrainHbyAK": 32 * B, "vectorFluxK": 48 * F + 24 * N, "boundaryFluxK": 56 * B,
           "adjustPhi (sums, decision, scale)": 8 * F + 20 * B}
    kernels, total_ms = {}, 0.0
    for name, ev in events.items():
        per = len(ev) // (reps + 1)
        ms = [a.elapsed_time(b) for a, b in ev[per:]]   # the first round warms up
        avg = sum(ms) / len(ms)
        total_ms += sum(ms) / reps
        gbs = alg[name] / (avg * 1e-3) / 1e9
        kernels[name] = {"launches_per_step": per, "avg_us": round(avg * 1e3, 2), "alg_bytes": int(alg[name]),
                         "alg_GBps": round(gbs, 1), "frac": round(gbs / PEAK_HBM_GBS, 4)}
    log("pressure predictor glue: %.3f ms per step (%s)" % (total_ms, ", ".join(
        "%s %.1f us" % (k, v["avg_us"]) for k, v in kernels.items())))
    return {"ms_per_step": round(total_ms, 3), "mass_correction": massCorr[-1], "per_kernel": kernels,
            "workload": "UEqn relax, linearUpwind correction, viscous stress, rAU, HbyA (3 components), constrainHbyA, phiHbyA, adjustPhi on the %dx%dx%d "
                        "channel"
                        % ch.shape}


def batch_run(args):
    """Many small solves at once: --batch hierarchies of each of three graded, jittered channels (8,192,
    16,000 and 65,536 cells) on one context, the FP16 V(2,1) scaled cycle to 1e-6 from p = 0, solved in
    one mfv_amg_pcg_solve_batch launch (one workgroup per solve) and one after another on the per-grid
    kernels."""
    import torch
    from brae_amd import hexmesh, mfv
    ctx = mfv.Context(0)
    out = []
    ctl = dict(tolerance=1e-6, flexible=1, nPreSweeps=2, nPostSweeps=1, scaleCorrection=1)
    for nx, ny, nz in [(32, 16, 16), (40, 25, 16), (64, 32, 32)]:
        m = hexmesh.box_mesh(nx, ny, nz, (4.0, 1.0, 0.5), (1.0, 20.0, 20.0), 0.1, seed=3)
        nI, nC = m["nInternal"], m["nCells"]
        own, nei, bfc = (np.ascontiguousarray(a, np.int32) for a in (m["owner"][:nI], m["neighbour"], m["owner"][nI:]))
        # -laplacian(1) held positive on the mesh, the outlet fixed at 0, a seeded source
        dev = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float64)).to("cuda:0")
        dc = dev(m["deltaCoeffs"])
        magSf = dev(m["magSf"][:nI])
        meshes, amgs = [], []
        for _ in range(args.batch):
            meshes.append(mfv.Mesh(ctx, nC, own, nei, bfc))
        D, U = torch.zeros(nC, dtype=torch.float64, device="cuda:0"), torch.zeros(nI, dtype=torch.float64,
                                                                                    device="cuda:0")
        gam = -torch.ones(nI, dtype=torch.float64, device="cuda:0")
        ctx.call("mfv_laplacian_coeffs", meshes[0].h, mfv.ptr(dc), mfv.ptr(gam), mfv.ptr(magSf), mfv.ptr(D),
                 mfv.ptr(U), None)
        D[0] += D[0]   # a reference cell: the walls leave the matrix singular
        src = dev(np.random.default_rng(5).normal(size=nC) * m["V"])
        src -= src.mean()
        for mesh in meshes:
            a = mfv.Amg(mesh, np.asarray(m["magSf"][:nI]), mfv.AGG_BRAE, 64)
            a.galerkin(D, U, mfv.PREC_FP16)
            amgs.append(a)
        psis = [torch.zeros(nC, dtype=torch.float64, device="cuda:0") for _ in amgs]
        jobs = [a.solve_batch_job(D, src, p, precision=mfv.PREC_FP16, **ctl) for a, p in zip(amgs, psis)]
        mfv.solve_batch(ctx, jobs[:2])   # first use
        for p in psis:
            p.zero_()
        torch.cuda.synchronize()
        t = time.perf_counter()
        perfs = mfv.solve_batch(ctx, jobs)
        tb = time.perf_counter() - t
        amgs[0].solve(D, U, src, psis[0], precision=mfv.PREC_FP16, checkEvery=4, **ctl)   # first use
        torch.cuda.synchronize()
        t = time.perf_counter()
        its = []
        for a, p in zip(amgs, psis):
            p.zero_()
            its.append(a.solve(D, U, src, p, precision=mfv.PREC_FP16, checkEvery=4, **ctl).nIterations)
        torch.cuda.synchronize()
        tk = time.perf_counter() - t
        out.append(dict(cells=nC, solves=len(amgs), batch_ms=round(tb * 1e3, 1), per_grid_ms=round(tk * 1e3, 1),
                        iterations_batch=perfs[0].nIterations, iterations_per_grid=its[0],
                        max_final_residual=max(q.finalResidual for q in perfs)))
        log("batch: %d solves of %d cells: %.1f ms batched, %.1f ms one after another" % (len(amgs), nC, tb * 1e3,
                                                                                          tk * 1e3))
        for a in amgs:
            a.close()
        for mesh in meshes:
            mesh.close()
    ctx.close()
    return dict(workload="small FP16 AMG-PCG solves at once, batched (one workgroup each) vs the per-grid "
                         "kernels in turn", runs=out)


def serving_run(args):
    """The serving stream: Poisson arrivals at --serve-rate of pressure-Poisson requests on boxes of
    log-uniform 32k-4M cells, x-aspect uniform 1-8, 30% with 10%-jittered points, 1-20 right-hand sides
    each, admitted by a 200 GB HBM budget into --serve-slots concurrent slots; AMG-PCG with the mixed
    path's V-cycle precision (V(2,1), two pairwise passes per level) to 1e-6. Latency = finish - arrival."""
    from brae_amd import mfv
    rng = np.random.default_rng(2025)
    t, reqs = 0.0, []
    for k in range(args.serving):
        t += rng.exponential(1.0 / args.serve_rate)
        cells = float(np.exp(rng.uniform(np.log(32e3), np.log(4e6))))
        a = rng.uniform(1.0, 8.0)
        ny = max(2, int(round((cells / a) ** (1.0 / 3.0))))
        nx = max(2, int(round(a * ny)))
        jit = 0.1 if rng.uniform() < 0.3 else 0.0
        reqs.append(dict(nx=nx, ny=ny, nz=ny, lx=a, ly=1.0, lz=1.0, jitter=jit, seed=1000 + k,
                         nSteps=int(rng.integers(1, 21)), tolerance=1e-6, arrival=t))
    t0 = time.perf_counter()
    prec = 16 if args.precision == "fp16" else 32
    rc, res = mfv.serve(reqs, nSlots=args.serve_slots, hbmBudgetBytes=200e9, precision=prec,
                        residentBytes=getattr(args, "serve_resident", 0.0) * 1e9)
    wall = time.perf_counter() - t0
    if rc != mfv.MFV_OK:
        raise RuntimeError("serving run failed (%d)" % rc)
    lat = np.array([r.finished - r.arrival for r in res])
    wait = np.array([r.admitted - r.arrival for r in res])
    setup = np.array([r.built - r.admitted for r in res])
    solve = np.array([r.finished - r.built for r in res])
    solves = sum(r.nSteps for r in res)
    makespan = max(r.finished for r in res) - min(r.arrival for r in res)
    log("serving: %d requests, %d solves in %.1f s; latency p50 %.3f s p99 %.3f s" % (
        len(res), solves, makespan, np.percentile(lat, 50), np.percentile(lat, 99)))
    return dict(requests=len(res), solves=solves, cells_total=int(sum(r.nCells for r in res)),
                solves_per_s=round(solves / makespan, 2), p50_latency_s=round(float(np.percentile(lat, 50)), 4),
                p99_latency_s=round(float(np.percentile(lat, 99)), 4), makespan_s=round(makespan, 2),
                wall_s=round(wall, 2), arrival_rate_per_s=args.serve_rate, slots=args.serve_slots,
                hbm_budget_gb=200, max_residual=max(r.finalResidual for r in res),
                resident_gb=getattr(args, "serve_resident", 0.0), resident_hits=int(sum(r.resident for r in res)),
                latency_parts_s={name: dict(p50=round(float(np.percentile(v, 50)), 4),
                                            p99=round(float(np.percentile(v, 99)), 4))
                                 for name, v in (("admission_wait", wait), ("system_build", setup), ("solves", solve))},
                workload="pressure Poisson on hex boxes, log-uniform 32k-4M cells, x-aspect 1-8, 30%% jittered, "
                         "1-20 RHS per request, %s V-cycle AMG-PCG to 1e-6" % args.precision.upper())


def calibration_kernels(torch, n):
    """Kernels of known HBM traffic for reading a profiler's byte counters: a plain copy of n doubles
    (8n read, 8n written) and a gather of n doubles through an int64 index that reverses every run of 64
    elements (8n index + 8n source read, 8n written: every cache line of the source is used whole)."""
    src = torch.rand(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    idx = torch.arange(n, device="cuda", dtype=torch.int64)
    m = n // 64 * 64
    idx[:m] = idx[:m].view(-1, 64).flip(1).reshape(-1)
    torch.cuda.synchronize()
    for _ in range(3):
        dst.copy_(src)
    for _ in range(3):
        torch.index_select(src, 0, idx, out=dst)
    torch.cuda.synchronize()
    del src, dst, idx


def copy_peak(torch, n):
    """Measured device-to-device copy rate (read + write bytes / time) of n doubles: the plain copy kernel
    the HBM byte counters are calibrated against, and the box's practical bandwidth ceiling."""
    src = torch.rand(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 20
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del src, dst
    return 16.0 * n / (ms * 1e-3) / 1e9


def triad_peak(torch, n):
    """Measured streaming rate of c = a + s*b over n doubles (24 bytes per element / time): a plain
    elementwise stream, the box's practical HBM ceiling beside the published 8 TB/s."""
    a = torch.rand(n, dtype=torch.float64, device="cuda")
    b = torch.rand(n, dtype=torch.float64, device="cuda")
    c = torch.empty_like(a)
    for _ in range(3):
        torch.add(a, b, alpha=0.5, out=c)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 20
    e0.record()
    for _ in range(reps):
        torch.add(a, b, alpha=0.5, out=c)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del a, b, c
    return 24.0 * n / (ms * 1e-3) / 1e9


TRAFFIC_TABLE = os.path.join(ROOT, "profiles", "hbm_traffic_pcg_iteration_20M.md")


def measured_traffic(label):
    """HBM bytes per launch of `label` from the committed FETCH_SIZE/WRITE_SIZE counter table (rows
    '| label | ... | traffic_bytes |'), or None when the table does not list the kernel."""
    try:
        with open(TRAFFIC_TABLE) as f:
            for line in f:
                cells = [c.strip() for c in line.strip().strip("|").split("|")]
                if len(cells) >= 2 and cells[0] == label:
                    return int(float(cells[-1]))
    except OSError:
        pass
    return None


def cpu_info():
    """(logical CPUs, CPU model) of this host"""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return os.cpu_count() or 0, model


def cpu_sample(sample, nu, tol_u, reltol_u, tolerance, non_orth):
    """The oracle's FP64 turbulent SIMPLE iteration on a reduced channel: (seconds of the second of two
    iterations, its pressure iterations, cells). Module level so a process pool can run it."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    from brae_amd import hexmesh
    nx, ny, nz = (int(v) for v in sample.split("x"))
    m = hexmesh.box_mesh(nx, ny, nz, (4.0, 1.0, 0.5), (1.0, 20.0, 20.0), 0.1, seed=2024)
    nC, nI = m["nCells"], m["nInternal"]
    nB = len(m["owner"]) - nI
    mesh = O.Mesh(m["points"], np.arange(0, 4 * m["faces"].shape[0] + 1, 4), m["faces"].ravel(), m["owner"],
                  m["neighbour"])
    uT = np.full(nB, O.BC_FIXED_VALUE, np.int32)
    pT = np.full(nB, O.BC_ZERO_GRADIENT, np.int32)
    uV = np.zeros((nB, 3))
    isWall = np.zeros(nB, np.int32)
    kT = np.full(nB, O.BC_ZERO_GRADIENT, np.int32)
    kV, eV = np.zeros(nB), np.zeros(nB)
    for name, start, size in m["patches"]:
        s = slice(start - nI, start - nI + size)
        if name == "inlet":
            uV[s] = (10.0, 0.0, 0.0)
            kT[s] = O.BC_FIXED_VALUE
            kV[s], eV[s] = 0.375, 14.855
        elif name == "outlet":
            uT[s] = O.BC_ZERO_GRADIENT
            pT[s] = O.BC_FIXED_VALUE
        else:
            isWall[s] = 1
    rng = np.random.default_rng(7)
    U0 = rng.uniform(-1.0, 1.0, (nC, 3))
    U0[:, 0] += 10.0
    U = [U0[:, k].copy() for k in range(3)]
    Ub = [O.evaluate_boundary(mesh.bFaceCell, uT, uV[:, k], U[k]) for k in range(3)]
    own, nei = mesh.own(), mesh.neighbour
    w = mesh.weights[:, None]
    phi = ((w * U0[own] + (1 - w) * U0[nei]) * mesh.Sf[:nI]).sum(1)
    phiB = (np.stack(Ub, 1) * mesh.Sf[nI:]).sum(1)
    st = O.SimpleState(U, Ub, np.zeros(nC), np.zeros(nB), phi, phiB)
    k = np.full(nC, 0.375)
    e = np.full(nC, 14.855)
    nut = 0.09 * k * k / e
    kb = O.evaluate_boundary(mesh.bFaceCell, kT, kV, k)
    eb = O.evaluate_boundary(mesh.bFaceCell, kT, eV, e)
    nutb = 0.09 * kb * kb / np.where(eb > 0, eb, 1.0)
    y = O.near_wall_dist(mesh, isWall)
    kcs = O.kepsilon_case(mesh, isWall, y, kT, kV, kT, eV, nu=nu, relaxEps=0.7, relaxK=0.7, tol=tol_u,
                          relTol=reltol_u, maxIter=1000)
    times, iters = [], []
    for it in range(2):
        t = time.perf_counter()
        cs = O.simple_case(mesh, uT, uV.ravel(), pT, np.zeros(nB), nu=nu, tolU=tol_u, relTolU=reltol_u,
                           tolP=tolerance, pSolver=O.P_GAMG_REF, linearUpwind=1, correctedLaplacian=1,
                           nNonOrthCorr=non_orth, nuEff=nu + nut, nuEffB=nu + nutb)
        r = O.simple_step(cs, st)
        O.kepsilon_correct(kcs, st.U, st.Ub, st.phi, st.phiB, k, kb, e, eb, nut, nutb)
        times.append(time.perf_counter() - t)
        iters.append(list(r.pIters[:non_orth + 1]))
    return times[-1], iters[-1], nC


def cpu_baseline(args, n_full):
    """The oracle's FP64 SIMPLE iteration -- the reference CPU path: PBiCGStab-DILU momentum solves,
    gamg.cu's GAMG (agglomeratePair, GaussSeidel, dense coarsest) for the pressure, then the k-epsilon
    correct() with PBiCGStab-DILU solves -- on a reduced channel of the same construction and boundaries,
    from the same kind of seeded fields; one untimed iteration, then one timed, scaled linearly to the full
    cell count. `value` is one thread (the reference's host path has no threading; its Gauss-Seidel and
    DILU sweeps are sequential). `all_cores` runs one sample per host core at once (the box's CPU share, at
    most 16): the throughput a perfectly decomposed all-core run could reach, as s/iteration at full size."""
    import concurrent.futures
    import multiprocessing
    a = (args.cpu_sample, args.nu, args.tol_u, args.reltol_u, args.tolerance, args.non_orth)
    sec, iters, nC = cpu_sample(*a)
    ncpu, model = cpu_info()
    out = dict(value=round(sec * n_full / nC, 3), unit="s/SIMPLE-iter @ full size, linear in cells",
               cores=1, kind="port", host_cpus=ncpu, cpu_model=model,
               sample="oracle FP64 turbulent SIMPLE iteration (kEpsilon, wall functions; U PBiCGStab-DILU to %g "
                      "relTol %g; p gamg.cu GAMG to %g; k/epsilon PBiCGStab-DILU) on the %s channel (%d cells), one "
                      "untimed then one timed iteration: %.2f s (pressure solves %s); scaled x%.1f"
                      % (args.tol_u, args.reltol_u, args.tolerance, args.cpu_sample, nC, sec, iters, n_full / nC))
    try:
        P = max(1, min(16, len(os.sched_getaffinity(0))))
    except AttributeError:
        P = max(1, min(16, os.cpu_count() or 1))
    if P > 1:
        t = time.perf_counter()
        with concurrent.futures.ProcessPoolExecutor(P, mp_context=multiprocessing.get_context("spawn")) as ex:
            res = list(ex.map(cpu_sample, *[[x] * P for x in a]))
        wall = time.perf_counter() - t
        per = max(r[0] for r in res)
        out["all_cores"] = dict(
            value=round(per * n_full / (P * nC), 3), unit="s/SIMPLE-iter @ full size (aggregate of %d samples)" % P,
            cores=P, sample="%d copies of the sample at once, one per core: the slowest timed iteration %.2f s "
                            "(wall %.1f s with set-up); %d x %d cells per iteration" % (P, per, wall, P, nC))
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py runs on one GPU: this build has no multi-GPU path (each box has one MI355X); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    nx, ny, nz = (int(v) for v in args.size.split("x"))
    r = gpu_run(args)
    if args.profile_simple:
        print(json.dumps(dict(r["profile_simple"], size=args.size, v_cycle=args.precision)), flush=True)
        return
    c5 = c5_run(args) if args.c5 else None
    batch = batch_run(args) if args.batch > 0 else None
    serving = serving_run(args) if args.serving > 0 else None
    cpu = None if args.no_cpu_baseline else cpu_baseline(args, nx * ny * nz)
    sim = r.get("simple")
    pc = {
        "s_per_step": round(r["dt_mixed"], 5), "fp64_s_per_step": round(r["dt_fp64"], 5),
        "speedup_mixed_vs_fp64": round(r["dt_fp64"] / r["dt_mixed"], 3),
        "pcg_iterations": {"mixed": r["iters_mixed"], "fp64": r["iters_fp64"]},
        "final_residual": {"mixed": r["final_mixed"], "fp64": r["final_fp64"]},
        "p_rel_l2_mixed_vs_fp64": r["p_rel_l2_mixed_vs_fp64"], "other_mixed": r["other"],
        "workload": "the pressure corrector alone (interpolate(rAU), laplacian, Galerkin, %d x (grad(p), correction, "
                    "div(phiHbyA), fold, AMG-PCG to %g from p = 0 then the first solution), flux) on seeded "
                    "phiHbyA and rAU" % (args.non_orth + 1, args.tolerance)}
    if sim is not None:
        value, fp64 = sim["dt_mixed"], sim["dt_fp64"]
        unit = "s/SIMPLE-iter"
        workload = ("SIMPLE iteration (%s, nu %g; linearUpwind U, corrected laplacian, nNonOrthogonalCorrectors "
                    "%d; U%s PBiCGStab-diagonal to %g relTol %g; p AMG-PCG to %g), %s graded (ratio 20) jittered "
                    "(10%%) channel 4x1x0.5, %d cells"
                    % ("kEpsilon with nutk/epsilon wall functions, inlet k 0.375 epsilon 14.855, relax k/epsilon 0.7"
                       if args.turbulence else "laminar", args.nu, args.non_orth, ", k, epsilon" if args.turbulence
                       else "", args.tol_u, args.reltol_u, args.tolerance, args.size, nx * ny * nz))
        data = "synthetic (seeded perturbed plug flow U = (10,0,0) + uniform(-1,1), p = 0; %d timed iterations " \
               "from the seeded fields)" % args.steps
    else:
        value, fp64 = r["dt_mixed"], r["dt_fp64"]
        unit = "s/SIMPLE-iter (pressure corrector)"
        workload = "SIMPLE pressure corrector, %s channel, %d cells" % (args.size, nx * ny * nz)
        data = "synthetic (seeded perturbed plug flow on the graded jittered channel; p solved from 0)"
    line = {
        "metric": METRIC,
        "value": round(value, 5),
        "unit": unit,
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(value * 1e3, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 outer / %s V-cycle" % args.precision,
        "data": data,
        "config": {"workload": workload, "cells": nx * ny * nz, "tolerance": args.tolerance,
                   "v_cycle": args.precision},
        "fp64_s_per_iter": round(fp64, 5),
        "speedup_mixed_vs_fp64": round(fp64 / value, 3),
        "simple": None if sim is None else {
            "other_mixed_s_per_iter": None if "dt_other" not in sim else round(sim["dt_other"], 5),
            "p_iterations": sim["p_iterations"], "U_iterations": sim["U_iterations"],
            "final_p_residual": sim["final_p_residual"], "p_rel_l2_vs_fp64": sim["p_rel_l2_vs_fp64"],
            "U_rel_l2_vs_fp64": sim["U_rel_l2_vs_fp64"], "phases_ms": sim["phases_ms"],
            "ke_iterations": sim.get("ke_iterations"), "k_rel_l2_vs_fp64": sim.get("k_rel_l2_vs_fp64"),
            "eps_rel_l2_vs_fp64": sim.get("eps_rel_l2_vs_fp64"), "wall_cells": sim.get("wall_cells")},
        "pressure_corrector": pc,
        "amg": {"grids": r["n_grids"], "coarsest_cells": r["coarsest"], "vcycle_avg_us": r["vcycle_avg_us"],
                "mergeLevels": args.merge_levels,
                "cycle": "V(%d,%d) weighted Jacobi%s" % (args.npre, args.npost,
                                                        ", scaled correction" if args.scale else ""),
                "per_grid_us": r["grids"]},
        "openfoam_gamg_gpu": r.get("gamg_step"),
        "pressure_predictor": r.get("predictor"),
        "hbm_used_gib": r["hbm_used_gib"],
        "roofline": r["roofline"],
        "per_kernel": r["per_kernel"],
        "cpu_baseline": cpu,
        "serving": serving,
        "c5": c5,
        "batch": batch,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
