"""Benchmark: the SIMPLE pressure corrector on one MI355X.

Workload (one step = one SIMPLE iteration's pressure corrector on the 500x250x160 graded, jittered
channel, 20,000,000 cells, nNonOrthogonalCorrectors 1): rAUf = interpolate(rAU); assemble
-laplacian(rAUf, p) (face coefficients on nonOrthDeltaCoeffs, diagonal, fixedValue/zeroGradient boundary
coefficients); Galerkin re-coarsening; then twice: grad(p), the explicit non-orthogonal correction flux
and its divergence into the source, source -= div(phiHbyA)*V, fold, normFactor, AMG-preconditioned CG to
a normalised residual of 1e-6 (from p = 0, then from the first solution); finally
phi = phiHbyA - pEqn.flux() with the stored correction flux. The mixed-precision path runs the V-cycle
in FP16 with per-grid power-of-two scaling (--precision fp32: in FP32) inside FP64 flexible CG; the FP64
path runs the same solve with an FP64 cycle, and the other mixed precision is timed beside them.
`value` is the mixed path's seconds per step (lower is better); the FP64 time, the ratio, iterations,
residuals and the mixed/FP64 solution difference are reported beside it.

Inputs are synthetic and seeded: perturbed plug flow U = (10,0,0) + uniform(-1,1) per component
(inlet fixedValue (10,0,0), outlet zeroGradient, walls no-slip), phiHbyA = interpolate(U).Sf,
rAU = 0.05*V^(1/3)/|U|; p: fixedValue 0 at the outlet, zeroGradient elsewhere.
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
    ap.add_argument("--cpu-sample", default="250x125x80", help="channel the CPU oracle is timed on")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--check-every", type=int, default=4)
    ap.add_argument("--graph", type=int, default=0, help="1: replay captured groups of PCG iterations")
    ap.add_argument("--no-scale", action="store_true", help="plain injection instead of the scaled correction")
    # V(2,1) on levels of two composed pairwise passes: the fastest measured cycle for the FP64 path and
    # within 1 % of the fastest for the FP16 one on the 20 M channel
    ap.add_argument("--npre", type=int, default=2)
    ap.add_argument("--npost", type=int, default=1)
    ap.add_argument("--merge-levels", type=int, default=2, help="pairwise passes per AMG level")
    ap.add_argument("--non-orth", type=int, default=1, help="nNonOrthogonalCorrectors (solves per step - 1)")
    ap.add_argument("--serving", type=int, default=512, help="requests of the serving stream (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=8.0, help="Poisson arrival rate, requests/s")
    ap.add_argument("--serve-slots", type=int, default=8)
    ap.add_argument("--predictor", type=int, default=1,
                    help="1: also time the pressure predictor's glue (rAU, HbyA, constrainHbyA, phiHbyA, adjustPhi)")
    ap.add_argument("--gamg-step", type=int, default=0,
                    help="1: also time the step with OpenFOAM's GAMG-PCG on the device (the CPU baseline's solver)")
    return ap.parse_args()


class Channel:
    """The graded, jittered channel and its synthetic pressure-corrector inputs on the device."""

    def __init__(self, nx, ny, nz, torch, mfv, hexmesh, ctx):
        t0 = time.time()
        self.shape = (nx, ny, nz)
        pts = hexmesh.box_points(nx, ny, nz, (4.0, 1.0, 0.5), (1.0, 20.0, 20.0), 0.1, seed=2024)
        faces, owner, nei, patches = hexmesh.box_topology(nx, ny, nz)
        self.nC, self.nI = nx * ny * nz, nei.size
        self.nB = owner.size - self.nI
        self.patches = patches
        log("mesh %dx%dx%d: %d cells, %d internal faces (%.1fs)" % (nx, ny, nz, self.nC, self.nI, time.time() - t0))
        self.mesh = mfv.Mesh(ctx, self.nC, np.ascontiguousarray(owner[:self.nI]), nei,
                             np.ascontiguousarray(owner[self.nI:]))
        dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")
        z = lambda n: torch.zeros(int(n), dtype=torch.float64, device="cuda:0")
        nF = owner.size
        dpts = dev(pts)
        dfo = dev(np.arange(0, 4 * nF + 1, 4, dtype=np.int32))
        dfv = dev(faces.reshape(-1))
        self.Cf, self.Sf, self.magSf = z(3 * nF), z(3 * nF), z(nF)
        self.C, self.V = z(3 * self.nC), z(self.nC)
        self.w, self.dc, self.ndc, self.corr, self.bDelta = z(self.nI), z(self.nI), z(self.nI), z(3 * self.nI), z(self.nB)
        torch.cuda.synchronize()
        ctx.call("mfv_mesh_geometry", self.mesh.h, *[mfv.ptr(t) for t in (
            dpts, dfo, dfv, self.Cf, self.Sf, self.magSf, self.C, self.V, self.w, self.dc, self.ndc, self.corr,
            self.bDelta)])
        ctx.sync()
        del dpts, dfo, dfv
        log("geometry on device (%.1fs)" % (time.time() - t0))
        # synthetic seeded fields
        g = torch.Generator(device="cuda:0")
        g.manual_seed(7)
        U = torch.rand((self.nC, 3), dtype=torch.float64, device="cuda:0", generator=g) * 2.0 - 1.0
        U[:, 0] += 10.0
        own_d, nei_d = dev(owner[:self.nI]).long(), dev(nei).long()
        Sf = self.Sf.view(nF, 3)
        wv = self.w.view(-1, 1)
        Uf = wv * U[own_d] + (1.0 - wv) * U[nei_d]
        self.phiHbyA = (Uf * Sf[:self.nI]).sum(1).contiguous()
        phiB = torch.zeros(self.nB, dtype=torch.float64, device="cuda:0")
        bcell = dev(owner[self.nI:]).long()
        self.types = torch.zeros(self.nB, dtype=torch.int32, device="cuda:0")
        for name, start, size in patches:
            s = slice(start - self.nI, start - self.nI + size)
            if name == "inlet":
                phiB[s] = 10.0 * Sf[start:start + size, 0]
            elif name == "outlet":
                phiB[s] = (U[bcell[s]] * Sf[start:start + size]).sum(1)
                self.types[s] = 1
        self.phiB = phiB.contiguous()
        self.bMagSf = self.magSf[self.nI:].contiguous()
        self.magSfI = self.magSf[:self.nI].contiguous()
        Umag = U.norm(dim=1)
        self.negRAU = (-0.05 * self.V.pow(1.0 / 3.0) / Umag).contiguous()
        self.bcValue = z(self.nB)
        del U, Uf, own_d, nei_d, bcell
        # work arrays
        self.gf, self.gb = z(self.nI), z(self.nB)
        self.diag, self.upper = z(self.nC), z(self.nI)
        self.iC, self.bC = z(self.nB), z(self.nB)
        self.source, self.diagC, self.b = z(self.nC), z(self.nC), z(self.nC)
        self.p, self.phi = z(self.nC), z(self.nI)
        self.gradP, self.ffc = z(3 * self.nC), z(self.nI)
        torch.cuda.synchronize()
        log("fields ready (%.1fs)" % (time.time() - t0))


def gpu_run(args):
    import torch
    from brae_amd import hexmesh, mfv

    nx, ny, nz = (int(v) for v in args.size.split("x"))
    # one stream (default priority) for torch's own work and the library's, so the two stay in order
    stream = torch.cuda.Stream(device="cuda:0")
    torch.cuda.set_stream(stream)
    ctx = mfv.Context(0, stream=stream.cuda_stream)
    ch = Channel(nx, ny, nz, torch, mfv, hexmesh, ctx)
    t0 = time.time()
    w_host = ch.magSfI.cpu().numpy()
    amg = mfv.Amg(ch.mesh, w_host, mfv.AGG_BRAE, 64, mergeLevels=args.merge_levels)
    sizes = [amg.grid_size(g) for g in range(amg.nGrids)]
    log("AMG hierarchy: %d grids, coarsest %d cells (%.1fs host build)" % (amg.nGrids, sizes[-1][0], time.time() - t0))
    P = mfv.ptr
    A = mfv.ldu(ch.diagC, ch.upper)
    prec_mixed = mfv.PREC_FP32 if args.precision == "fp32" else mfv.PREC_FP16

    def step(prec, gamg=None):
        """one SIMPLE pressure corrector: nNonOrthogonalCorrectors + 1 assemblies and solves of the same
        matrix, the explicit non-orthogonal correction rebuilt from the latest p before each, and the flux
        (with the stored correction flux) after the last. With `gamg` the solve is OpenFOAM's PCG with
        the GAMG preconditioner (GaussSeidel, nVcycles 2) on that hierarchy, as the CPU path solves it."""
        ch.p.zero_()
        ctx.call("mfv_interpolate", ch.mesh.h, P(ch.w), P(ch.negRAU), P(ch.gf), P(ch.gb))
        ctx.call("mfv_laplacian_coeffs", ch.mesh.h, P(ch.ndc), P(ch.gf), P(ch.magSfI), P(ch.diag), P(ch.upper), None)
        ctx.call("mfv_laplacian_bcoeffs", ch.mesh.h, P(ch.types), P(ch.bcValue), P(ch.gb), P(ch.bMagSf),
                 P(ch.bDelta), P(ch.iC), P(ch.bC))
        perfs = []
        for corr in range(args.non_orth + 1):
            ch.source.zero_()
            ctx.call("mfv_grad_gauss", ch.mesh.h, P(ch.types), P(ch.bcValue), P(ch.Sf), P(ch.w), P(ch.V), P(ch.p),
                     P(ch.gradP))
            ctx.call("mfv_laplacian_corr", ch.mesh.h, P(ch.gradP), P(ch.w), P(ch.corr), P(ch.gf), P(ch.magSfI),
                     P(ch.ffc), P(ch.source))
            ctx.call("mfv_div_source", ch.mesh.h, P(ch.phiHbyA), P(ch.phiB), P(ch.V), -1.0, P(ch.source))
            ctx.call("mfv_fold", ch.mesh.h, P(ch.diag), P(ch.source), P(ch.iC), P(ch.bC), P(ch.diagC), P(ch.b))
            if gamg is not None:
                perfs.append(gamg.pcg_gamg_solve(ch.diagC, ch.upper, ch.b, ch.p, args.tolerance,
                                                 smoother="GaussSeidel"))
                continue
            if corr == 0:
                amg.galerkin(ch.diagC, ch.upper, prec)
            perfs.append(amg.solve(ch.diagC, ch.upper, ch.b, ch.p, tolerance=args.tolerance, precision=prec,
                                   flexible=int(prec != mfv.PREC_FP64 or not args.no_scale),
                                   checkEvery=args.check_every, scaleCorrection=int(not args.no_scale),
                                   nPreSweeps=args.npre, nPostSweeps=args.npost, captureVcycle=args.graph))
        ctx.call("mfv_flux_correct", ch.mesh.h, ctypes.byref(A), P(ch.p), P(ch.ffc), P(ch.phiHbyA), -1.0, P(ch.phi))
        return perfs

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
    other = mfv.PREC_FP16 if prec_mixed == mfv.PREC_FP32 else mfv.PREC_FP32
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
        mfv.PROF_AMG_CORR_SMOOTH0: ("corrSmoothK grid 0", 5 * zb * N + dw * N, iters),
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
        hbm_used_gib=round(hbm_used_gib, 2), other=other_out,
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
    rc, res = mfv.serve(reqs, nSlots=args.serve_slots, hbmBudgetBytes=200e9, precision=prec)
    wall = time.perf_counter() - t0
    if rc != mfv.MFV_OK:
        raise RuntimeError("serving run failed (%d)" % rc)
    lat = np.array([r.finished - r.arrival for r in res])
    solves = sum(r.nSteps for r in res)
    makespan = max(r.finished for r in res) - min(r.arrival for r in res)
    log("serving: %d requests, %d solves in %.1f s; latency p50 %.3f s p99 %.3f s" % (
        len(res), solves, makespan, np.percentile(lat, 50), np.percentile(lat, 99)))
    return dict(requests=len(res), solves=solves, cells_total=int(sum(r.nCells for r in res)),
                solves_per_s=round(solves / makespan, 2), p50_latency_s=round(float(np.percentile(lat, 50)), 4),
                p99_latency_s=round(float(np.percentile(lat, 99)), 4), makespan_s=round(makespan, 2),
                wall_s=round(wall, 2), arrival_rate_per_s=args.serve_rate, slots=args.serve_slots,
                hbm_budget_gb=200, max_residual=max(r.finalResidual for r in res),
                workload="pressure Poisson on hex boxes, log-uniform 32k-4M cells, x-aspect 1-8, 30%% jittered, "
                         "1-20 RHS per request, %s V-cycle AMG-PCG to 1e-6" % args.precision.upper())


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


TRAFFIC_TABLE = os.path.join(ROOT, "profiles", "hbm_traffic_pressure_corrector_20M.md")


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


def cpu_baseline(args, n_full):
    """The oracle's FP64 pressure corrector on a reduced channel of the same construction,
    single-threaded, scaled to the full cell count: the same passes as the GPU step (assembly, then per
    non-orthogonal pass grad(p), correction flux and source, div(phiHbyA), fold and OpenFOAM's
    GAMG-preconditioned PCG (faceAreaPair, GaussSeidel, V(0,2,2), nVcycles 2), then the flux)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    from brae_amd import hexmesh
    nx, ny, nz = (int(v) for v in args.cpu_sample.split("x"))
    m = hexmesh.box_mesh(nx, ny, nz, (4.0, 1.0, 0.5), (1.0, 20.0, 20.0), 0.1, seed=2024)
    nC, nI = m["nCells"], m["nInternal"]
    nB = m["faces"].shape[0] - nI
    rng = np.random.default_rng(7)
    U = rng.uniform(-1.0, 1.0, (nC, 3))
    U[:, 0] += 10.0
    own, nei = np.ascontiguousarray(m["owner"][:nI]), m["neighbour"]
    w = m["weights"][:, None]
    phi = ((w * U[own] + (1 - w) * U[nei]) * m["Sf"][:nI]).sum(1)
    phiB = np.zeros(nB)
    types = np.zeros(nB, np.int32)
    bcell = np.ascontiguousarray(m["owner"][nI:])
    for name, start, size in m["patches"]:
        s = slice(start - nI, start - nI + size)
        if name == "inlet":
            phiB[s] = 10.0 * m["Sf"][start:start + size, 0]
        elif name == "outlet":
            phiB[s] = (U[bcell[s]] * m["Sf"][start:start + size]).sum(1)
            types[s] = 1
    negRAU = -0.05 * np.cbrt(m["V"]) / np.linalg.norm(U, axis=1)
    wts = O.faceareapair_weights(m["Sf"][:nI], m["magSf"][:nI])
    ag = O.of_agglomeration(nC, own, nei, wts, nCellsInCoarsestLevel=10)
    t = time.perf_counter()
    gf = m["weights"] * negRAU[own] + (1 - m["weights"]) * negRAU[nei]
    diag, up = np.zeros(nC), np.zeros(nI)
    O.lib().orc_laplacian(nC, nI, O.p(own), O.p(nei), O.p(m["nonOrthDeltaCoeffs"]), O.p(gf),
                          O.p(O.f64(m["magSf"][:nI])), O.p(diag), O.p(up), None)
    iC, bC = np.zeros(nB), np.zeros(nB)
    O.lib().orc_laplacian_bcoeffs(nB, O.p(types), O.p(np.zeros(nB)), O.p(O.f64(negRAU[bcell])),
                                  O.p(O.f64(m["magSf"][nI:])), O.p(O.f64(m["bDeltaCoeffs"])), O.p(iC), O.p(bC))
    div = np.zeros(nC)
    np.add.at(div, own, phi)
    np.add.at(div, nei, -phi)
    np.add.at(div, bcell, phiB)
    psi = np.zeros(nC)
    iters = []
    for _ in range(args.non_orth + 1):
        g = O.grad_gauss(nC, own, nei, bcell, types, np.zeros(nB), m["Sf"], m["weights"], m["V"], psi)
        ffc, src = O.laplacian_corr(nC, own, nei, g, m["weights"], m["corrVec"], gf, m["magSf"][:nI],
                                    source=np.zeros(nC))
        src = src - (div / m["V"]) * m["V"]
        diagC, b = np.zeros(nC), np.zeros(nC)
        O.lib().orc_fold(nC, O.p(diag), O.p(src), nB, O.p(bcell), O.p(types), O.p(iC), O.p(bC), O.p(diagC), O.p(b))
        psi, perf = O.pcg_gamg_solve(ag, diagC, up, b, psi, args.tolerance, smoother="GaussSeidel")
        iters.append(perf.nIterations)
    O.flux_correct(own, nei, up, None, psi, ffc, phi)
    sec = time.perf_counter() - t
    scaled = sec * n_full / nC
    return dict(value=round(scaled, 3), unit="s/SIMPLE-iter (pressure corrector) @ full size, linear in cells",
                cores=1, kind="port",
                sample="oracle FP64 pressure corrector (assembly, %d non-orthogonal passes of grad/correction/"
                       "fold/OpenFOAM GAMG-PCG GaussSeidel V(0,2,2) nVcycles 2 to %g, flux) on the %s channel "
                       "(%d cells): %.2f s, PCG iterations %s; scaled x%.1f (hierarchy build excluded)"
                       % (args.non_orth + 1, args.tolerance, args.cpu_sample, nC, sec, iters, n_full / nC))


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py runs on one GPU: this build has no multi-GPU path (each box has one MI355X); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    nx, ny, nz = (int(v) for v in args.size.split("x"))
    r = gpu_run(args)
    serving = serving_run(args) if args.serving > 0 else None
    cpu = None if args.no_cpu_baseline else cpu_baseline(args, nx * ny * nz)
    line = {
        "metric": METRIC,
        "value": round(r["dt_mixed"], 5),
        "unit": "s/SIMPLE-iter (pressure corrector)",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["dt_mixed"] * 1e3, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 outer / %s V-cycle" % args.precision,
        "data": "synthetic (seeded perturbed plug flow on the graded jittered channel; p solved from 0)",
        "config": {"workload": "SIMPLE pressure corrector (nNonOrthogonalCorrectors %d), %s graded (ratio 20) "
                               "jittered (10%%) channel 4x1x0.5, %d cells, AMG-PCG to %g"
                               % (args.non_orth, args.size, nx * ny * nz, args.tolerance),
                   "cells": nx * ny * nz, "tolerance": args.tolerance, "v_cycle": args.precision},
        "fp64_s_per_iter": round(r["dt_fp64"], 5),
        "speedup_mixed_vs_fp64": round(r["dt_fp64"] / r["dt_mixed"], 3),
        "pcg_iterations": {"mixed": r["iters_mixed"], "fp64": r["iters_fp64"]},
        "final_residual": {"mixed": r["final_mixed"], "fp64": r["final_fp64"]},
        "p_rel_l2_mixed_vs_fp64": r["p_rel_l2_mixed_vs_fp64"],
        "amg": {"grids": r["n_grids"], "coarsest_cells": r["coarsest"], "vcycle_avg_us": r["vcycle_avg_us"],
                "mergeLevels": args.merge_levels,
                "cycle": "V(%d,%d) weighted Jacobi%s" % (args.npre, args.npost,
                                                        "" if args.no_scale else ", scaled correction"),
                "per_grid_us": r["grids"]},
        "other_mixed": r["other"],
        "openfoam_gamg_gpu": r.get("gamg_step"),
        "pressure_predictor": r.get("predictor"),
        "hbm_used_gib": r["hbm_used_gib"],
        "roofline": r["roofline"],
        "per_kernel": r["per_kernel"],
        "cpu_baseline": cpu,
        "serving": serving,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
