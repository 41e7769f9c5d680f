"""The workload configurations at the size one GPU runs them, checked through size-independent properties.

- 20 M-cell channel (500 x 250 x 160, graded 20, jittered 10 %): the SIMPLE pressure corrector with
  nNonOrthogonalCorrectors 1 in FP64, FP32 and FP16 V-cycles. Each final normalised residual, recomputed
  here in FP64 from mfv_amul and mfv_norm_factor (sum|b - A p| / normFactor), is below the 1e-6 tolerance;
  the FP32 and FP16 pressures differ from the FP64 one by less than 1e-4 in relative L2 (all three stop
  anywhere below 1e-6); iteration counts are printed.
- 2.1 M-cell channel (256 x 128 x 64, kEpsilon with wall functions): three turbulent SIMPLE iterations with the
  FP16 and the FP64 pressure cycle from the same seeded fields; U, p, k and epsilon agree within 1e-3 relative
  L2 and every pressure solve reaches 1e-6.
- 120 M-cell channel (800 x 400 x 375, the largest single-GPU configuration): one turbulent SIMPLE iteration with
  the FP32 cycle; the pressure solves reach 1e-6, the fields stay finite, the HBM in use fits the card.
- Serving at the stream's sizes: 16 requests of log-uniform 32 k - 4 M cells (aspect 1-8, 30 % jittered,
  1-20 steps) with the FP32 and the FP16 cycle; each request's fingerprint is the same alone and in the mix,
  and every residual is below 1e-6.
"""
import numpy as np
import pytest

from brae_amd import channel as C
from brae_amd import hexmesh, mfv

pytestmark = pytest.mark.gpu


def torch_cuda():
    import torch
    assert torch.cuda.is_available()
    return torch


def test_pressure_corrector_20M_three_precisions():
    torch = torch_cuda()
    ctx = mfv.Context(0)
    print("case: pressure corrector 500x250x160", flush=True)
    ch = C.Channel(500, 250, 160, torch, mfv, hexmesh, ctx, log=print)
    amg = mfv.Amg(ch.mesh, ch.magSfI.cpu().numpy(), mfv.AGG_BRAE, 64, mergeLevels=2)
    out = {}
    for prec in (mfv.PREC_FP64, mfv.PREC_FP32, mfv.PREC_FP16):
        perfs = C.pressure_corrector(ch, ctx, mfv, amg, prec, tolerance=1e-6, nonOrth=1)
        torch.cuda.synchronize()
        # the final residual recomputed in FP64 from the library's product and normFactor
        A = mfv.ldu(ch.diagC, ch.upper)
        Ap = torch.zeros_like(ch.p)
        nf = torch.zeros(1, dtype=torch.float64, device="cuda:0")
        torch.cuda.synchronize()
        import ctypes
        ctx.call("mfv_amul", ch.mesh.h, ctypes.byref(A), mfv.ptr(ch.p), mfv.ptr(Ap))
        ctx.call("mfv_norm_factor", ch.mesh.h, ctypes.byref(A), mfv.ptr(ch.p), mfv.ptr(ch.b), mfv.ptr(nf))
        ctx.sync()
        res = float((ch.b - Ap).abs().sum() / nf[0])
        print("precision", prec, "iterations", [q.nIterations for q in perfs], "final", perfs[-1].finalResidual,
              "recomputed", res, flush=True)
        assert all(q.finalResidual < 1e-6 for q in perfs)
        assert res < 1e-6
        out[prec] = ch.p.clone()
    ref = out[mfv.PREC_FP64]
    for prec in (mfv.PREC_FP32, mfv.PREC_FP16):
        d = float((out[prec] - ref).norm() / ref.norm())
        print("p rel L2 vs FP64", prec, d, flush=True)
        assert d < 1e-4, (prec, d)
    amg.close()
    ch.mesh.close()
    ctx.close()


def test_turbulent_channel_2M_mixed_against_fp64():
    torch = torch_cuda()
    ctx = mfv.Context(0)
    print("case: turbulent SIMPLE 256x128x64", flush=True)
    ch = C.Channel(256, 128, 64, torch, mfv, hexmesh, ctx, log=print)
    uT, uV, pT, pV = C.channel_bcs(ch, torch)
    turb = C.channel_kepsilon(ch, torch, mfv, tol=1e-5, relTol=0.1, log=print)
    geom = mfv.Geometry(*(mfv.ptr(t) for t in (ch.Sf, ch.Cf, ch.magSf, ch.C, ch.V, ch.w, ch.dc, ch.ndc, ch.corr,
                                                ch.bDelta)))
    results = {}
    for prec in (mfv.PREC_FP64, mfv.PREC_FP16):
        ctl = mfv.simple_controls(U=dict(tolerance=1e-5, relTol=0.1, preconditioner=mfv.PRECOND_DIAGONAL),
                                  p=dict(tolerance=1e-6, relTol=0.0, precision=prec, nPreSweeps=2, nPostSweeps=1,
                                         flexible=1, scaleCorrection=1),
                                  mergeLevels=2, nu=1e-5, linearUpwind=1, correctedLaplacian=1, nNonOrthCorr=1)
        sim = mfv.Simple(ch.mesh, geom, uT, uV.ravel(), pT, pV, ctl)
        ke = mfv.KEpsilon(ch.mesh, geom, turb["isWall"], turb["y"], turb["type"], turb["kV"], turb["type"],
                          turb["eV"], turb["controls"])
        f = C.simple_fields(ch, torch, uT, uV)
        f.update({k: v.clone() for k, v in turb["fields0"].items()})
        for it in range(3):
            r, kr = mfv.simple_step_kepsilon(sim, ke, f["U"], f["Ub"], f["p"], f["pb"], f["phi"], f["phiB"], f["k"],
                                             f["kb"], f["eps"], f["epsb"], f["nut"], f["nutb"])
            print("precision", prec, "iteration", it, "p", list(r.pIters[:2]), "U", list(r.UIters), "ke",
                  kr.epsIters, kr.kIters, "wall cells", kr.nWallCells, flush=True)
            assert r.pFinal < 1e-6 and kr.nWallCells > 0
        results[prec] = {k: (torch.stack(f[k]) if k == "U" else f[k]).clone() for k in ("U", "p", "k", "eps")}
        sim.close()
        ke.close()
    for k in ("U", "p", "k", "eps"):
        a, b = results[mfv.PREC_FP16][k], results[mfv.PREC_FP64][k]
        d = float((a - b).norm() / b.norm())
        print(k, "rel L2 FP16 vs FP64", d, flush=True)
        assert d < 1e-3, (k, d)
    ch.mesh.close()
    ctx.close()


def release(torch):
    """the previous tests' device memory back to the device: their objects collected, torch's cache emptied"""
    import gc
    gc.collect()
    torch.cuda.synchronize()
    torch.cuda.empty_cache()


def test_turbulent_channel_120M_fp32():
    """the largest single-GPU configuration: one turbulent SIMPLE iteration on 800 x 400 x 375 (120 M cells) with
    the FP32 cycle; every pressure solve reaches 1e-6, the fields stay finite, and the HBM the case takes (the
    device's in-use memory over what was in use before it, the library's allocations included) fits the card"""
    torch = torch_cuda()
    release(torch)
    free_b, total_b = torch.cuda.mem_get_info()
    before = (total_b - free_b) / 2 ** 30
    ctx = mfv.Context(0)
    print("case: turbulent SIMPLE 800x400x375 FP32 (%.1f GiB in use before)" % before, flush=True)
    objs = {}
    try:
        ch = objs["ch"] = C.Channel(800, 400, 375, torch, mfv, hexmesh, ctx, log=print, corrector=False)
        uT, uV, pT, pV = C.channel_bcs(ch, torch)
        turb = objs["turb"] = C.channel_kepsilon(ch, torch, mfv, tol=1e-5, relTol=0.1, log=print)
        geom = mfv.Geometry(*(mfv.ptr(t) for t in (ch.Sf, ch.Cf, ch.magSf, ch.C, ch.V, ch.w, ch.dc, ch.ndc, ch.corr,
                                                    ch.bDelta)))
        ctl = mfv.simple_controls(U=dict(tolerance=1e-5, relTol=0.1, preconditioner=mfv.PRECOND_DIAGONAL),
                                  p=dict(tolerance=1e-6, relTol=0.0, precision=mfv.PREC_FP32, nPreSweeps=1,
                                         nPostSweeps=1, flexible=1, scaleCorrection=0),
                                  mergeLevels=2, nu=1e-5, linearUpwind=1, correctedLaplacian=1, nNonOrthCorr=1)
        sim = objs["sim"] = mfv.Simple(ch.mesh, geom, uT, uV.ravel(), pT, pV, ctl)
        ke = objs["ke"] = mfv.KEpsilon(ch.mesh, geom, turb["isWall"], turb["y"], turb["type"], turb["kV"],
                                       turb["type"], turb["eV"], turb["controls"])
        f = objs["f"] = C.simple_fields(ch, torch, uT, uV)
        f.update({k: v.clone() for k, v in turb["fields0"].items()})
        r, kr = mfv.simple_step_kepsilon(sim, ke, f["U"], f["Ub"], f["p"], f["pb"], f["phi"], f["phiB"], f["k"],
                                         f["kb"], f["eps"], f["epsb"], f["nut"], f["nutb"])
        torch.cuda.synchronize()
        free_b, total_b = torch.cuda.mem_get_info()
        used = (total_b - free_b) / 2 ** 30 - before
        print("p iterations", list(r.pIters[:2]), "final", r.pFinal, "U", list(r.UIters), "k-epsilon", kr.epsIters,
              kr.kIters, "HBM taken by the case %.1f GiB" % used, flush=True)
        assert r.pFinal < 1e-6 and kr.nWallCells > 0
        for name in ("p", "k", "eps", "nut"):
            assert bool(torch.isfinite(f[name]).all()), name
        assert all(bool(torch.isfinite(u).all()) for u in f["U"])
        assert used < 250.0   # of the card's 288 GB (268 GiB)
    finally:
        for k in ("sim", "ke"):
            if k in objs:
                objs[k].close()
        if "ch" in objs:
            objs["ch"].mesh.close()
        objs.clear()
        ctx.close()
        release(torch)


def serving_stream(n, seed=2025):
    """n requests of the serving stream's construction (bench.py serving_run), arriving at once"""
    rng = np.random.default_rng(seed)
    reqs = []
    for k in range(n):
        cells = float(np.exp(rng.uniform(np.log(32e3), np.log(4e6))))
        a = rng.uniform(1.0, 8.0)
        ny = max(2, int(round((cells / a) ** (1.0 / 3.0))))
        nx = max(2, int(round(a * ny)))
        jit = 0.1 if rng.uniform() < 0.3 else 0.0
        reqs.append(dict(nx=nx, ny=ny, nz=ny, lx=a, ly=1.0, lz=1.0, jitter=jit, seed=1000 + k,
                         nSteps=int(rng.integers(1, 21)), tolerance=1e-6, arrival=0.0))
    return reqs


@pytest.mark.parametrize("precision", [32, 16])
def test_serving_at_stream_sizes(precision):
    reqs = serving_stream(16)
    sizes = [r["nx"] * r["ny"] * r["nz"] for r in reqs]
    print("case: serving 16 requests, cells", min(sizes), "-", max(sizes), "precision", precision, flush=True)
    assert min(sizes) >= 25000 and max(sizes) <= 5000000
    rc, mixed = mfv.serve(reqs, nSlots=8, hbmBudgetBytes=200e9, precision=precision)
    assert rc == mfv.MFV_OK
    for r in mixed:
        assert r.status == 0 and r.finalResidual < 1e-6
    for i in (0, 5, 11):
        rc, alone = mfv.serve([reqs[i]], nSlots=1, hbmBudgetBytes=200e9, precision=precision)
        assert rc == mfv.MFV_OK
        assert alone[0].fingerprint == mixed[i].fingerprint, i
