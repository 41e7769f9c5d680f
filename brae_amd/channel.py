"""The graded, jittered channel of the pressure and SIMPLE workloads on the device: 4 x 1 x 0.5, two-sided wall
grading ratio 20 on y and z, interior points jittered by +-10 % of the local spacing (seed 2024); inlet (x-min)
U = (10, 0, 0), outlet (x-max) zeroGradient with p = 0, the four other faces walls. Built by the bench and
the GPU tests at the configurations' sizes (20 M cells: 500 x 250 x 160; 2.1 M: 256 x 128 x 64).

Channel           the mesh, its geometry (mfv_mesh_geometry) and, unless built for SIMPLE alone, synthetic
                  pressure-corrector inputs (seeded perturbed plug flow U = (10, 0, 0) + uniform(-1, 1):
                  phiHbyA = interpolate(U).Sf, rAU) with their work arrays
pressure_corrector one SIMPLE pressure corrector on those inputs: nNonOrthogonalCorrectors + 1 assemblies and
                  AMG-PCG solves, then the flux
channel_bcs       U and p boundary conditions per boundary face
simple_fields     seeded initial fields of a SIMPLE run
channel_kepsilon  the k-epsilon set-up (wall functions on the four walls, inlet k and epsilon)
"""
import ctypes
import time

import numpy as np


class Channel:
    """The graded, jittered channel and its synthetic pressure-corrector inputs on the device. With
    corrector=False only the mesh and its geometry are built (what a SIMPLE run reads): at 120 M cells that
    leaves out about 24 GiB of inputs and work arrays. Torch's cached blocks from the set-up's temporaries
    are handed back to the device at the end, so the library's own allocations can use them."""

    def __init__(self, nx, ny, nz, torch, mfv, hexmesh, ctx, log=None, corrector=True):
        log = log or (lambda *a: None)
        t0 = time.time()
        self.shape = (nx, ny, nz)
        pts = hexmesh.box_points(nx, ny, nz, (4.0, 1.0, 0.5), (1.0, 20.0, 20.0), 0.1, seed=2024)
        faces, owner, nei, patches = hexmesh.box_topology(nx, ny, nz)
        self.points, self.faces = pts, faces
        self.nC, self.nI = nx * ny * nz, nei.size
        self.nB = owner.size - self.nI
        self.patches = patches
        log("mesh %dx%dx%d: %d cells, %d internal faces (%.1fs)" % (nx, ny, nz, self.nC, self.nI, time.time() - t0))
        self.own = np.ascontiguousarray(owner[:self.nI])
        self.nei = nei
        self.bcell = np.ascontiguousarray(owner[self.nI:])
        self.mesh = mfv.Mesh(ctx, self.nC, self.own, nei, self.bcell)
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
        self.corrector = corrector
        if not corrector:
            torch.cuda.empty_cache()
            return
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
        torch.cuda.empty_cache()
        log("fields ready (%.1fs)" % (time.time() - t0))


def channel_bcs(ch, torch):
    """The channel's boundary conditions per boundary face: inlet U fixedValue (10,0,0) with p
    zeroGradient; outlet U zeroGradient with p fixedValue 0; the four walls noSlip with p zeroGradient."""
    from brae_amd import mfv
    nB, nI = ch.nB, ch.nI
    uT = np.full(nB, mfv.BC_FIXED_VALUE, np.int32)
    pT = np.full(nB, mfv.BC_ZERO_GRADIENT, np.int32)
    uV = np.zeros((nB, 3))
    pV = np.zeros(nB)
    for name, start, size in ch.patches:
        s = slice(start - nI, start - nI + size)
        if name == "inlet":
            uV[s] = (10.0, 0.0, 0.0)
        elif name == "outlet":
            uT[s] = mfv.BC_ZERO_GRADIENT
            pT[s] = mfv.BC_FIXED_VALUE
    return uT, uV, pT, pV


def simple_fields(ch, torch, uT, uV, seed=7):
    """Seeded initial fields: U = (10,0,0) + uniform(-1,1) per component, U's patch values from the
    boundary conditions, phi = interpolate(U).Sf inside and Ub.Sf on the boundary, p = 0."""
    g = torch.Generator(device="cuda:0")
    g.manual_seed(seed)
    nC, nI, nB = ch.nC, ch.nI, ch.nB
    U = torch.rand((nC, 3), dtype=torch.float64, device="cuda:0", generator=g) * 2.0 - 1.0
    U[:, 0] += 10.0
    bcell = torch.from_numpy(ch.bcell).to("cuda:0").long()
    fixed = torch.from_numpy(uT == 1).to("cuda:0")
    Ub = torch.where(fixed[:, None], torch.from_numpy(uV).to("cuda:0"), U[bcell])
    own_d = torch.from_numpy(ch.own).to("cuda:0").long()
    nei_d = torch.from_numpy(ch.nei).to("cuda:0").long()
    nF = nI + nB
    Sf = ch.Sf.view(nF, 3)
    wv = ch.w.view(-1, 1)
    phi = ((wv * U[own_d] + (1.0 - wv) * U[nei_d]) * Sf[:nI]).sum(1).contiguous()
    phiB = (Ub * Sf[nI:]).sum(1).contiguous()
    fields = dict(U=[U[:, k].contiguous() for k in range(3)], Ub=[Ub[:, k].contiguous() for k in range(3)],
                  p=torch.zeros(nC, dtype=torch.float64, device="cuda:0"),
                  pb=torch.zeros(nB, dtype=torch.float64, device="cuda:0"), phi=phi, phiB=phiB)
    del U, Ub, own_d, nei_d, bcell
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    return fields


def channel_kepsilon(ch, torch, mfv, nu=1e-5, tol=1e-5, relTol=0.1, checkEvery=4, relax=0.7, log=None):
    """The channel's k-epsilon set-up: wall functions on the four walls (their near-wall distance from the
    library's host nearWallDist), inlet k = 0.375 and epsilon = 14.855, zeroGradient elsewhere; initial k and
    epsilon uniform at the inlet values, nut = Cmu k^2/epsilon (cells and boundary faces)."""
    t0 = time.time()
    nB, nI, nC = ch.nB, ch.nI, ch.nC
    isWall = np.zeros(nB, np.int32)
    ktype = np.full(nB, mfv.BC_ZERO_GRADIENT, np.int32)
    kV, eV = np.zeros(nB), np.zeros(nB)
    for name, start, size in ch.patches:
        s = slice(start - nI, start - nI + size)
        if name in ("lowerWall", "upperWall", "back", "front"):
            isWall[s] = 1
        elif name == "inlet":
            ktype[s] = mfv.BC_FIXED_VALUE
            kV[s], eV[s] = 0.375, 14.855
    nF = nI + nB
    y = mfv.near_wall_dist(nI, ch.points, np.arange(0, 4 * nF + 1, 4, dtype=np.int32), ch.faces.reshape(-1),
                           ch.bcell, ch.C.cpu().numpy(), isWall)
    if log:
        log("near-wall distance on %d wall faces (%.1fs host)" % (int(isWall.sum()), time.time() - t0))
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")
    k0 = torch.full((nC,), 0.375, dtype=torch.float64, device="cuda:0")
    e0 = torch.full((nC,), 14.855, dtype=torch.float64, device="cuda:0")
    bc = torch.from_numpy(ch.bcell).to("cuda:0").long()
    fixed = dev(ktype == mfv.BC_FIXED_VALUE)
    kb0 = torch.where(fixed, dev(kV), k0[bc])
    eb0 = torch.where(fixed, dev(eV), e0[bc])
    fields0 = dict(k=k0, kb=kb0.contiguous(), eps=e0, epsb=eb0.contiguous(), nut=(0.09 * k0 * k0 / e0).contiguous(),
                   nutb=(0.09 * kb0 * kb0 / eb0).contiguous())
    controls = mfv.kepsilon_controls(solver=dict(tolerance=tol, relTol=relTol, maxIter=1000,
                                                 preconditioner=mfv.PRECOND_DIAGONAL, checkEvery=checkEvery),
                                     nu=nu, relaxEps=relax, relaxK=relax)
    return dict(isWall=isWall, y=y, type=ktype, kV=kV, eV=eV, fields0=fields0, controls=controls)


def pressure_corrector(ch, ctx, mfv, amg, prec, tolerance=1e-6, nonOrth=1, checkEvery=4, nPre=2, nPost=1,
                       scale=True, graph=0, gamg=None):
    """One SIMPLE pressure corrector on the channel's inputs, from p = 0: interpolate(rAU), the laplacian on
    nonOrthDeltaCoeffs with its boundary coefficients, Galerkin once; then nonOrth + 1 times grad(p), the
    correction flux and its divergence, -div(phiHbyA)*V, fold and the solve (AMG-PCG in `prec`, or with
    `gamg` OpenFOAM's PCG + GAMG preconditioner on that hierarchy); finally phi = phiHbyA - (flux + ffc).
    Returns the solves' mfv.Perf."""
    P = mfv.ptr
    A = mfv.ldu(ch.diagC, ch.upper)
    ch.p.zero_()
    ctx.call("mfv_interpolate", ch.mesh.h, P(ch.w), P(ch.negRAU), P(ch.gf), P(ch.gb))
    ctx.call("mfv_laplacian_coeffs", ch.mesh.h, P(ch.ndc), P(ch.gf), P(ch.magSfI), P(ch.diag), P(ch.upper), None)
    ctx.call("mfv_laplacian_bcoeffs", ch.mesh.h, P(ch.types), P(ch.bcValue), P(ch.gb), P(ch.bMagSf),
             P(ch.bDelta), P(ch.iC), P(ch.bC))
    perfs = []
    for corr in range(nonOrth + 1):
        ch.source.zero_()
        ctx.call("mfv_grad_gauss", ch.mesh.h, P(ch.types), P(ch.bcValue), P(ch.Sf), P(ch.w), P(ch.V), P(ch.p),
                 P(ch.gradP))
        ctx.call("mfv_laplacian_corr", ch.mesh.h, P(ch.gradP), P(ch.w), P(ch.corr), P(ch.gf), P(ch.magSfI),
                 P(ch.ffc), P(ch.source))
        ctx.call("mfv_div_source", ch.mesh.h, P(ch.phiHbyA), P(ch.phiB), P(ch.V), -1.0, P(ch.source))
        ctx.call("mfv_fold", ch.mesh.h, P(ch.diag), P(ch.source), P(ch.iC), P(ch.bC), P(ch.diagC), P(ch.b))
        if gamg is not None:
            perfs.append(gamg.pcg_gamg_solve(ch.diagC, ch.upper, ch.b, ch.p, tolerance, smoother="GaussSeidel"))
            continue
        if corr == 0:
            amg.galerkin(ch.diagC, ch.upper, prec)
        perfs.append(amg.solve(ch.diagC, ch.upper, ch.b, ch.p, tolerance=tolerance, precision=prec,
                               flexible=int(prec != mfv.PREC_FP64 or scale), checkEvery=checkEvery,
                               scaleCorrection=int(scale), nPreSweeps=nPre, nPostSweeps=nPost, captureVcycle=graph))
    ctx.call("mfv_flux_correct", ch.mesh.h, ctypes.byref(A), P(ch.p), P(ch.ffc), P(ch.phiHbyA), -1.0, P(ch.phi))
    return perfs
