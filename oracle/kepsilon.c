/*
 * kepsilon.c -- TEST INFRASTRUCTURE ONLY: the FP64 CPU restatement of the reference's standard k-epsilon
 * model as its host SIMPLE driver calls it (kepsilon::correct,
 * src/TurbulenceModels/turbulenceModels/RAS/kEpsilon/k_epsilon.cuh:221-337), with the near-wall distance
 * it takes (nearWallDist, src/finiteVolume/fvMesh/near_wall_dist.cuh:31-177) and the nutkWallFunction
 * (src/TurbulenceModels/turbulenceModels/derivedFvPatchFields/wallFunctions/nut_wall_function.cu:8-37,
 * nut_wall_function.cuh:31-35). Incompressible, upwind convection (no `bounded`, no limitedLinear), uncoupled
 * fixedValue/zeroGradient/empty patches, the wall functions on the faces of `wall` patches. The parity tests
 * and bench.py's CPU-baseline leg check the HIP library against it; the product never links or loads it.
 *
 * Every loop keeps the reference's traversal order, because that order fixes the floating-point bits.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static void* kzalloc(size_t n, size_t sz)
{
    void* p = calloc(n ? n : 1, sz);
    if (!p) abort();
    return p;
}
#define KD(n) ((double*)kzalloc((size_t)(n), sizeof(double)))
#define KI(n) ((int*)kzalloc((size_t)(n), sizeof(int)))

void orc_kepsilon_default_coeffs(orc_kepsilon_coeffs* co)
{
    /* kepsilon_coeffs.cuh:11-22: OpenFOAM v2412's kEpsilon and wall-function defaults */
    co->Cmu = 0.09;
    co->C1 = 1.44;
    co->C2 = 1.92;
    co->C3 = 0.0;
    co->sigmaK = 1.0;
    co->sigmaEps = 1.3;
    co->kappa = 0.41;
    co->E = 9.8;
}

/* ---- near-wall distance (near_wall_dist.cuh) ---- */

static void sub3(const double* a, const double* b, double* o)
{
    o[0] = a[0] - b[0];
    o[1] = a[1] - b[1];
    o[2] = a[2] - b[2];
}
static double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* closest point on triangle (a, b, c) to p: near_wall_dist.cuh:33-56 (Ericson 5.1.5) */
static void closest_on_triangle(const double* p, const double* a, const double* b, const double* c, double* out)
{
    double ab[3], ac[3], ap[3];
    sub3(b, a, ab);
    sub3(c, a, ac);
    sub3(p, a, ap);
    const double d1 = dot3(ab, ap), d2 = dot3(ac, ap);
    if (d1 <= 0.0 && d2 <= 0.0)
    {
        memcpy(out, a, 3 * sizeof(double));
        return;
    }
    double bp[3];
    sub3(p, b, bp);
    const double d3 = dot3(ab, bp), d4 = dot3(ac, bp);
    if (d3 >= 0.0 && d4 <= d3)
    {
        memcpy(out, b, 3 * sizeof(double));
        return;
    }
    const double vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0)
    {
        const double v = d1 / (d1 - d3);
        for (int i = 0; i < 3; ++i) out[i] = a[i] + v * ab[i];
        return;
    }
    double cp[3];
    sub3(p, c, cp);
    const double d5 = dot3(ab, cp), d6 = dot3(ac, cp);
    if (d6 >= 0.0 && d5 <= d6)
    {
        memcpy(out, c, 3 * sizeof(double));
        return;
    }
    const double vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0)
    {
        const double w = d2 / (d2 - d6);
        for (int i = 0; i < 3; ++i) out[i] = a[i] + w * ac[i];
        return;
    }
    const double va = d3 * d6 - d5 * d4;
    if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0)
    {
        const double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        for (int i = 0; i < 3; ++i) out[i] = b[i] + w * (c[i] - b[i]);
        return;
    }
    const double denom = 1.0 / (va + vb + vc);
    const double v = vb * denom, w = vc * denom;
    for (int i = 0; i < 3; ++i) out[i] = a[i] + v * ab[i] + w * ac[i];
}

/* distance from p to a face polygon through the fan about its area-weighted centre:
 * near_wall_dist.cuh:63-98 */
static double point_to_face(const double* p, const double* pts, const int* fv, int v0, int v1)
{
    const int n = v1 - v0;
    double cp[3] = {0.0, 0.0, 0.0};
    for (int j = v0; j < v1; ++j)
        for (int i = 0; i < 3; ++i) cp[i] += pts[3 * (size_t)fv[j] + i];
    for (int i = 0; i < 3; ++i) cp[i] = (1.0 / n) * cp[i];
    double sumA = 0.0, sumAc[3] = {0.0, 0.0, 0.0};
    for (int j = 0; j < n; ++j)
    {
        const double* a = pts + 3 * (size_t)fv[v0 + j];
        const double* b = pts + 3 * (size_t)fv[v0 + (j + 1) % n];
        double u[3], v[3];
        sub3(a, cp, u);
        sub3(b, cp, v);
        const double cr[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
        const double ta = sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
        sumA += ta;
        for (int i = 0; i < 3; ++i) sumAc[i] += ta * (a[i] + b[i] + cp[i]);
    }
    double ctr[3];
    if (sumA > 1.0e-30)
        for (int i = 0; i < 3; ++i) ctr[i] = sumAc[i] / (3.0 * sumA);
    else
        memcpy(ctr, cp, sizeof(ctr));
    double best = 1.0e15;
    for (int j = 0; j < n; ++j)
    {
        const double* a = pts + 3 * (size_t)fv[v0 + j];
        const double* b = pts + 3 * (size_t)fv[v0 + (j + 1) % n];
        double q[3], d[3];
        closest_on_triangle(p, ctr, a, b, q);
        sub3(p, q, d);
        const double dist = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        if (dist < best) best = dist;
    }
    return best;
}

/* nearWallDist (near_wall_dist.cuh:120-177): for each face of a wall patch, the smallest distance from its
 * cell's centre to that face or to any wall face sharing a vertex with it, over all wall patches. The wall
 * faces are given as boundary-face indices (ascending: the patches in order, each patch's faces in order);
 * y is written per boundary face (wall faces only). Vertex-neighbour faces are visited per vertex of the
 * face, in the order the combined wall patch lists them; the minimum does not depend on that order. */
void orc_near_wall_dist(int nFaces, int nBFaces, const double* points, const int* faceOffsets,
                        const int* faceVerts, const int* bFaceCell, const double* C, const int* isWall, double* y)
{
    /* the combined wall patch and, per point, the wall faces on it */
    int nW = 0;
    for (int i = 0; i < nBFaces; ++i) nW += isWall[i] != 0;
    int* wf = KI(nW);
    {
        int k = 0;
        for (int i = 0; i < nBFaces; ++i)
            if (isWall[i]) wf[k++] = i;
    }
    int maxPt = -1;
    for (int w = 0; w < nW; ++w)
    {
        const int f = nFaces + wf[w];
        for (int j = faceOffsets[f]; j < faceOffsets[f + 1]; ++j)
            if (faceVerts[j] > maxPt) maxPt = faceVerts[j];
    }
    const int nP = maxPt + 1;
    int* cnt = KI(nP + 1);
    for (int w = 0; w < nW; ++w)
    {
        const int f = nFaces + wf[w];
        for (int j = faceOffsets[f]; j < faceOffsets[f + 1]; ++j) cnt[faceVerts[j] + 1]++;
    }
    for (int p = 0; p < nP; ++p) cnt[p + 1] += cnt[p];
    int* fill = KI(nP);
    int* pf = KI(cnt[nP]);
    for (int w = 0; w < nW; ++w)
    {
        const int f = nFaces + wf[w];
        for (int j = faceOffsets[f]; j < faceOffsets[f + 1]; ++j)
        {
            const int p = faceVerts[j];
            pf[cnt[p] + fill[p]++] = w;
        }
    }
    for (int i = 0; i < nBFaces; ++i) y[i] = 0.0;
    for (int w = 0; w < nW; ++w)
    {
        const int i = wf[w];
        const double* cc = C + 3 * (size_t)bFaceCell[i];
        const int f = nFaces + i;
        double best = point_to_face(cc, points, faceVerts, faceOffsets[f], faceOffsets[f + 1]);
        for (int j = faceOffsets[f]; j < faceOffsets[f + 1]; ++j)
        {
            const int p = faceVerts[j];
            for (int q = cnt[p]; q < cnt[p + 1]; ++q)
            {
                const int nw = pf[q];
                if (nw == w) continue;
                const int g = nFaces + wf[nw];
                const double d = point_to_face(cc, points, faceVerts, faceOffsets[g], faceOffsets[g + 1]);
                if (d < best) best = d;
            }
        }
        y[i] = best;
    }
    free(wf);
    free(cnt);
    free(fill);
    free(pf);
}

/* yPlusLam: the fixed point of yPlusLam = log(E yPlusLam)/kappa, ten iterations from 11
 * (nut_wall_function.cu:8-15) */
double orc_y_plus_lam(double kappa, double E)
{
    double ypl = 11.0;
    for (int i = 0; i < 10; ++i) ypl = log(fmax(E * ypl, 1.0)) / kappa;
    return ypl;
}

/* nutkWallFunction on one face (nut_wall_function.cuh:31-35, nut_wall_function.cu:19-37) */
double orc_nutk_wall(double Cmu, double kappa, double E, double y, double kCell, double nu)
{
    const double Cmu25 = pow(Cmu, 0.25);
    const double yplLam = orc_y_plus_lam(kappa, E);
    const double yPlus = Cmu25 * y * sqrt(kCell) / nu;
    return (yPlus > yplLam) ? (nu * yPlus * kappa / log(fmax(E * yPlus, 1.0 + 1e-4)) - nu) : 0.0;
}

/* GbyNu = gradU && devTwoSymm(gradU) per cell (k_epsilon.cuh:62-76); gradU component-major,
 * gradU[(i*3 + j)*n + c] = dU_j/dx_i */
void orc_gbynu(int n, const double* gradU, double* out)
{
    const size_t N = (size_t)n;
    for (int c = 0; c < n; ++c)
    {
        double t[9];
        for (int q = 0; q < 9; ++q) t[q] = gradU[q * N + c];
        const double t23 = (2.0 / 3.0) * (t[0] + t[4] + t[8]);
        double d[9];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) d[i * 3 + j] = t[i * 3 + j] + t[j * 3 + i];
        d[0] -= t23;
        d[4] -= t23;
        d[8] -= t23;
        double s = t[0] * d[0];
        for (int q = 1; q < 9; ++q) s += t[q] * d[q];
        out[c] = s;
    }
}

/* fvc::div(phi) (fvc.cu:706-733): internal faces in order, then non-empty boundary faces, over V */
static void div_phi(const orc_kepsilon_case* cs, const double* phi, const double* phiB, double* d)
{
    memset(d, 0, sizeof(double) * (size_t)cs->nCells);
    for (int f = 0; f < cs->nFaces; ++f)
    {
        d[cs->owner[f]] += phi[f];
        d[cs->neighbour[f]] -= phi[f];
    }
    for (int i = 0; i < cs->nBFaces; ++i)
        if (cs->kType[i] != ORC_BC_EMPTY) d[cs->bFaceCell[i]] += phiB[i];
    for (int c = 0; c < cs->nCells; ++c) d[c] /= cs->V[c];
}

/* div(phi, psi) - laplacian(D, psi) (k_epsilon.cuh:150-157, 173-181): upwind convection with its boundary
 * coefficients, minus the orthogonal laplacian of the face diffusivity (cells interpolated inside, the
 * patch's own on the boundary), through addEqual(M, L, -1) */
static void transport_matrix(const orc_kepsilon_case* cs, const int* type, const double* value, const double* phi,
                             const double* phiB, const double* Dc, const double* Db, double* diag, double* upper,
                             double* lower, double* iC, double* bC)
{
    const int nC = cs->nCells, nF = cs->nFaces, nB = cs->nBFaces;
    const int* own = cs->owner;
    const int* nei = cs->neighbour;
    orc_div_upwind(nC, nF, own, nei, phi, diag, upper, lower);
    orc_div_bcoeffs(nB, type, value, phiB, iC, bC);
    double* Df = KD(nF);
    for (int f = 0; f < nF; ++f) Df[f] = cs->weights[f] * Dc[own[f]] + (1.0 - cs->weights[f]) * Dc[nei[f]];
    double* lD = KD(nC);
    double* lU = KD(nF);
    orc_laplacian(nC, nF, own, nei, cs->deltaCoeffs, Df, cs->magSf, lD, lU, NULL);
    double* liC = KD(nB);
    double* lbC = KD(nB);
    orc_laplacian_bcoeffs(nB, type, value, Db, cs->magSf + nF, cs->bDeltaCoeffs, liC, lbC);
    for (int c = 0; c < nC; ++c) diag[c] += -1.0 * lD[c];
    for (int f = 0; f < nF; ++f)
    {
        upper[f] += -1.0 * lU[f];
        lower[f] += -1.0 * lU[f];
    }
    for (int i = 0; i < nB; ++i)
    {
        iC[i] += -1.0 * liC[i];
        bC[i] += -1.0 * lbC[i];
    }
    free(Df); free(lD); free(lU); free(liC); free(lbC);
}

void orc_kepsilon_eqns(const orc_kepsilon_case* cs, const double* k, const double* kb, const double* eps,
                       const double* epsb, const double* nut, const double* nutb, const double* G,
                       const double* GbyNu, const double* phi, const double* phiB, double* kD, double* kU,
                       double* kL, double* kS, double* kIC, double* kBC, double* eD, double* eU, double* eL,
                       double* eS, double* eIC, double* eBC)
{
    const int nC = cs->nCells, nB = cs->nBFaces;
    const orc_kepsilon_coeffs* co = &cs->co;
    const double nu = cs->nu;
    (void)kb;
    (void)epsb;
    double* Dc = KD(nC);
    double* Db = KD(nB);
    /* kEqn (k_epsilon.cuh:137-160): DkEff = nut/sigmak + nu, the patch's own nut on the boundary (or the
     * cell diffusivity interpolated to the face, the owner's, when nutb is NULL) */
    for (int c = 0; c < nC; ++c) Dc[c] = nut[c] / co->sigmaK + nu;
    for (int i = 0; i < nB; ++i) Db[i] = nutb ? nutb[i] / co->sigmaK + nu : Dc[cs->bFaceCell[i]];
    transport_matrix(cs, cs->kType, cs->kValue, phi, phiB, Dc, Db, kD, kU, kL, kIC, kBC);
    for (int c = 0; c < nC; ++c)
    {
        kD[c] += (eps[c] / k[c]) * cs->V[c];
        kS[c] = 0.0;
        kS[c] += G[c] * cs->V[c];
    }
    /* epsilonEqn (k_epsilon.cuh:163-186): DepsilonEff = nut/sigmaEps + nu */
    for (int c = 0; c < nC; ++c) Dc[c] = nut[c] / co->sigmaEps + nu;
    for (int i = 0; i < nB; ++i) Db[i] = nutb ? nutb[i] / co->sigmaEps + nu : Dc[cs->bFaceCell[i]];
    transport_matrix(cs, cs->epsType, cs->epsValue, phi, phiB, Dc, Db, eD, eU, eL, eIC, eBC);
    for (int c = 0; c < nC; ++c)
    {
        eD[c] += (co->C2 * eps[c] / k[c]) * cs->V[c];
        eS[c] = 0.0;
        eS[c] += (co->C1 * co->Cmu * k[c] * GbyNu[c]) * cs->V[c];
    }
    free(Dc);
    free(Db);
}

/* fvMatrix::setValues at the near-wall cells (fv_matrix_ops.cuh:110-148): a constrained cell's internal
 * faces move their coupling to the other cell's source and are zeroed, its boundary coefficients are zeroed,
 * then psi and source = value*diag are written. Cells ascending, each cell's faces in face order. */
static void set_values(const orc_kepsilon_case* cs, int nWall, const int* cells, const double* values,
                       double* diag, double* upper, double* lower, double* source, double* iC, double* bC,
                       double* psi)
{
    const int nC = cs->nCells, nF = cs->nFaces, nB = cs->nBFaces;
    const int* own = cs->owner;
    const int* nei = cs->neighbour;
    /* cell -> its internal faces in face order, and its boundary faces */
    int* st = KI(nC + 1);
    for (int f = 0; f < nF; ++f)
    {
        st[own[f] + 1]++;
        st[nei[f] + 1]++;
    }
    for (int c = 0; c < nC; ++c) st[c + 1] += st[c];
    int* fill = KI(nC);
    int* adj = KI(2 * (size_t)nF);
    for (int f = 0; f < nF; ++f)
    {
        adj[st[own[f]] + fill[own[f]]++] = f;
        adj[st[nei[f]] + fill[nei[f]]++] = f;
    }
    int* bs = KI(nC + 1);
    for (int i = 0; i < nB; ++i) bs[cs->bFaceCell[i] + 1]++;
    for (int c = 0; c < nC; ++c) bs[c + 1] += bs[c];
    int* bl = KI(nB);
    memset(fill, 0, sizeof(int) * (size_t)nC);
    for (int i = 0; i < nB; ++i) bl[bs[cs->bFaceCell[i]] + fill[cs->bFaceCell[i]]++] = i;
    for (int w = 0; w < nWall; ++w)
    {
        const int c = cells[w];
        const double v = values[w];
        for (int q = st[c]; q < st[c + 1]; ++q)
        {
            const int f = adj[q];
            if (own[f] == c)
                source[nei[f]] -= lower[f] * v;
            else
                source[own[f]] -= upper[f] * v;
            upper[f] = 0.0;
            lower[f] = 0.0;
        }
        for (int q = bs[c]; q < bs[c + 1]; ++q)
        {
            iC[bl[q]] = 0.0;
            bC[bl[q]] = 0.0;
        }
    }
    for (int w = 0; w < nWall; ++w)
    {
        const int c = cells[w];
        psi[c] = values[w];
        source[c] = values[w] * diag[c];
    }
    free(st); free(fill); free(adj); free(bs); free(bl);
}

/* patch values after evaluate(): fixedValue keeps its value, zeroGradient takes the owner cell's, empty 0 */
static void evaluate_scalar(int nB, const int* bfc, const int* type, const double* value, const double* cell,
                            double* out)
{
    for (int i = 0; i < nB; ++i)
    {
        if (type[i] == ORC_BC_FIXED_VALUE)
            out[i] = value[i];
        else if (type[i] == ORC_BC_ZERO_GRADIENT)
            out[i] = cell[bfc[i]];
        else
            out[i] = 0.0;
    }
}

/* relax (fv_matrix_ops.cuh:161-209) of a scalar equation: cmptMagMax(iC) = |iC|, cmptMin(iC) = iC */
static void relax_scalar(const orc_kepsilon_case* cs, double alpha, double* diag, const double* upper,
                         const double* lower, const double* iC, const double* psi, double* source)
{
    const int nB = cs->nBFaces;
    double* mag = KD(nB);
    for (int i = 0; i < nB; ++i) mag[i] = fabs(iC[i]);
    const double* ps[1] = {psi};
    double* sr[1] = {source};
    orc_relax_matrix(cs->nCells, cs->nFaces, cs->owner, cs->neighbour, upper, lower, nB, cs->bFaceCell, mag, iC,
                     NULL, alpha, diag, 1, ps, sr, NULL);
    free(mag);
}

int orc_kepsilon_correct(const orc_kepsilon_case* cs, const double* const* U, const double* const* Ub,
                         const double* phi, const double* phiB, double* k, double* kb, double* eps, double* epsb,
                         double* nut, double* nutb, orc_kepsilon_result* res)
{
    const int nC = cs->nCells, nF = cs->nFaces, nB = cs->nBFaces;
    const int* bfc = cs->bFaceCell;
    const orc_kepsilon_coeffs* co = &cs->co;
    const double nu = cs->nu;
    memset(res, 0, sizeof(*res));

    /* nut's boundary as it enters: the diffusivities' patch values */
    double* nutB0 = KD(nB);
    memcpy(nutB0, nutb, sizeof(double) * (size_t)nB);
    const double kappa = co->kappa;
    const double Cmu25 = pow(co->Cmu, 0.25), Cmu75 = pow(co->Cmu, 0.75);

    /* production from the current nut: grad(U) with every non-empty face at U's patch value */
    int* gType = KI(nB);
    for (int i = 0; i < nB; ++i) gType[i] = cs->kType[i] == ORC_BC_EMPTY ? ORC_BC_EMPTY : ORC_BC_FIXED_VALUE;
    double* gradU = KD(9 * (size_t)nC);
    double* g3 = KD(3 * (size_t)nC);
    for (int j = 0; j < 3; ++j)
    {
        orc_grad_gauss(nC, nF, cs->owner, cs->neighbour, nB, bfc, gType, Ub[j], cs->Sf, cs->weights, cs->V, U[j],
                       g3);
        for (int c = 0; c < nC; ++c)
            for (int i = 0; i < 3; ++i) gradU[(size_t)(i * 3 + j) * nC + c] = g3[3 * (size_t)c + i];
    }
    double* gByNu = KD(nC);
    orc_gbynu(nC, gradU, gByNu);
    double* G = KD(nC);
    for (int c = 0; c < nC; ++c) G[c] = nut[c] * gByNu[c];
    double* divU = KD(nC);
    div_phi(cs, phi, phiB, divU);

    /* epsilonWallFunction: near-wall epsilon and production, corner weights 1/nWallFaces */
    int* nw = KI(nC);
    for (int i = 0; i < nB; ++i)
        if (cs->isWall[i]) ++nw[bfc[i]];
    double* eps0 = KD(nC);
    double* G0 = KD(nC);
    for (int i = 0; i < nB; ++i)
    {
        if (!cs->isWall[i]) continue;
        const int c = bfc[i];
        const double y = cs->yWall[i];
        const double nutw = orc_nutk_wall(co->Cmu, co->kappa, co->E, y, k[c], nu);
        const double w = 1.0 / nw[c], kc = k[c];
        const double dc = cs->bDeltaCoeffs[i];
        const double gx = (Ub[0][i] - U[0][c]) * dc, gy = (Ub[1][i] - U[1][c]) * dc, gz = (Ub[2][i] - U[2][c]) * dc;
        const double magGradUw = sqrt(gx * gx + gy * gy + gz * gz);
        eps0[c] += w * Cmu75 * pow(kc, 1.5) / (kappa * y);
        G0[c] += w * (nutw + nu) * magGradUw * Cmu25 * sqrt(kc) / (kappa * y);
    }
    int nWall = 0;
    for (int c = 0; c < nC; ++c) nWall += nw[c] > 0;
    int* wallCells = KI(nWall);
    double* epsVals = KD(nWall);
    {
        int q = 0;
        for (int c = 0; c < nC; ++c)
            if (nw[c] > 0)
            {
                G[c] = G0[c];
                eps[c] = eps0[c];
                wallCells[q] = c;
                epsVals[q] = eps0[c];
                ++q;
            }
    }
    res->nWallCells = nWall;

    double* D = KD(nC);
    double* Up = KD(nF);
    double* Lo = KD(nF);
    double* S = KD(nC);
    double* iC = KD(nB);
    double* bC = KD(nB);
    double* Dc = KD(nC);
    double* Db = KD(nB);

    /* ---- the epsilon equation ---- */
    for (int c = 0; c < nC; ++c) Dc[c] = nut[c] / co->sigmaEps + nu;
    for (int i = 0; i < nB; ++i) Db[i] = nutB0[i] / co->sigmaEps + nu;
    transport_matrix(cs, cs->epsType, cs->epsValue, phi, phiB, Dc, Db, D, Up, Lo, iC, bC);
    for (int c = 0; c < nC; ++c)
    {
        D[c] += (co->C2 * eps[c] / k[c]) * cs->V[c];
        S[c] = 0.0;
        S[c] += (co->C1 * co->Cmu * k[c] * gByNu[c]) * cs->V[c];
    }
    for (int c = 0; c < nC; ++c)
    {
        const double sp = ((2.0 / 3.0) * co->C1 - co->C3) * divU[c];
        D[c] += cs->V[c] * fmax(sp, 0.0);
        S[c] -= cs->V[c] * fmin(sp, 0.0) * eps[c];
    }
    relax_scalar(cs, cs->relaxEps, D, Up, Lo, iC, eps, S);
    set_values(cs, nWall, wallCells, epsVals, D, Up, Lo, S, iC, bC, eps);
    {
        const orc_perf pf = orc_pbicgstab(nC, nF, cs->owner, cs->neighbour, D, Up, Lo, S, nB, bfc, iC, bC, eps,
                                          cs->tol, cs->relTol, cs->maxIter, cs->minIter);
        res->epsInitial = pf.initialResidual;
        res->epsFinal = pf.finalResidual;
        res->epsIters = pf.nIterations;
    }
    for (int c = 0; c < nC; ++c) eps[c] = fmax(eps[c], 1e-15);
    evaluate_scalar(nB, bfc, cs->epsType, cs->epsValue, eps, epsb);

    /* ---- the k equation (the wall-overridden G, the new epsilon) ---- */
    for (int c = 0; c < nC; ++c) Dc[c] = nut[c] / co->sigmaK + nu;
    for (int i = 0; i < nB; ++i) Db[i] = nutB0[i] / co->sigmaK + nu;
    transport_matrix(cs, cs->kType, cs->kValue, phi, phiB, Dc, Db, D, Up, Lo, iC, bC);
    for (int c = 0; c < nC; ++c)
    {
        D[c] += (eps[c] / k[c]) * cs->V[c];
        S[c] = 0.0;
        S[c] += G[c] * cs->V[c];
    }
    for (int c = 0; c < nC; ++c)
    {
        const double sp = (2.0 / 3.0) * divU[c];
        D[c] += cs->V[c] * fmax(sp, 0.0);
        S[c] -= cs->V[c] * fmin(sp, 0.0) * k[c];
    }
    relax_scalar(cs, cs->relaxK, D, Up, Lo, iC, k, S);
    {
        const orc_perf pf = orc_pbicgstab(nC, nF, cs->owner, cs->neighbour, D, Up, Lo, S, nB, bfc, iC, bC, k,
                                          cs->tol, cs->relTol, cs->maxIter, cs->minIter);
        res->kInitial = pf.initialResidual;
        res->kFinal = pf.finalResidual;
        res->kIters = pf.nIterations;
    }
    for (int c = 0; c < nC; ++c) k[c] = fmax(k[c], 1e-15);
    evaluate_scalar(nB, bfc, cs->kType, cs->kValue, k, kb);

    /* correctNut: Cmu k^2/epsilon in the cells, the nutkWallFunction of the new k on wall faces */
    for (int c = 0; c < nC; ++c) nut[c] = co->Cmu * k[c] * k[c] / eps[c];
    for (int i = 0; i < nB; ++i)
        if (cs->isWall[i]) nutb[i] = orc_nutk_wall(co->Cmu, co->kappa, co->E, cs->yWall[i], k[bfc[i]], nu);

    free(nutB0); free(gType); free(gradU); free(g3); free(gByNu); free(G); free(divU); free(nw); free(eps0);
    free(G0); free(wallCells); free(epsVals); free(D); free(Up); free(Lo); free(S); free(iC); free(bC); free(Dc);
    free(Db);
    return 0;
}
