/*
 * oracle.c -- TEST INFRASTRUCTURE ONLY: the FP64 CPU restatement of the reference path that the
 * parity tests, smoke() and bench.py's CPU-baseline leg check the HIP library against. The product
 * never links or loads it. Each function cites the reference file:line it restates; the loops keep
 * the reference's traversal order because that order fixes the floating-point bits.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ROOTVSMALL 1.0e-150
#define VSMALL 1.0e-300
#define PERF_SMALL 1e-20
#define PERF_VSMALL 1e-300
#define PERF_GREAT 1e+20

typedef struct { double x, y, z; } vec3;

static vec3 v_add(vec3 a, vec3 b) { vec3 r = {a.x + b.x, a.y + b.y, a.z + b.z}; return r; }
static vec3 v_sub(vec3 a, vec3 b) { vec3 r = {a.x - b.x, a.y - b.y, a.z - b.z}; return r; }
static vec3 v_scale(double s, vec3 a) { vec3 r = {s * a.x, s * a.y, s * a.z}; return r; }
static vec3 v_div(vec3 a, double s) { vec3 r = {a.x / s, a.y / s, a.z / s}; return r; }
static double v_dot(vec3 a, vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static double v_mag(vec3 a) { return sqrt(v_dot(a, a)); }
static vec3 v_cross(vec3 a, vec3 b)
{
    vec3 r = {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
    return r;
}
static vec3 v_get(const double* p, long i) { vec3 r = {p[3 * i], p[3 * i + 1], p[3 * i + 2]}; return r; }
static void v_put(double* p, long i, vec3 v) { p[3 * i] = v.x; p[3 * i + 1] = v.y; p[3 * i + 2] = v.z; }

static void* xcalloc(size_t n, size_t sz)
{
    void* p = calloc(n ? n : 1, sz);
    if (!p) abort();
    return p;
}

/* ------------------------------------------------------------------------------------------------
 * Geometry. fv_geometry.cu:13-77 (makeFaceCentresAndAreas), :79-127 (makeCellCentresAndVols),
 * :129-158 (makeInterpolation).
 * ---------------------------------------------------------------------------------------------- */
int orc_geometry(int nPoints, const double* points, int nFaces, const int* faceOffsets, const int* faceVerts,
                 int nCells, int nInternal, const int* owner, const int* neighbour,
                 double* Cf, double* Sf, double* magSf, double* C, double* V,
                 double* weights, double* deltaCoeffs, double* nonOrthDeltaCoeffs, double* corrVec)
{
    (void)nPoints;
    for (int f = 0; f < nFaces; ++f)
    {
        const int k = faceOffsets[f + 1] - faceOffsets[f];
        const int* fv = faceVerts + faceOffsets[f];
        if (k == 3)
        {
            vec3 a = v_get(points, fv[0]), b = v_get(points, fv[1]), c = v_get(points, fv[2]);
            v_put(Cf, f, v_scale(1.0 / 3.0, v_add(v_add(a, b), c)));
            v_put(Sf, f, v_scale(0.5, v_cross(v_sub(b, a), v_sub(c, a))));
        }
        else
        {
            vec3 fCentre = v_get(points, fv[0]);
            for (int pi = 1; pi < k; ++pi) fCentre = v_add(fCentre, v_get(points, fv[pi]));
            fCentre = v_div(fCentre, (double)k);
            vec3 sumN = {0, 0, 0}, sumAc = {0, 0, 0};
            double sumA = 0.0;
            for (int pi = 0; pi < k; ++pi)
            {
                vec3 thisP = v_get(points, fv[pi]);
                vec3 nextP = v_get(points, fv[pi == k - 1 ? 0 : pi + 1]);
                vec3 c = v_add(v_add(thisP, nextP), fCentre);
                vec3 n = v_cross(v_sub(nextP, thisP), v_sub(fCentre, thisP));
                double a = v_mag(n);
                sumN = v_add(sumN, n);
                sumA += a;
                sumAc = v_add(sumAc, v_scale(a, c));
            }
            if (sumA < ROOTVSMALL)
            {
                vec3 z = {0, 0, 0};
                v_put(Cf, f, fCentre);
                v_put(Sf, f, z);
            }
            else
            {
                v_put(Cf, f, v_div(v_scale(1.0 / 3.0, sumAc), sumA));
                v_put(Sf, f, v_scale(0.5, sumN));
            }
        }
    }
    for (int f = 0; f < nFaces; ++f) magSf[f] = v_mag(v_get(Sf, f));

    /* cell centres and volumes: owned faces pass, then neighbour faces pass */
    double* cEst = (double*)xcalloc(3 * (size_t)nCells, sizeof(double));
    int* nCellFaces = (int*)xcalloc((size_t)nCells, sizeof(int));
    for (int f = 0; f < nFaces; ++f)
    {
        v_put(cEst, owner[f], v_add(v_get(cEst, owner[f]), v_get(Cf, f)));
        ++nCellFaces[owner[f]];
    }
    for (int f = 0; f < nInternal; ++f)
    {
        v_put(cEst, neighbour[f], v_add(v_get(cEst, neighbour[f]), v_get(Cf, f)));
        ++nCellFaces[neighbour[f]];
    }
    for (int c = 0; c < nCells; ++c) v_put(cEst, c, v_div(v_get(cEst, c), (double)nCellFaces[c]));
    memset(C, 0, 3 * (size_t)nCells * sizeof(double));
    memset(V, 0, (size_t)nCells * sizeof(double));
    for (int f = 0; f < nFaces; ++f)
    {
        const int o = owner[f];
        vec3 cf = v_get(Cf, f), ce = v_get(cEst, o);
        double pyr3Vol = v_dot(v_get(Sf, f), v_sub(cf, ce));
        vec3 pc = v_add(v_scale(0.75, cf), v_scale(0.25, ce));
        v_put(C, o, v_add(v_get(C, o), v_scale(pyr3Vol, pc)));
        V[o] += pyr3Vol;
    }
    for (int f = 0; f < nInternal; ++f)
    {
        const int n = neighbour[f];
        vec3 cf = v_get(Cf, f), ce = v_get(cEst, n);
        double pyr3Vol = v_dot(v_get(Sf, f), v_sub(ce, cf));
        vec3 pc = v_add(v_scale(0.75, cf), v_scale(0.25, ce));
        v_put(C, n, v_add(v_get(C, n), v_scale(pyr3Vol, pc)));
        V[n] += pyr3Vol;
    }
    for (int c = 0; c < nCells; ++c)
        v_put(C, c, fabs(V[c]) > VSMALL ? v_div(v_get(C, c), V[c]) : v_get(cEst, c));
    for (int c = 0; c < nCells; ++c) V[c] *= (1.0 / 3.0);
    free(cEst);
    free(nCellFaces);

    /* interpolation quantities on internal faces */
    for (int f = 0; f < nInternal; ++f)
    {
        const int o = owner[f], n = neighbour[f];
        vec3 sf = v_get(Sf, f), cf = v_get(Cf, f), co = v_get(C, o), cn = v_get(C, n);
        double SfdOwn = fabs(v_dot(sf, v_sub(cf, co)));
        double SfdNei = fabs(v_dot(sf, v_sub(cn, cf)));
        weights[f] = fabs(SfdOwn + SfdNei) > ROOTVSMALL ? SfdNei / (SfdOwn + SfdNei) : 0.5;
        vec3 delta = v_sub(cn, co);
        deltaCoeffs[f] = 1.0 / v_mag(delta);
        vec3 unitArea = v_div(sf, magSf[f]);
        nonOrthDeltaCoeffs[f] = 1.0 / fmax(v_dot(unitArea, delta), 0.05 * v_mag(delta));
        v_put(corrVec, f, v_sub(unitArea, v_scale(nonOrthDeltaCoeffs[f], delta)));
    }
    return 0;
}

/* fv_patch.cu:144-166: deltaCoeffs = 1/|(Cf - C).nHat| */
void orc_patch_delta(int nBFaces, const int* bFaceCell, const double* bCf, const double* bSf,
                     const double* bMagSf, const double* C, double* bDeltaCoeffs)
{
    for (int i = 0; i < nBFaces; ++i)
    {
        vec3 nHat = v_div(v_get(bSf, i), bMagSf[i]);
        double normalDelta = v_dot(v_sub(v_get(bCf, i), v_get(C, bFaceCell[i])), nHat);
        bDeltaCoeffs[i] = 1.0 / fabs(normalDelta);
    }
}

/* ------------------------------------------------------------------------------------------------
 * lduAddressing. device_ldu.cuh:109-146: ownerStart is the prefix count of faces per owner;
 * losort is a STABLE counting sort of the faces by neighbour, losortStart its starts.
 * ---------------------------------------------------------------------------------------------- */
int orc_ldu_addressing(int nCells, int nFaces, const int* owner, const int* neighbour,
                       int* ownerStart, int* losortStart, int* losort)
{
    memset(ownerStart, 0, ((size_t)nCells + 1) * sizeof(int));
    memset(losortStart, 0, ((size_t)nCells + 1) * sizeof(int));
    for (int f = 0; f < nFaces; ++f)
    {
        if (owner[f] >= neighbour[f] || (f > 0 && owner[f] < owner[f - 1])) return -1;
        ++ownerStart[owner[f] + 1];
        ++losortStart[neighbour[f] + 1];
    }
    for (int c = 0; c < nCells; ++c)
    {
        ownerStart[c + 1] += ownerStart[c];
        losortStart[c + 1] += losortStart[c];
    }
    int* at = (int*)xcalloc((size_t)nCells, sizeof(int));
    memcpy(at, losortStart, (size_t)nCells * sizeof(int));
    for (int f = 0; f < nFaces; ++f) losort[at[neighbour[f]]++] = f;
    free(at);
    return 0;
}

/* ------------------------------------------------------------------------------------------------
 * fvm::laplacian(gamma, psi), interior: fvm.cuh:113-151. diag accumulates -coeff in face order.
 * ---------------------------------------------------------------------------------------------- */
void orc_laplacian(int nCells, int nFaces, const int* owner, const int* neighbour,
                   const double* dc, const double* gammaf, const double* magSf,
                   double* diag, double* upper, double* lower)
{
    for (int c = 0; c < nCells; ++c) diag[c] = 0.0;
    for (int f = 0; f < nFaces; ++f)
    {
        const double coeff = dc[f] * gammaf[f] * magSf[f];
        upper[f] = coeff;
        if (lower) lower[f] = coeff;
        diag[owner[f]] -= coeff;
        diag[neighbour[f]] -= coeff;
    }
}

/* fvm.cuh:43-103 with the fixedValue/zeroGradient patch coefficients:
 * fixedValue gradientInternalCoeffs = -deltaCoeffs, gradientBoundaryCoeffs = deltaCoeffs*value;
 * zeroGradient both zero. iC = gamma*magSf*gIC, bC = -gamma*magSf*gBC. */
void orc_laplacian_bcoeffs(int nBFaces, const int* bcType, const double* bcValue, const double* gammab,
                           const double* bMagSf, const double* bDeltaCoeffs, double* iC, double* bC)
{
    for (int i = 0; i < nBFaces; ++i)
    {
        const double pGamma = gammab[i] * bMagSf[i];
        double gIC = 0.0, gBC = 0.0;
        if (bcType[i] == ORC_BC_FIXED_VALUE)
        {
            gIC = -bDeltaCoeffs[i];
            gBC = bDeltaCoeffs[i] * bcValue[i];
        }
        if (bcType[i] == ORC_BC_EMPTY)
        {
            iC[i] = 0.0;
            bC[i] = 0.0;
            continue;
        }
        iC[i] = pGamma * gIC;
        bC[i] = (-pGamma) * gBC;
    }
}

/* fvm::div(phi, psi) upwind: fvm.cuh:476-554 (weights = pos0(phi)) */
void orc_div_upwind(int nCells, int nFaces, const int* owner, const int* neighbour, const double* phi,
                    double* diag, double* upper, double* lower)
{
    for (int c = 0; c < nCells; ++c) diag[c] = 0.0;
    for (int f = 0; f < nFaces; ++f)
    {
        const double w = phi[f] >= 0.0 ? 1.0 : 0.0;
        lower[f] = -w * phi[f];
        upper[f] = lower[f] + phi[f];
        diag[owner[f]] -= lower[f];
        diag[neighbour[f]] -= upper[f];
    }
}

/* fvm::div(phi, psi) Gauss linear (deviceDivCentralCoeffs, src/cuda/device_fvm.cu:43-61,628-643): the upwind
 * coefficients with the geometric weights w in place of pos0(phi) */
void orc_div_linear(int nCells, int nFaces, const int* owner, const int* neighbour, const double* phi,
                    const double* weights, double* diag, double* upper, double* lower)
{
    for (int c = 0; c < nCells; ++c) diag[c] = 0.0;
    for (int f = 0; f < nFaces; ++f)
    {
        lower[f] = -weights[f] * phi[f];
        upper[f] = lower[f] + phi[f];
        diag[owner[f]] -= lower[f];
        diag[neighbour[f]] -= upper[f];
    }
}

/* fvm.cuh:545-552: iC = phi_b*valueInternalCoeffs, bC = -phi_b*valueBoundaryCoeffs; fixedValue
 * (0, value), zeroGradient (1, 0). */
void orc_div_bcoeffs(int nBFaces, const int* bcType, const double* bcValue, const double* phib,
                     double* iC, double* bC)
{
    for (int i = 0; i < nBFaces; ++i)
    {
        double vIC = 1.0, vBC = 0.0;
        if (bcType[i] == ORC_BC_FIXED_VALUE) { vIC = 0.0; vBC = bcValue[i]; }
        if (bcType[i] == ORC_BC_EMPTY) { iC[i] = 0.0; bC[i] = 0.0; continue; }
        iC[i] = phib[i] * vIC;
        bC[i] = (-phib[i]) * vBC;
    }
}

/* fvMatrix::solveSegregated fold, gamg_solver_cpp.cu:769-797 (no coupled patches here) */
void orc_fold(int nCells, const double* diag, const double* source, int nBFaces, const int* bFaceCell,
              const int* bcType, const double* iC, const double* bC, double* diagC, double* b)
{
    memcpy(diagC, diag, (size_t)nCells * sizeof(double));
    memcpy(b, source, (size_t)nCells * sizeof(double));
    for (int i = 0; i < nBFaces; ++i)
    {
        if (bcType && bcType[i] == ORC_BC_EMPTY) continue;
        diagC[bFaceCell[i]] += iC[i];
        b[bFaceCell[i]] += bC[i];
    }
}

/* fvMatrix::setReference as the device path applies it, device_simple.cuh:104-106 */
void orc_set_reference(double* diag, double* b, int refCell, double refValue)
{
    if (refCell < 0) return;
    b[refCell] += diag[refCell] * refValue;
    diag[refCell] += diag[refCell];
}

/* ------------------------------------------------------------------------------------------------
 * LDU products: ldu_spmv.cu:5-24 (Amul), gamg_solver_cpp.cu:88-112 (residual), :115-159 (normFactor)
 * ---------------------------------------------------------------------------------------------- */
void orc_amul(int nCells, int nFaces, const int* owner, const int* neighbour, const double* diag,
              const double* upper, const double* lower, const double* psi, double* Apsi)
{
    if (!lower) lower = upper;
    for (int c = 0; c < nCells; ++c) Apsi[c] = diag[c] * psi[c];
    for (int f = 0; f < nFaces; ++f)
    {
        Apsi[neighbour[f]] += lower[f] * psi[owner[f]];
        Apsi[owner[f]] += upper[f] * psi[neighbour[f]];
    }
}

void orc_residual(int nCells, int nFaces, const int* owner, const int* neighbour, const double* diag,
                  const double* upper, const double* lower, const double* psi, const double* source, double* r)
{
    if (!lower) lower = upper;
    for (int c = 0; c < nCells; ++c) r[c] = source[c] - diag[c] * psi[c];
    for (int f = 0; f < nFaces; ++f)
    {
        r[neighbour[f]] -= lower[f] * psi[owner[f]];
        r[owner[f]] -= upper[f] * psi[neighbour[f]];
    }
}

double orc_sum_mag(int n, const double* x)
{
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += fabs(x[i]);
    return s;
}

double orc_sum_prod(int n, const double* a, const double* b)
{
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

/* lduMatrix::solver::normFactor: sum(|A.psi - sumA*xRef| + |b - sumA*xRef|) + small,
 * pcg.cu:68-97 and gamg_solver_cpp.cu:115-159 (symmetric row sums in face order) */
static double norm_factor_with(int nCells, int nFaces, const int* owner, const int* neighbour, const double* diag,
                               const double* upper, const double* lower, const double* psi, const double* source,
                               const double* Apsi)
{
    if (!lower) lower = upper;
    double* sumA = (double*)xcalloc((size_t)nCells, sizeof(double));
    for (int c = 0; c < nCells; ++c) sumA[c] = diag[c];
    for (int f = 0; f < nFaces; ++f)
    {
        sumA[neighbour[f]] += lower[f];
        sumA[owner[f]] += upper[f];
    }
    double xRef = 0.0;
    for (int c = 0; c < nCells; ++c) xRef += psi[c];
    xRef /= (double)nCells;
    double nf = 0.0;
    for (int c = 0; c < nCells; ++c)
    {
        const double t = sumA[c] * xRef;
        nf += fabs(Apsi[c] - t) + fabs(source[c] - t);
    }
    free(sumA);
    return nf + PERF_SMALL;
}

double orc_norm_factor(int nCells, int nFaces, const int* owner, const int* neighbour, const double* diag,
                       const double* upper, const double* lower, const double* psi, const double* source)
{
    double* Apsi = (double*)xcalloc((size_t)nCells, sizeof(double));
    orc_amul(nCells, nFaces, owner, neighbour, diag, upper, lower, psi, Apsi);
    double nf = norm_factor_with(nCells, nFaces, owner, neighbour, diag, upper, lower, psi, source, Apsi);
    free(Apsi);
    return nf;
}

void orc_face_flux(int nFaces, const int* owner, const int* neighbour, const double* upper,
                   const double* lower, const double* psi, double* flux)
{
    if (!lower) lower = upper;
    for (int f = 0; f < nFaces; ++f) flux[f] = upper[f] * psi[neighbour[f]] - lower[f] * psi[owner[f]];
}

/* ------------------------------------------------------------------------------------------------
 * Non-orthogonal correction. Gauss linear gradient of a cell field with its patch values
 * (fvc.cu:67-121, gaussGrad of a GeometricField): internal faces in face order, face value
 * w*(P - N) + N, +Sf*pf to the owner and -Sf*pf to the neighbour; then boundary faces in order (an
 * empty patch contributes nothing; the patch value is the bc value for fixedValue, the cell value for
 * zeroGradient); then divided by V. Sf is per face, internal then boundary, xyz triples.
 * ---------------------------------------------------------------------------------------------- */
void orc_grad_gauss(int nCells, int nFaces, const int* owner, const int* neighbour, int nBFaces,
                    const int* bFaceCell, const int* bcType, const double* bcValue, const double* Sf,
                    const double* weights, const double* V, const double* psi, double* grad)
{
    memset(grad, 0, sizeof(double) * 3 * (size_t)nCells);
    for (int f = 0; f < nFaces; ++f)
    {
        const int o = owner[f], n = neighbour[f];
        const double pf = weights[f] * (psi[o] - psi[n]) + psi[n];
        const vec3 s = v_scale(pf, v_get(Sf, f));
        v_put(grad, o, v_add(v_get(grad, o), s));
        v_put(grad, n, v_sub(v_get(grad, n), s));
    }
    for (int i = 0; i < nBFaces; ++i)
    {
        if (bcType[i] == ORC_BC_EMPTY) continue;
        const int c = bFaceCell[i];
        const double pv = bcType[i] == ORC_BC_FIXED_VALUE ? bcValue[i] : psi[c];
        v_put(grad, c, v_add(v_get(grad, c), v_scale(pv, v_get(Sf, nFaces + i))));
    }
    for (int c = 0; c < nCells; ++c) v_put(grad, c, v_div(v_get(grad, c), V[c]));
}

/* The correction as a face flux (fvm.cuh:170-222, unlimited):
 *   ffc_f = (gamma_f*magSf_f) * (cv_f . (w_f*grad[own] + (1 - w_f)*grad[nei]))
 * and, when source is given, its divergence taken off the source in face order (fvm.cuh:276-299 and
 * pEqn_cpp.cu:223-229): src[own] += ffc, src[nei] -= ffc, source -= src. */
void orc_laplacian_corr(int nCells, int nFaces, const int* owner, const int* neighbour, const double* grad,
                        const double* weights, const double* corrVec, const double* gammaf, const double* magSf,
                        double* ffc, double* source)
{
    for (int f = 0; f < nFaces; ++f)
    {
        const double w = weights[f];
        const vec3 gf = v_add(v_scale(w, v_get(grad, owner[f])), v_scale(1.0 - w, v_get(grad, neighbour[f])));
        const vec3 cv = v_get(corrVec, f);
        const double corr = cv.x * gf.x + cv.y * gf.y + cv.z * gf.z;
        ffc[f] = (gammaf[f] * magSf[f]) * corr;
    }
    if (!source) return;
    double* src = (double*)xcalloc((size_t)nCells, sizeof(double));
    for (int f = 0; f < nFaces; ++f)
    {
        src[owner[f]] += ffc[f];
        src[neighbour[f]] -= ffc[f];
    }
    for (int c = 0; c < nCells; ++c) source[c] -= src[c];
    free(src);
}

/* `limited k corrected` (fvm.cuh:170-211, OpenFOAM limitedSnGrad): per face the correction corr is capped
 * against the orthogonal snGrad orth = nonOrthDc*(vf[nei] - vf[own]),
 *   lim = min(k*|orth| / ((1 - k)*|corr| + 1e-15), 1),  ffc = (gamma*magSf)*(lim*corr);
 * k outside (0, 1) is the unlimited correction. The source as in orc_laplacian_corr. */
void orc_laplacian_corr_limited(int nCells, int nFaces, const int* owner, const int* neighbour, const double* grad,
                                const double* weights, const double* corrVec, const double* gammaf,
                                const double* magSf, const double* nonOrthDc, const double* vf, double limitCoeff,
                                double* ffc, double* source)
{
    const int limited = limitCoeff > 0.0 && limitCoeff < 1.0;
    for (int f = 0; f < nFaces; ++f)
    {
        const double w = weights[f];
        const vec3 gf = v_add(v_scale(w, v_get(grad, owner[f])), v_scale(1.0 - w, v_get(grad, neighbour[f])));
        const vec3 cv = v_get(corrVec, f);
        double corr = cv.x * gf.x + cv.y * gf.y + cv.z * gf.z;
        if (limited)
        {
            const double orth = nonOrthDc[f] * (vf[neighbour[f]] - vf[owner[f]]);
            const double lim = fmin(limitCoeff * fabs(orth) / ((1.0 - limitCoeff) * fabs(corr) + 1e-15), 1.0);
            corr = lim * corr;
        }
        ffc[f] = (gammaf[f] * magSf[f]) * corr;
    }
    if (!source) return;
    double* src = (double*)xcalloc((size_t)nCells, sizeof(double));
    for (int f = 0; f < nFaces; ++f)
    {
        src[owner[f]] += ffc[f];
        src[neighbour[f]] -= ffc[f];
    }
    for (int c = 0; c < nCells; ++c) source[c] -= src[c];
    free(src);
}

/* phi = phiHbyA - (upper*psi[nei] - lower*psi[own] + ffc) (fv_matrix_ops.cuh:33-50 with the stored
 * faceFluxCorrection, then pEqn_cpp.cu:255-258); ffc may be NULL */
void orc_flux_correct(int nFaces, const int* owner, const int* neighbour, const double* upper, const double* lower,
                      const double* psi, const double* ffc, const double* phiHbyA, double* phi)
{
    if (!lower) lower = upper;
    for (int f = 0; f < nFaces; ++f)
    {
        double fl = upper[f] * psi[neighbour[f]] - lower[f] * psi[owner[f]];
        if (ffc) fl += ffc[f];
        phi[f] = phiHbyA[f] - fl;
    }
}

/* ------------------------------------------------------------------------------------------------
 * PCG with DIC, pcg.cu:6-149 (OpenFOAM PCG + DICPreconditioner + normFactor). Folded system in.
 * ---------------------------------------------------------------------------------------------- */
orc_perf orc_pcg_dic(int nCells, int nFaces, const int* owner, const int* neighbour, const double* diag,
                     const double* upper, const double* b, double* psi, double tol, double relTol,
                     int maxIter, int minIter)
{
    const int nC = nCells;
    double* pA = (double*)xcalloc((size_t)nC, sizeof(double));
    double* wA = (double*)xcalloc((size_t)nC, sizeof(double));
    double* rA = (double*)xcalloc((size_t)nC, sizeof(double));
    double* rD = (double*)xcalloc((size_t)nC, sizeof(double));
    orc_amul(nC, nFaces, owner, neighbour, diag, upper, upper, psi, wA);
    for (int c = 0; c < nC; ++c) rA[c] = b[c] - wA[c];
    const double nf = norm_factor_with(nC, nFaces, owner, neighbour, diag, upper, upper, psi, b, wA);
    orc_perf perf;
    perf.initialResidual = orc_sum_mag(nC, rA) / nf;
    perf.finalResidual = perf.initialResidual;
    perf.nIterations = 0;
    memcpy(rD, diag, (size_t)nC * sizeof(double));
    for (int f = 0; f < nFaces; ++f) rD[neighbour[f]] -= upper[f] * upper[f] / rD[owner[f]];
    for (int c = 0; c < nC; ++c) rD[c] = 1.0 / rD[c];
#define ORC_CONV(fr) ((fr) < tol || (relTol > 0.0 && (fr) < relTol * perf.initialResidual))
    double wArA = 1e300, wArAold;
    int nIter = 0;
    if (minIter > 0 || !ORC_CONV(perf.finalResidual))
    {
        do
        {
            wArAold = wArA;
            for (int c = 0; c < nC; ++c) wA[c] = rD[c] * rA[c];
            for (int f = 0; f < nFaces; ++f) wA[neighbour[f]] -= rD[neighbour[f]] * upper[f] * wA[owner[f]];
            for (int f = nFaces - 1; f >= 0; --f) wA[owner[f]] -= rD[owner[f]] * upper[f] * wA[neighbour[f]];
            wArA = orc_sum_prod(nC, wA, rA);
            if (nIter == 0)
                memcpy(pA, wA, (size_t)nC * sizeof(double));
            else
            {
                const double beta = wArA / wArAold;
                for (int c = 0; c < nC; ++c) pA[c] = wA[c] + beta * pA[c];
            }
            orc_amul(nC, nFaces, owner, neighbour, diag, upper, upper, pA, wA);
            const double wApA = orc_sum_prod(nC, wA, pA);
            const double alpha = wArA / wApA;
            for (int c = 0; c < nC; ++c)
            {
                psi[c] += alpha * pA[c];
                rA[c] -= alpha * wA[c];
            }
            perf.finalResidual = orc_sum_mag(nC, rA) / nf;
            ++nIter;
        } while ((nIter < maxIter && !ORC_CONV(perf.finalResidual)) || nIter < minIter);
    }
#undef ORC_CONV
    perf.nIterations = nIter;
    free(pA); free(wA); free(rA); free(rD);
    return perf;
}

/* ------------------------------------------------------------------------------------------------
 * Agglomeration hierarchies.
 * ---------------------------------------------------------------------------------------------- */
typedef struct
{
    int nFine, nFineFaces, nCoarse, nCoarseFaces;
    int* map;          /* [nFine] */
    int* cOwn;         /* [nCoarseFaces] */
    int* cNei;         /* [nCoarseFaces] */
    int* faceRestrict; /* [nFineFaces] */
    signed char* faceFlip;
} orc_level;

struct orc_aggl
{
    int nLevels;
    int cap;
    orc_level* lv;
    int fineCells, fineFaces;
    int* fineOwn; /* copies of the mesh's internal-face addressing */
    int* fineNei;
};

static orc_aggl* aggl_new(int nCells, int nFaces, const int* owner, const int* neighbour)
{
    orc_aggl* a = (orc_aggl*)xcalloc(1, sizeof(orc_aggl));
    a->fineCells = nCells;
    a->fineFaces = nFaces;
    a->fineOwn = (int*)xcalloc((size_t)nFaces, sizeof(int));
    a->fineNei = (int*)xcalloc((size_t)nFaces, sizeof(int));
    memcpy(a->fineOwn, owner, (size_t)nFaces * sizeof(int));
    memcpy(a->fineNei, neighbour, (size_t)nFaces * sizeof(int));
    return a;
}

static void aggl_push(orc_aggl* a, orc_level L)
{
    if (a->nLevels == a->cap)
    {
        a->cap = a->cap ? 2 * a->cap : 8;
        a->lv = (orc_level*)realloc(a->lv, (size_t)a->cap * sizeof(orc_level));
        if (!a->lv) abort();
    }
    a->lv[a->nLevels++] = L;
}

void orc_faceareapair_weights(int nFaces, const double* Sf, const double* magSf, double* w)
{
    for (int f = 0; f < nFaces; ++f)
    {
        const double rootMagSf = sqrt(magSf[f]);
        const double wx = Sf[3 * f] / rootMagSf;
        const double wy = (Sf[3 * f + 1] / rootMagSf) * 1.01;
        const double wz = (Sf[3 * f + 2] / rootMagSf) * 1.02;
        w[f] = sqrt(wx * wx + wy * wy + wz * wz);
    }
}

/* pair_gamg_agglomeration_cpp.cu:158-284: one pairing pass. cellFaces lists the faces a cell is the
 * UPPER (neighbour) cell of first, then those it owns, each in face order. Flips *forward. */
static int* of_pair_pass(int nFineCells, int nFaces, const int* lowerAddr, const int* upperAddr,
                         const double* fw, int* forward, int* nCoarseOut)
{
    int* off = (int*)xcalloc((size_t)nFineCells + 1, sizeof(int));
    int* nNbrs = (int*)xcalloc((size_t)nFineCells, sizeof(int));
    int* cellFaces = (int*)xcalloc(2 * (size_t)nFaces, sizeof(int));
    for (int f = 0; f < nFaces; ++f) ++nNbrs[upperAddr[f]];
    for (int f = 0; f < nFaces; ++f) ++nNbrs[lowerAddr[f]];
    for (int c = 0; c < nFineCells; ++c) off[c + 1] = off[c] + nNbrs[c];
    memset(nNbrs, 0, (size_t)nFineCells * sizeof(int));
    for (int f = 0; f < nFaces; ++f) { int c = upperAddr[f]; cellFaces[off[c] + nNbrs[c]++] = f; }
    for (int f = 0; f < nFaces; ++f) { int c = lowerAddr[f]; cellFaces[off[c] + nNbrs[c]++] = f; }

    int* map = (int*)xcalloc((size_t)nFineCells, sizeof(int));
    for (int c = 0; c < nFineCells; ++c) map[c] = -1;
    int nCoarse = 0;
    const double great = 1.0e+15;
    const int fwd = *forward;
    for (int cellfi = 0; cellfi < nFineCells; ++cellfi)
    {
        const int celli = fwd ? cellfi : nFineCells - cellfi - 1;
        if (map[celli] >= 0) continue;
        int matchFaceNo = -1;
        double maxFaceWeight = -great;
        for (int o = off[celli]; o < off[celli + 1]; ++o)
        {
            const int f = cellFaces[o];
            if (map[upperAddr[f]] < 0 && map[lowerAddr[f]] < 0 && fw[f] > maxFaceWeight)
            {
                matchFaceNo = f;
                maxFaceWeight = fw[f];
            }
        }
        if (matchFaceNo >= 0)
        {
            map[upperAddr[matchFaceNo]] = nCoarse;
            map[lowerAddr[matchFaceNo]] = nCoarse;
            ++nCoarse;
            continue;
        }
        int clusterMatchFaceNo = -1;
        double clusterMaxFaceCoeff = -great;
        for (int o = off[celli]; o < off[celli + 1]; ++o)
        {
            const int f = cellFaces[o];
            if (fw[f] > clusterMaxFaceCoeff)
            {
                clusterMatchFaceNo = f;
                clusterMaxFaceCoeff = fw[f];
            }
        }
        if (clusterMatchFaceNo >= 0)
        {
            const int mf = clusterMatchFaceNo;
            const int mu = map[upperAddr[mf]], ml = map[lowerAddr[mf]];
            map[celli] = mu > ml ? mu : ml;
        }
    }
    for (int cellfi = 0; cellfi < nFineCells; ++cellfi)
    {
        const int celli = fwd ? cellfi : nFineCells - cellfi - 1;
        if (map[celli] < 0) map[celli] = nCoarse++;
    }
    if (!fwd)
    {
        --nCoarse;
        for (int c = 0; c < nFineCells; ++c) map[c] = nCoarse - map[c];
        ++nCoarse;
    }
    *forward = !fwd;
    free(off); free(nNbrs); free(cellFaces);
    *nCoarseOut = nCoarse;
    return map;
}

/* pair_gamg_agglomeration_cpp.cu:28-154: coarse faces owner-major in first-appearance order,
 * faceRestrict (>=0 coarse face, <0 -(coarse cell+1)), flip where the fine UPPER cell is in the
 * coarse owner. */
static void of_coarse_addressing(orc_level* L, const int* lowerAddr, const int* upperAddr)
{
    const int nFineFaces = L->nFineFaces, nCoarse = L->nCoarse;
    const int* rm = L->map;
    int* cCellnFaces = (int*)xcalloc((size_t)nCoarse, sizeof(int));
    int* cStart = (int*)xcalloc((size_t)nCoarse + 1, sizeof(int));
    int* initNei = (int*)xcalloc((size_t)nFineFaces, sizeof(int));
    int* faceRestrict = (int*)xcalloc((size_t)nFineFaces, sizeof(int));
    /* Each coarse owner's list of coarse faces in the order the walk creates them. A linked list per
     * owner keeps first-appearance order without the reference's doubling buffer; the lookup scans
     * the owner's faces in the same order, so the result is the same. */
    int* headOf = (int*)xcalloc((size_t)nCoarse, sizeof(int));
    int* tailOf = (int*)xcalloc((size_t)nCoarse, sizeof(int));
    int* nextOf = (int*)xcalloc((size_t)nFineFaces, sizeof(int));
    for (int c = 0; c < nCoarse; ++c) { headOf[c] = -1; tailOf[c] = -1; }
    int nCoarseFaces = 0;
    for (int f = 0; f < nFineFaces; ++f)
    {
        const int rU = rm[upperAddr[f]], rL = rm[lowerAddr[f]];
        if (rU == rL)
        {
            faceRestrict[f] = -(rU + 1);
            continue;
        }
        const int cOwn = rU < rL ? rU : rL;
        const int cNei = rU < rL ? rL : rU;
        int found = -1;
        for (int cf = headOf[cOwn]; cf >= 0; cf = nextOf[cf])
            if (initNei[cf] == cNei) { found = cf; break; }
        if (found >= 0) { faceRestrict[f] = found; continue; }
        initNei[nCoarseFaces] = cNei;
        nextOf[nCoarseFaces] = -1;
        if (tailOf[cOwn] >= 0) nextOf[tailOf[cOwn]] = nCoarseFaces; else headOf[cOwn] = nCoarseFaces;
        tailOf[cOwn] = nCoarseFaces;
        ++cCellnFaces[cOwn];
        faceRestrict[f] = nCoarseFaces;
        ++nCoarseFaces;
    }
    L->nCoarseFaces = nCoarseFaces;
    L->cOwn = (int*)xcalloc((size_t)nCoarseFaces, sizeof(int));
    L->cNei = (int*)xcalloc((size_t)nCoarseFaces, sizeof(int));
    int* coarseFaceMap = (int*)xcalloc((size_t)nCoarseFaces, sizeof(int));
    int k = 0;
    for (int c = 0; c < nCoarse; ++c)
        for (int cf = headOf[c]; cf >= 0; cf = nextOf[cf])
        {
            L->cOwn[k] = c;
            L->cNei[k] = initNei[cf];
            coarseFaceMap[cf] = k;
            ++k;
        }
    for (int f = 0; f < nFineFaces; ++f)
        if (faceRestrict[f] >= 0) faceRestrict[f] = coarseFaceMap[faceRestrict[f]];
    L->faceFlip = (signed char*)xcalloc((size_t)nFineFaces, 1);
    for (int f = 0; f < nFineFaces; ++f)
    {
        const int cFace = faceRestrict[f];
        if (cFace < 0) continue;
        const int rU = rm[upperAddr[f]];
        if (L->cOwn[cFace] == rU) L->faceFlip[f] = 1;
    }
    L->faceRestrict = faceRestrict;
    free(cCellnFaces); free(cStart); free(initNei); free(headOf); free(tailOf); free(nextOf);
    free(coarseFaceMap);
}

orc_aggl* orc_of_agglomeration(int nCells, int nFaces, const int* owner, const int* neighbour,
                               const double* weights, int nCellsInCoarsestLevel, int* forward)
{
    orc_aggl* a = aggl_new(nCells, nFaces, owner, neighbour);
    /* GAMGAgglomeration.C:285 clamp, pair_gamg_agglomeration_cpp.cu:298 */
    int half = nCells / 2;
    int nc = nCellsInCoarsestLevel < half ? nCellsInCoarsestLevel : half;
    nCellsInCoarsestLevel = nc > 1 ? nc : 1;
    const int maxLevels = 50;
    double* fw = (double*)xcalloc((size_t)nFaces, sizeof(double));
    memcpy(fw, weights, (size_t)nFaces * sizeof(double));
    const int* lo = owner;
    const int* up = neighbour;
    int nFine = nCells, nFineFaces = nFaces;
    while (a->nLevels < maxLevels - 1)
    {
        int nCoarse = -1;
        int* map = of_pair_pass(nFine, nFineFaces, lo, up, fw, forward, &nCoarse);
        /* continueAgglomerating: nCoarse >= nCellsInCoarsestLevel && nCoarse < nFine */
        if (!(nCoarse >= nCellsInCoarsestLevel && nCoarse < nFine))
        {
            free(map);
            break;
        }
        orc_level L;
        memset(&L, 0, sizeof(L));
        L.nFine = nFine;
        L.nFineFaces = nFineFaces;
        L.nCoarse = nCoarse;
        L.map = map;
        of_coarse_addressing(&L, lo, up);
        /* restrictFaceField: coarse weights summed in ascending fine face */
        double* cfw = (double*)xcalloc((size_t)L.nCoarseFaces, sizeof(double));
        for (int f = 0; f < nFineFaces; ++f)
            if (L.faceRestrict[f] >= 0) cfw[L.faceRestrict[f]] += fw[f];
        free(fw);
        fw = cfw;
        aggl_push(a, L);
        orc_level* P = &a->lv[a->nLevels - 1];
        lo = P->cOwn;
        up = P->cNei;
        nFine = P->nCoarse;
        nFineFaces = P->nCoarseFaces;
    }
    free(fw);
    return a;
}

/* device_amg.cu:533-685 (agglomerate: forward greedy over interleaved own/nei cellFaces, strict >
 * with -1e300 seed) and :444-522 (coarseFacesByBucket: sorted unique (lo,hi) pairs). */
static void brae_agglomerate_pass(int nC, int nFaces, const int* owner, const int* nei, const double* fw,
                                  orc_level* L, double** coarseWeights)
{
    int* nNbr = (int*)xcalloc((size_t)nC, sizeof(int));
    int* off = (int*)xcalloc((size_t)nC + 1, sizeof(int));
    int* cellFaces = (int*)xcalloc(2 * (size_t)nFaces, sizeof(int));
    for (int f = 0; f < nFaces; ++f) { nNbr[owner[f]]++; nNbr[nei[f]]++; }
    for (int c = 0; c < nC; ++c) off[c + 1] = off[c] + nNbr[c];
    memset(nNbr, 0, (size_t)nC * sizeof(int));
    for (int f = 0; f < nFaces; ++f)
    {
        cellFaces[off[owner[f]] + nNbr[owner[f]]++] = f;
        cellFaces[off[nei[f]] + nNbr[nei[f]]++] = f;
    }
    int* map = (int*)xcalloc((size_t)nC, sizeof(int));
    for (int c = 0; c < nC; ++c) map[c] = -1;
    int nCoarse = 0;
    for (int c = 0; c < nC; ++c)
    {
        if (map[c] >= 0) continue;
        int mf = -1;
        double mw = -1e300;
        for (int o = off[c]; o < off[c + 1]; ++o)
        {
            const int f = cellFaces[o];
            if (map[owner[f]] < 0 && map[nei[f]] < 0 && fw[f] > mw) { mf = f; mw = fw[f]; }
        }
        if (mf >= 0)
        {
            map[owner[mf]] = nCoarse;
            map[nei[mf]] = nCoarse;
            ++nCoarse;
        }
        else
        {
            int cm = -1;
            double cmw = -1e300;
            for (int o = off[c]; o < off[c + 1]; ++o)
            {
                const int f = cellFaces[o];
                if (fw[f] > cmw) { cm = f; cmw = fw[f]; }
            }
            if (cm >= 0)
            {
                const int m1 = map[owner[cm]], m2 = map[nei[cm]];
                map[c] = m1 > m2 ? m1 : m2;
            }
        }
    }
    for (int c = 0; c < nC; ++c)
        if (map[c] < 0) map[c] = nCoarse++;
    free(nNbr); free(off); free(cellFaces);

    /* coarse faces: bucket crossing faces by lower coarse cell, sort each bucket by higher cell */
    int* start = (int*)xcalloc((size_t)nCoarse + 1, sizeof(int));
    for (int f = 0; f < nFaces; ++f)
    {
        const int co = map[owner[f]], cn = map[nei[f]];
        if (co != cn) ++start[(co < cn ? co : cn) + 1];
    }
    for (int c = 0; c < nCoarse; ++c) start[c + 1] += start[c];
    const int nCross = start[nCoarse];
    int* eHi = (int*)xcalloc((size_t)nCross, sizeof(int));
    int* eF = (int*)xcalloc((size_t)nCross, sizeof(int));
    int* at = (int*)xcalloc((size_t)nCoarse + 1, sizeof(int));
    memcpy(at, start, (size_t)nCoarse * sizeof(int));
    for (int f = 0; f < nFaces; ++f)
    {
        const int co = map[owner[f]], cn = map[nei[f]];
        if (co != cn)
        {
            const int lo = co < cn ? co : cn, hi = co < cn ? cn : co;
            eHi[at[lo]] = hi;
            eF[at[lo]] = f;
            at[lo]++;
        }
    }
    int* faceRestrict = (int*)xcalloc((size_t)nFaces, sizeof(int));
    int* cOwn = (int*)xcalloc((size_t)nCross, sizeof(int));
    int* cNei = (int*)xcalloc((size_t)nCross, sizeof(int));
    int nCF = 0;
    for (int c = 0; c < nCoarse; ++c)
    {
        /* insertion sort of the bucket by (hi, f): std::sort of pairs (hi, f) */
        for (int i = start[c] + 1; i < start[c + 1]; ++i)
        {
            int h = eHi[i], ff = eF[i], j = i - 1;
            while (j >= start[c] && (eHi[j] > h || (eHi[j] == h && eF[j] > ff)))
            {
                eHi[j + 1] = eHi[j];
                eF[j + 1] = eF[j];
                --j;
            }
            eHi[j + 1] = h;
            eF[j + 1] = ff;
        }
        int last = -1;
        for (int i = start[c]; i < start[c + 1]; ++i)
        {
            if (eHi[i] != last)
            {
                last = eHi[i];
                cOwn[nCF] = c;
                cNei[nCF] = last;
                ++nCF;
            }
            faceRestrict[eF[i]] = nCF - 1;
        }
    }
    signed char* faceFlip = (signed char*)xcalloc((size_t)nFaces, 1);
    double* cfw = (double*)xcalloc((size_t)nCF, sizeof(double));
    for (int f = 0; f < nFaces; ++f)
    {
        const int co = map[owner[f]], cn = map[nei[f]];
        if (co == cn)
        {
            faceRestrict[f] = -1 - co;
            faceFlip[f] = 0;
        }
        else
        {
            faceFlip[f] = co > cn ? 1 : 0;
            cfw[faceRestrict[f]] += fw[f];
        }
    }
    free(start); free(eHi); free(eF); free(at);
    L->nFine = nC;
    L->nFineFaces = nFaces;
    L->nCoarse = nCoarse;
    L->nCoarseFaces = nCF;
    L->map = map;
    L->cOwn = cOwn;
    L->cNei = cNei;
    L->faceRestrict = faceRestrict;
    L->faceFlip = faceFlip;
    *coarseWeights = cfw;
}

orc_aggl* orc_brae_agglomeration(int nCells, int nFaces, const int* owner, const int* neighbour,
                                 const double* weights, int target)
{
    orc_aggl* a = aggl_new(nCells, nFaces, owner, neighbour);
    const int* own = owner;
    const int* nei = neighbour;
    double* fw = (double*)xcalloc((size_t)nFaces, sizeof(double));
    memcpy(fw, weights, (size_t)nFaces * sizeof(double));
    int n = nCells, nf = nFaces;
    while (n > target)
    {
        orc_level L;
        memset(&L, 0, sizeof(L));
        double* cfw = NULL;
        brae_agglomerate_pass(n, nf, own, nei, fw, &L, &cfw);
        if (L.nCoarse >= n || L.nCoarseFaces == 0)
        {
            free(L.map); free(L.cOwn); free(L.cNei); free(L.faceRestrict); free(L.faceFlip); free(cfw);
            break;
        }
        aggl_push(a, L);
        free(fw);
        fw = cfw;
        orc_level* P = &a->lv[a->nLevels - 1];
        own = P->cOwn;
        nei = P->cNei;
        n = P->nCoarse;
        nf = P->nCoarseFaces;
    }
    free(fw);
    return a;
}

int orc_aggl_nlevels(const orc_aggl* a) { return a->nLevels; }

void orc_aggl_sizes(const orc_aggl* a, int i, int* nFineCells, int* nFineFaces, int* nCoarseCells, int* nCoarseFaces)
{
    const orc_level* L = &a->lv[i];
    if (nFineCells) *nFineCells = L->nFine;
    if (nFineFaces) *nFineFaces = L->nFineFaces;
    if (nCoarseCells) *nCoarseCells = L->nCoarse;
    if (nCoarseFaces) *nCoarseFaces = L->nCoarseFaces;
}

void orc_aggl_level(const orc_aggl* a, int i, int* restrictMap, int* coarseOwner, int* coarseNeighbour,
                    int* faceRestrict, signed char* faceFlip)
{
    const orc_level* L = &a->lv[i];
    if (restrictMap) memcpy(restrictMap, L->map, (size_t)L->nFine * sizeof(int));
    if (coarseOwner) memcpy(coarseOwner, L->cOwn, (size_t)L->nCoarseFaces * sizeof(int));
    if (coarseNeighbour) memcpy(coarseNeighbour, L->cNei, (size_t)L->nCoarseFaces * sizeof(int));
    if (faceRestrict) memcpy(faceRestrict, L->faceRestrict, (size_t)L->nFineFaces * sizeof(int));
    if (faceFlip) memcpy(faceFlip, L->faceFlip, (size_t)L->nFineFaces);
}

void orc_aggl_free(orc_aggl* a)
{
    if (!a) return;
    for (int i = 0; i < a->nLevels; ++i)
    {
        free(a->lv[i].map); free(a->lv[i].cOwn); free(a->lv[i].cNei);
        free(a->lv[i].faceRestrict); free(a->lv[i].faceFlip);
    }
    free(a->lv);
    free(a->fineOwn);
    free(a->fineNei);
    free(a);
}

/* ------------------------------------------------------------------------------------------------
 * Galerkin coarse matrices.
 * OpenFOAM order, gamg_solver_cpp.cu:474-526: diag = restrict(fine diag) ascending fine cell, then
 * the fine faces in order: coarse face += upper, or coarse diag += 2*upper.
 * brae order, device_amg.cu:156-223: per coarse cell sum(fine diag ascending) then
 * sum(upper+lower of interior faces ascending); per coarse face sum(fine faces ascending).
 * ---------------------------------------------------------------------------------------------- */
static size_t level_cells_offset(const orc_aggl* a, int upto)
{
    size_t s = 0;
    for (int i = 0; i < upto; ++i) s += (size_t)a->lv[i].nCoarse;
    return s;
}
static size_t level_faces_offset(const orc_aggl* a, int upto)
{
    size_t s = 0;
    for (int i = 0; i < upto; ++i) s += (size_t)a->lv[i].nCoarseFaces;
    return s;
}

void orc_galerkin_of(const orc_aggl* a, const double* fineDiag, const double* fineUpper,
                     double* coarseDiag, double* coarseUpper)
{
    const double* fd = fineDiag;
    const double* fu = fineUpper;
    for (int i = 0; i < a->nLevels; ++i)
    {
        const orc_level* L = &a->lv[i];
        double* cd = coarseDiag + level_cells_offset(a, i);
        double* cu = coarseUpper + level_faces_offset(a, i);
        for (int c = 0; c < L->nCoarse; ++c) cd[c] = 0.0;
        for (int c = 0; c < L->nFine; ++c) cd[L->map[c]] += fd[c];
        for (int f = 0; f < L->nCoarseFaces; ++f) cu[f] = 0.0;
        for (int f = 0; f < L->nFineFaces; ++f)
        {
            const int cf = L->faceRestrict[f];
            if (cf >= 0) cu[cf] += fu[f];
            else cd[-1 - cf] += 2 * fu[f];
        }
        fd = cd;
        fu = cu;
    }
}

void orc_galerkin_brae(const orc_aggl* a, const double* fineDiag, const double* fineUpper,
                       double* coarseDiag, double* coarseUpper)
{
    const double* fd = fineDiag;
    const double* fu = fineUpper;
    for (int i = 0; i < a->nLevels; ++i)
    {
        const orc_level* L = &a->lv[i];
        double* cd = coarseDiag + level_cells_offset(a, i);
        double* cu = coarseUpper + level_faces_offset(a, i);
        /* two accumulators per coarse cell: fine diags first, then interior faces, both ascending */
        double* sd = (double*)xcalloc((size_t)L->nCoarse, sizeof(double));
        double* sf = (double*)xcalloc((size_t)L->nCoarse, sizeof(double));
        int* hasF = (int*)xcalloc((size_t)L->nCoarse, sizeof(int));
        for (int c = 0; c < L->nFine; ++c) sd[L->map[c]] += fd[c];
        for (int c = 0; c < L->nCoarse; ++c) sf[c] = sd[c];
        (void)hasF;
        for (int f = 0; f < L->nFineFaces; ++f)
        {
            const int cf = L->faceRestrict[f];
            if (cf < 0) sf[-1 - cf] += fu[f] + fu[f];
        }
        for (int c = 0; c < L->nCoarse; ++c) cd[c] = sf[c];
        for (int f = 0; f < L->nCoarseFaces; ++f) cu[f] = 0.0;
        for (int f = 0; f < L->nFineFaces; ++f)
            if (L->faceRestrict[f] >= 0) cu[L->faceRestrict[f]] += fu[f];
        free(sd); free(sf); free(hasF);
        fd = cd;
        fu = cu;
    }
}

/* ------------------------------------------------------------------------------------------------
 * OpenFOAM GAMG on a symmetric folded system. gamg_solver_cpp.cu:155-526 (smoothers, pcgDic,
 * scale, restrict/prolong), :627-735 (vCycle), :979-1023, :1027-1143.
 * ---------------------------------------------------------------------------------------------- */
typedef struct
{
    int nCells, nFaces;
    const int* lo; /* owner side */
    const int* up; /* neighbour side */
    double* diag;
    double* upper;
    int* ownStart;
    double* rD; /* DIC reciprocal diagonal (smoother) */
} of_mat;

static void of_amul(const of_mat* A, const double* psi, double* Apsi)
{
    for (int c = 0; c < A->nCells; ++c) Apsi[c] = A->diag[c] * psi[c];
    for (int f = 0; f < A->nFaces; ++f)
    {
        Apsi[A->up[f]] += A->upper[f] * psi[A->lo[f]];
        Apsi[A->lo[f]] += A->upper[f] * psi[A->up[f]];
    }
}

static void of_residual(const of_mat* A, double* rA, const double* psi, const double* source)
{
    for (int c = 0; c < A->nCells; ++c) rA[c] = source[c] - A->diag[c] * psi[c];
    for (int f = 0; f < A->nFaces; ++f)
    {
        rA[A->up[f]] -= A->upper[f] * psi[A->lo[f]];
        rA[A->lo[f]] -= A->upper[f] * psi[A->up[f]];
    }
}

static double of_normfactor(const of_mat* A, const double* psi, const double* source, const double* Apsi)
{
    double* tmp = (double*)xcalloc((size_t)A->nCells, sizeof(double));
    for (int c = 0; c < A->nCells; ++c) tmp[c] = A->diag[c];
    for (int f = 0; f < A->nFaces; ++f)
    {
        tmp[A->up[f]] += A->upper[f];
        tmp[A->lo[f]] += A->upper[f];
    }
    double average = 0;
    for (int c = 0; c < A->nCells; ++c) average += psi[c];
    average /= (double)A->nCells;
    double sum = 0;
    for (int c = 0; c < A->nCells; ++c)
    {
        const double t = tmp[c] * average;
        sum += fabs(Apsi[c] - t) + fabs(source[c] - t);
    }
    free(tmp);
    return sum + PERF_SMALL;
}

static double* of_dic_rd(const of_mat* A)
{
    double* rD = (double*)xcalloc((size_t)A->nCells, sizeof(double));
    memcpy(rD, A->diag, (size_t)A->nCells * sizeof(double));
    for (int f = 0; f < A->nFaces; ++f) rD[A->up[f]] -= A->upper[f] * A->upper[f] / rD[A->lo[f]];
    for (int c = 0; c < A->nCells; ++c) rD[c] = 1.0 / rD[c];
    return rD;
}

static void of_dic_substitute(const of_mat* A, const double* rD, double* w)
{
    for (int f = 0; f < A->nFaces; ++f) w[A->up[f]] -= rD[A->up[f]] * A->upper[f] * w[A->lo[f]];
    for (int f = A->nFaces - 1; f >= 0; --f) w[A->lo[f]] -= rD[A->lo[f]] * A->upper[f] * w[A->up[f]];
}

/* smooth_solver_cpp.cu:87-155: GaussSeidel sweeps on a folded system (symmetric adds a reverse sweep) */
static void of_gauss_seidel(const of_mat* A, double* psi, const double* b, int nSweeps, int symmetric)
{
    const int nC = A->nCells;
    double* bPrime = (double*)xcalloc((size_t)nC, sizeof(double));
    for (int sweep = 0; sweep < nSweeps; ++sweep)
    {
        memcpy(bPrime, b, (size_t)nC * sizeof(double));
        for (int c = 0; c < nC; ++c)
        {
            const int fStart = A->ownStart[c], fEnd = A->ownStart[c + 1];
            double psii = bPrime[c];
            for (int f = fStart; f < fEnd; ++f) psii -= A->upper[f] * psi[A->up[f]];
            psii /= A->diag[c];
            for (int f = fStart; f < fEnd; ++f) bPrime[A->up[f]] -= A->upper[f] * psii;
            psi[c] = psii;
        }
        if (symmetric)
        {
            /* the reverse walk starts from the bPrime the forward walk left (smooth_solver_cpp.cu:141-163):
             * each cell's lower-neighbour terms carry the forward sweep's values */
            for (int c = nC - 1; c >= 0; --c)
            {
                const int fStart = A->ownStart[c], fEnd = A->ownStart[c + 1];
                double psii = bPrime[c];
                for (int f = fStart; f < fEnd; ++f) psii -= A->upper[f] * psi[A->up[f]];
                psii /= A->diag[c];
                psi[c] = psii;   /* nothing to distribute: the upper neighbours are not visited again */
            }
        }
    }
    free(bPrime);
}

static void of_smooth(const of_mat* A, int smoother, double* psi, const double* source, int nSweeps)
{
    const int dic = (smoother == 2 || smoother == 3);
    const int gs = (smoother == 0 || smoother == 3);
    const int sgs = (smoother == 1);
    if (dic)
    {
        double* rA = (double*)xcalloc((size_t)A->nCells, sizeof(double));
        for (int sweep = 0; sweep < nSweeps; ++sweep)
        {
            of_residual(A, rA, psi, source);
            for (int i = 0; i < A->nCells; ++i) rA[i] *= A->rD[i];
            of_dic_substitute(A, A->rD, rA);
            for (int i = 0; i < A->nCells; ++i) psi[i] += rA[i];
        }
        free(rA);
    }
    if (gs || sgs) of_gauss_seidel(A, psi, source, nSweeps, sgs);
}

/* gamg_solver_cpp.cu:321-383: PCG with DIC, the coarsest-level solver (maxIter 1000, singularity guard) */
static orc_perf of_pcg_dic(const of_mat* A, double* psi, const double* source, double tol, double relTol)
{
    const int maxIter = 1000;
    const int nC = A->nCells;
    double* pA = (double*)xcalloc((size_t)nC, sizeof(double));
    double* wA = (double*)xcalloc((size_t)nC, sizeof(double));
    double* rA = (double*)xcalloc((size_t)nC, sizeof(double));
    double wArA = PERF_GREAT, wArAold = wArA;
    of_amul(A, psi, wA);
    for (int c = 0; c < nC; ++c) rA[c] = source[c] - wA[c];
    const double nf = of_normfactor(A, psi, source, wA);
    orc_perf perf;
    perf.initialResidual = orc_sum_mag(nC, rA) / nf;
    perf.finalResidual = perf.initialResidual;
    perf.nIterations = 0;
#define OF_CONV(p) ((p).finalResidual < tol || (relTol > PERF_SMALL && (p).finalResidual < relTol * (p).initialResidual))
    if (!OF_CONV(perf))
    {
        double* rD = of_dic_rd(A);
        do
        {
            wArAold = wArA;
            for (int c = 0; c < nC; ++c) wA[c] = rD[c] * rA[c];
            of_dic_substitute(A, rD, wA);
            wArA = orc_sum_prod(nC, wA, rA);
            if (perf.nIterations == 0) memcpy(pA, wA, (size_t)nC * sizeof(double));
            else
            {
                const double beta = wArA / wArAold;
                for (int c = 0; c < nC; ++c) pA[c] = wA[c] + beta * pA[c];
            }
            of_amul(A, pA, wA);
            const double wApA = orc_sum_prod(nC, wA, pA);
            if (fabs(wApA) / nf < PERF_VSMALL) break;
            const double alpha = wArA / wApA;
            for (int c = 0; c < nC; ++c)
            {
                psi[c] += alpha * pA[c];
                rA[c] -= alpha * wA[c];
            }
            perf.finalResidual = orc_sum_mag(nC, rA) / nf;
        } while (++perf.nIterations < maxIter && !OF_CONV(perf));
        free(rD);
    }
    free(pA); free(wA); free(rA);
    return perf;
}

/* GAMGSolver::scale, gamg_solver_cpp.cu:386-407 */
static void of_scale(double* field, double* Acf, const of_mat* A, const double* source)
{
    of_amul(A, field, Acf);
    double sfNum = 0, sfDen = 0;
    for (int i = 0; i < A->nCells; ++i)
    {
        sfNum += field[i] * source[i];
        sfDen += field[i] * Acf[i];
    }
    const double sf = sfNum / (sfDen >= 0 ? sfDen + PERF_VSMALL : sfDen - PERF_VSMALL);
    for (int i = 0; i < A->nCells; ++i) field[i] = sf * field[i] + (source[i] - sf * Acf[i]) / A->diag[i];
}

static int* own_start(int nCells, int nFaces, const int* lo)
{
    int* os = (int*)xcalloc((size_t)nCells + 1, sizeof(int));
    for (int f = 0; f < nFaces; ++f) ++os[lo[f] + 1];
    for (int c = 0; c < nCells; ++c) os[c + 1] += os[c];
    return os;
}

typedef struct
{
    const orc_aggl* a;
    of_mat fine;
    of_mat* lv;               /* coarse matrices, nLevels */
    double** corr;            /* coarseCorrFields */
    double** src;             /* coarseSources */
    double* ACf;
    double* preSmoothed;
} of_hier;

static void of_hier_build(of_hier* h, const orc_aggl* a, const double* diag, const double* upper, int smoother,
                          const int* fineOwn, const int* fineNei)
{
    memset(h, 0, sizeof(*h));
    h->a = a;
    const int nL = a->nLevels;
    of_mat* F = &h->fine;
    F->nCells = a->fineCells;
    F->nFaces = a->fineFaces;
    F->lo = fineOwn;
    F->up = fineNei;
    F->diag = (double*)xcalloc((size_t)F->nCells, sizeof(double));
    F->upper = (double*)xcalloc((size_t)F->nFaces, sizeof(double));
    memcpy(F->diag, diag, (size_t)F->nCells * sizeof(double));
    memcpy(F->upper, upper, (size_t)F->nFaces * sizeof(double));
    h->lv = (of_mat*)xcalloc((size_t)nL, sizeof(of_mat));
    h->corr = (double**)xcalloc((size_t)nL, sizeof(double*));
    h->src = (double**)xcalloc((size_t)nL, sizeof(double*));
    for (int i = 0; i < nL; ++i)
    {
        const orc_level* L = &a->lv[i];
        const of_mat* fm = i == 0 ? F : &h->lv[i - 1];
        of_mat* cm = &h->lv[i];
        cm->nCells = L->nCoarse;
        cm->nFaces = L->nCoarseFaces;
        cm->lo = L->cOwn;
        cm->up = L->cNei;
        cm->diag = (double*)xcalloc((size_t)cm->nCells, sizeof(double));
        cm->upper = (double*)xcalloc((size_t)cm->nFaces, sizeof(double));
        for (int c = 0; c < L->nFine; ++c) cm->diag[L->map[c]] += fm->diag[c];
        for (int f = 0; f < L->nFineFaces; ++f)
        {
            const int cf = L->faceRestrict[f];
            if (cf >= 0) cm->upper[cf] += fm->upper[f];
            else cm->diag[-1 - cf] += 2 * fm->upper[f];
        }
        h->corr[i] = (double*)xcalloc((size_t)cm->nCells, sizeof(double));
        h->src[i] = (double*)xcalloc((size_t)cm->nCells, sizeof(double));
    }
    for (int i = -1; i < nL; ++i)
    {
        of_mat* m = i < 0 ? F : &h->lv[i];
        m->ownStart = own_start(m->nCells, m->nFaces, m->lo);
        m->rD = (smoother == 2 || smoother == 3) ? of_dic_rd(m) : NULL;
    }
    h->ACf = (double*)xcalloc((size_t)F->nCells, sizeof(double));
    h->preSmoothed = (double*)xcalloc((size_t)F->nCells, sizeof(double));
}

static void of_mat_free(of_mat* m)
{
    free(m->diag); free(m->upper); free(m->ownStart); free(m->rD);
}

/* The GAMG restatement's own Gauss-Seidel sweeps and DIC preconditioner, on a plain symmetric system
 * (exported so the device's level-scheduled sweeps can be held to them). */
static of_mat of_wrap(int nCells, int nFaces, const int* owner, const int* neighbour, const double* diag,
                      const double* upper)
{
    of_mat m;
    m.nCells = nCells;
    m.nFaces = nFaces;
    m.lo = owner;
    m.up = neighbour;
    m.diag = (double*)xcalloc((size_t)nCells, sizeof(double));
    m.upper = (double*)xcalloc((size_t)(nFaces ? nFaces : 1), sizeof(double));
    memcpy(m.diag, diag, (size_t)nCells * sizeof(double));
    if (nFaces) memcpy(m.upper, upper, (size_t)nFaces * sizeof(double));
    m.ownStart = own_start(nCells, nFaces, owner);
    m.rD = NULL;
    return m;
}

void orc_gauss_seidel(int nCells, int nFaces, const int* owner, const int* neighbour, const double* diag,
                      const double* upper, const double* b, double* psi, int nSweeps)
{
    of_mat m = of_wrap(nCells, nFaces, owner, neighbour, diag, upper);
    of_gauss_seidel(&m, psi, b, nSweeps, 0);
    of_mat_free(&m);
}

void orc_sym_gauss_seidel(int nCells, int nFaces, const int* owner, const int* neighbour, const double* diag,
                          const double* upper, const double* b, double* psi, int nSweeps)
{
    of_mat m = of_wrap(nCells, nFaces, owner, neighbour, diag, upper);
    of_gauss_seidel(&m, psi, b, nSweeps, 1);
    of_mat_free(&m);
}

void orc_dic_precondition(int nCells, int nFaces, const int* owner, const int* neighbour, const double* diag,
                          const double* upper, const double* r, double* w)
{
    of_mat m = of_wrap(nCells, nFaces, owner, neighbour, diag, upper);
    double* rD = of_dic_rd(&m);
    for (int c = 0; c < nCells; ++c) w[c] = rD[c] * r[c];
    of_dic_substitute(&m, rD, w);
    free(rD);
    of_mat_free(&m);
}


static void of_hier_free(of_hier* h)
{
    of_mat_free(&h->fine);
    for (int i = 0; i < h->a->nLevels; ++i)
    {
        of_mat_free(&h->lv[i]);
        free(h->corr[i]);
        free(h->src[i]);
    }
    free(h->lv); free(h->corr); free(h->src); free(h->ACf); free(h->preSmoothed);
}

static void restrict_field(double* cf, int nCoarse, const double* ff, int nFine, const int* map)
{
    for (int c = 0; c < nCoarse; ++c) cf[c] = 0.0;
    for (int i = 0; i < nFine; ++i) cf[map[i]] += ff[i];
}

static void prolong_field(double* ff, int nFine, const double* cf, const int* map)
{
    for (int i = 0; i < nFine; ++i) ff[i] = cf[map[i]];
}

static int imin(int a, int b) { return a < b ? a : b; }

static void of_vcycle(of_hier* h, const orc_gamg_controls* c, double* psi, const double* source, double* Apsi,
                      double* finestCorrection, const double* finestResidual)
{
    const orc_aggl* a = h->a;
    const int coarsest = a->nLevels - 1;
    restrict_field(h->src[0], a->lv[0].nCoarse, finestResidual, a->lv[0].nFine, a->lv[0].map);
    for (int li = 0; li < coarsest; ++li)
    {
        const int n = h->lv[li].nCells;
        if (c->nPreSweeps)
        {
            memset(h->corr[li], 0, (size_t)n * sizeof(double));
            of_smooth(&h->lv[li], c->smoother, h->corr[li], h->src[li],
                      imin(c->nPreSweeps + c->preSweepsLevelMultiplier * li, c->maxPreSweeps));
            if (c->scaleCorrection && li < coarsest - 1) of_scale(h->corr[li], h->ACf, &h->lv[li], h->src[li]);
            of_residual(&h->lv[li], h->src[li], h->corr[li], h->src[li]);
        }
        restrict_field(h->src[li + 1], a->lv[li + 1].nCoarse, h->src[li], a->lv[li + 1].nFine, a->lv[li + 1].map);
    }
    memset(h->corr[coarsest], 0, (size_t)h->lv[coarsest].nCells * sizeof(double));
    of_pcg_dic(&h->lv[coarsest], h->corr[coarsest], h->src[coarsest], c->tolerance, c->relTol);
    for (int li = coarsest - 1; li >= 0; --li)
    {
        const int n = h->lv[li].nCells;
        if (c->nPreSweeps) memcpy(h->preSmoothed, h->corr[li], (size_t)n * sizeof(double));
        prolong_field(h->corr[li], n, h->corr[li + 1], a->lv[li + 1].map);
        if (c->scaleCorrection && li < coarsest - 1) of_scale(h->corr[li], h->ACf, &h->lv[li], h->src[li]);
        if (c->nPreSweeps)
            for (int i = 0; i < n; ++i) h->corr[li][i] += h->preSmoothed[i];
        of_smooth(&h->lv[li], c->smoother, h->corr[li], h->src[li],
                  imin(c->nPostSweeps + c->postSweepsLevelMultiplier * li, c->maxPostSweeps));
    }
    prolong_field(finestCorrection, h->fine.nCells, h->corr[0], a->lv[0].map);
    if (c->scaleCorrection) of_scale(finestCorrection, Apsi, &h->fine, finestResidual);
    for (int i = 0; i < h->fine.nCells; ++i) psi[i] += finestCorrection[i];
    of_smooth(&h->fine, c->smoother, psi, source, c->nFinestSweeps);
}

void orc_gamg_default_controls(orc_gamg_controls* c)
{
    c->smoother = 0;
    c->tolerance = 1e-6;
    c->relTol = 0;
    c->maxIter = 1000;
    c->minIter = 0;
    c->nPreSweeps = 0;
    c->preSweepsLevelMultiplier = 1;
    c->maxPreSweeps = 4;
    c->nPostSweeps = 2;
    c->postSweepsLevelMultiplier = 1;
    c->maxPostSweeps = 4;
    c->nFinestSweeps = 2;
    c->scaleCorrection = 1;
    c->nVcycles = 2;
}

static void aggl_fine_addressing(const orc_aggl* a, const int** own, const int** nei)
{
    *own = a->fineOwn;
    *nei = a->fineNei;
}

orc_perf orc_gamg_solve(const orc_aggl* a, const double* diag, const double* upper, const double* b,
                        double* psi, const orc_gamg_controls* c)
{
    const int *own, *nei;
    aggl_fine_addressing(a, &own, &nei);
    of_hier h;
    of_hier_build(&h, a, diag, upper, c->smoother, own, nei);
    const int nC = a->fineCells;
    double* Apsi = (double*)xcalloc((size_t)nC, sizeof(double));
    double* fCorr = (double*)xcalloc((size_t)nC, sizeof(double));
    double* fRes = (double*)xcalloc((size_t)nC, sizeof(double));
    of_amul(&h.fine, psi, Apsi);
    const double nf = of_normfactor(&h.fine, psi, b, Apsi);
    for (int i = 0; i < nC; ++i) fRes[i] = b[i] - Apsi[i];
    orc_perf perf;
    perf.initialResidual = orc_sum_mag(nC, fRes) / nf;
    perf.finalResidual = perf.initialResidual;
    perf.nIterations = 0;
    const double tol = c->tolerance, relTol = c->relTol;
    if (!(c->minIter <= 0 && OF_CONV(perf)))
    {
        do
        {
            of_vcycle(&h, c, psi, b, Apsi, fCorr, fRes);
            of_amul(&h.fine, psi, Apsi);
            for (int i = 0; i < nC; ++i)
            {
                fRes[i] = b[i];
                fRes[i] -= Apsi[i];
            }
            perf.finalResidual = orc_sum_mag(nC, fRes) / nf;
        } while ((++perf.nIterations < c->maxIter && !OF_CONV(perf)) || perf.nIterations < c->minIter);
    }
    free(Apsi); free(fCorr); free(fRes);
    of_hier_free(&h);
    return perf;
}

orc_perf orc_pcg_gamg_solve(const orc_aggl* a, const double* diag, const double* upper, const double* b,
                            double* psi, double tolerance, double relTolerance, int maxIter, int minIter,
                            const orc_gamg_controls* c)
{
    const int *own, *nei;
    aggl_fine_addressing(a, &own, &nei);
    const int nC = a->fineCells;
    of_mat A;
    memset(&A, 0, sizeof(A));
    A.nCells = nC;
    A.nFaces = a->fineFaces;
    A.lo = own;
    A.up = nei;
    A.diag = (double*)diag;
    A.upper = (double*)upper;
    double* pA = (double*)xcalloc((size_t)nC, sizeof(double));
    double* wA = (double*)xcalloc((size_t)nC, sizeof(double));
    double* rA = (double*)xcalloc((size_t)nC, sizeof(double));
    double wArA = PERF_GREAT, wArAold = wArA;
    of_amul(&A, psi, wA);
    for (int i = 0; i < nC; ++i) rA[i] = b[i] - wA[i];
    const double nf = of_normfactor(&A, psi, b, wA);
    orc_perf perf;
    perf.initialResidual = orc_sum_mag(nC, rA) / nf;
    perf.finalResidual = perf.initialResidual;
    perf.nIterations = 0;
    const double tol = tolerance, relTol = relTolerance;
    if (minIter <= 0 && OF_CONV(perf))
    {
        free(pA); free(wA); free(rA);
        return perf;
    }
    of_hier h;
    of_hier_build(&h, a, diag, upper, c->smoother, own, nei);
    double* AwA = (double*)xcalloc((size_t)nC, sizeof(double));
    double* fCorr = (double*)xcalloc((size_t)nC, sizeof(double));
    double* fRes = (double*)xcalloc((size_t)nC, sizeof(double));
    do
    {
        wArAold = wArA;
        /* GAMGPreconditioner::precondition: nVcycles from zero */
        memset(wA, 0, (size_t)nC * sizeof(double));
        memcpy(fRes, rA, (size_t)nC * sizeof(double));
        for (int cyc = 0; cyc < c->nVcycles; ++cyc)
        {
            of_vcycle(&h, c, wA, rA, AwA, fCorr, fRes);
            if (cyc < c->nVcycles - 1)
            {
                of_amul(&A, wA, AwA);
                memcpy(fRes, rA, (size_t)nC * sizeof(double));
                for (int i = 0; i < nC; ++i) fRes[i] -= AwA[i];
            }
        }
        wArA = orc_sum_prod(nC, wA, rA);
        if (perf.nIterations == 0) memcpy(pA, wA, (size_t)nC * sizeof(double));
        else
        {
            const double beta = wArA / wArAold;
            for (int i = 0; i < nC; ++i) pA[i] = wA[i] + beta * pA[i];
        }
        of_amul(&A, pA, wA);
        const double wApA = orc_sum_prod(nC, wA, pA);
        if (fabs(wApA) / nf < PERF_VSMALL) break;
        const double alpha = wArA / wApA;
        for (int i = 0; i < nC; ++i)
        {
            psi[i] += alpha * pA[i];
            rA[i] -= alpha * wA[i];
        }
        perf.finalResidual = orc_sum_mag(nC, rA) / nf;
    } while ((++perf.nIterations < maxIter && !OF_CONV(perf)) || perf.nIterations < minIter);
    free(pA); free(wA); free(rA); free(AwA); free(fCorr); free(fRes);
    of_hier_free(&h);
    return perf;
}
#undef OF_CONV

/* ------------------------------------------------------------------------------------------------
 * brae AMG-PCG (device path semantics). Jacobi smoother with the safeDiag floor
 * (device_amg_detail.cuh:248-254), omega 0.8; dense LU with partial pivoting at the coarsest level
 * (gamg.cu:222-263 style elimination). `single` runs the cycle in float as the FP32 path does.
 * ---------------------------------------------------------------------------------------------- */
typedef struct
{
    int n, nf;
    const int* own;
    const int* nei;
    double* diag;
    double* upper;
    float* fdiag;
    float* fupper;
} br_mat;

static double safe_diag_d(double d)
{
    const double fl = 1e-300;
    if (d > fl || d < -fl) return d;
    return d < 0 ? -fl : fl;
}
static float safe_diag_f(float d)
{
    const float fl = 1e-30f;
    if (d > fl || d < -fl) return d;
    return d < 0 ? -fl : fl;
}

/* face-order product: neighbour faces ascending then owned faces ascending per cell == global face order */
static void br_amul_d(const br_mat* A, const double* x, double* y)
{
    for (int c = 0; c < A->n; ++c) y[c] = A->diag[c] * x[c];
    for (int f = 0; f < A->nf; ++f)
    {
        y[A->nei[f]] += A->upper[f] * x[A->own[f]];
        y[A->own[f]] += A->upper[f] * x[A->nei[f]];
    }
}
static void br_amul_f(const br_mat* A, const float* x, float* y)
{
    for (int c = 0; c < A->n; ++c) y[c] = A->fdiag[c] * x[c];
    for (int f = 0; f < A->nf; ++f)
    {
        y[A->nei[f]] += A->fupper[f] * x[A->own[f]];
        y[A->own[f]] += A->fupper[f] * x[A->nei[f]];
    }
}

/* Explicit inverse of the coarsest grid: the reference's default coarse solve kind is Inverse
 * (device_amg_coarse_default.cuh:13); formed here by Gauss-Jordan with partial pivoting (first
 * largest |a_ik|, i >= k) and the pivot floor 1e-14*max|diag| (device_amg_detail.cuh:93-94). */
static void gauss_jordan_inverse(int n, const double* A, const double* diag, double* inv)
{
    const int w = 2 * n;
    double* M = (double*)xcalloc((size_t)n * w, sizeof(double));
    double* fcol = (double*)xcalloc((size_t)n, sizeof(double));
    for (int i = 0; i < n; ++i)
    {
        for (int j = 0; j < n; ++j) M[(size_t)i * w + j] = A[(size_t)i * n + j];
        M[(size_t)i * w + n + i] = 1.0;
    }
    double scale = 0.0;
    for (int c = 0; c < n; ++c) scale = fmax(scale, fabs(diag[c]));
    if (!(scale > 0.0)) scale = 1.0;
    for (int k = 0; k < n; ++k)
    {
        double bv = -1.0;
        int p = k;
        for (int i = k; i < n; ++i)
            if (fabs(M[(size_t)i * w + k]) > bv) { bv = fabs(M[(size_t)i * w + k]); p = i; }
        if (p != k)
            for (int j = 0; j < w; ++j)
            {
                double t = M[(size_t)k * w + j];
                M[(size_t)k * w + j] = M[(size_t)p * w + j];
                M[(size_t)p * w + j] = t;
            }
        double d = M[(size_t)k * w + k];
        const double fl = 1e-14 * scale;
        if (!(fabs(d) > fl)) d = d < 0.0 ? -fl : fl;
        M[(size_t)k * w + k] = d;
        const double rp = 1.0 / d;
        for (int j = 0; j < w; ++j) M[(size_t)k * w + j] *= rp;
        for (int i = 0; i < n; ++i) fcol[i] = M[(size_t)i * w + k];
        for (int i = 0; i < n; ++i)
        {
            if (i == k) continue;
            for (int j = 0; j < w; ++j) M[(size_t)i * w + j] -= fcol[i] * M[(size_t)k * w + j];
        }
    }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) inv[(size_t)i * n + j] = M[(size_t)i * w + n + j];
    free(M);
    free(fcol);
}

static void dense_apply(int n, const double* inv, const double* b, double* x)
{
    for (int i = 0; i < n; ++i)
    {
        double s = 0.0;
        for (int j = 0; j < n; ++j) s += inv[(size_t)i * n + j] * b[j];
        x[i] = s;
    }
}

typedef struct
{
    const orc_aggl* a;
    int nG;            /* grids = levels + 1 */
    br_mat* g;         /* g[0] fine */
    double* dense;     /* coarsest dense inverse */
    int nPre, nPost, single, scale;
    double **b, **x, **r, **Ax;
    float **bf, **xf, **rf, **Axf;
} br_hier;

static void br_vcycle_d(br_hier* H, int g)
{
    br_mat* A = &H->g[g];
    const int n = A->n;
    double *b = H->b[g], *x = H->x[g], *r = H->r[g], *Ax = H->Ax[g];
    if (g == H->nG - 1)
    {
        dense_apply(n, H->dense, b, x);
        return;
    }
    for (int s = 0; s < H->nPre; ++s)
    {
        if (s == 0)
        {
            for (int i = 0; i < n; ++i) x[i] = 0.0 + 0.8 * (b[i] - 0.0) / safe_diag_d(A->diag[i]);
            continue;
        }
        br_amul_d(A, x, Ax);
        for (int i = 0; i < n; ++i) x[i] += 0.8 * (b[i] - Ax[i]) / safe_diag_d(A->diag[i]);
    }
    if (H->nPre == 0)
        for (int i = 0; i < n; ++i) x[i] = 0.0;
    br_amul_d(A, x, Ax);
    for (int i = 0; i < n; ++i) r[i] = b[i] - Ax[i];
    const orc_level* L = &H->a->lv[g];
    /* restrictGatherT: coarse rhs sums fine residuals in ascending fine cell */
    double* bc = H->b[g + 1];
    for (int c = 0; c < L->nCoarse; ++c) bc[c] = 0.0;
    for (int i = 0; i < n; ++i) bc[L->map[i]] += r[i];
    br_vcycle_d(H, g + 1);
    double* xc = H->x[g + 1];
    if (H->scale == 2)
    {
        /* the reference's own order (device_amg_vcycle.cu:1113-1126): c = P xc (prolongToK), Ac,
         * alpha = (r.c)/(c.Ac) (deviceDotInto x2, scaleFactorK), xg += alpha*c (deviceAxpyDev), then
         * the post-sweeps each with a fresh product */
        double* e = (double*)xcalloc((size_t)n, sizeof(double));
        for (int i = 0; i < n; ++i) e[i] = xc[L->map[i]];
        br_amul_d(A, e, Ax);
        double num = 0.0, den = 0.0;
        for (int i = 0; i < n; ++i) num += r[i] * e[i];
        for (int i = 0; i < n; ++i) den += e[i] * Ax[i];
        const double al = den > 1e-30 ? num / den : 1.0;
        for (int i = 0; i < n; ++i) x[i] = x[i] + al * e[i];
        free(e);
        for (int s = 0; s < H->nPost; ++s)
        {
            br_amul_d(A, x, Ax);
            for (int i = 0; i < n; ++i) x[i] += 0.8 * (b[i] - Ax[i]) / safe_diag_d(A->diag[i]);
        }
        return;
    }
    if (H->scale)
    {
        /* scaled correction fused with the first post-sweep: e = P xc, alpha = (r.e)/(e.Ae)
         * (corrScaling, device_amg_vcycle.cu:1113-1126; guard scaleFactorK :69-79), then
         * x = x + alpha*e + omega*(r - alpha*Ae)/D with r the residual restricted above */
        double* e = (double*)xcalloc((size_t)n, sizeof(double));
        for (int i = 0; i < n; ++i) e[i] = xc[L->map[i]];
        br_amul_d(A, e, Ax);
        double num = 0.0, den = 0.0;
        for (int i = 0; i < n; ++i) num += e[i] * r[i];
        for (int i = 0; i < n; ++i) den += e[i] * Ax[i];
        const double al = den > 1e-30 ? num / den : 1.0;
        for (int i = 0; i < n; ++i)
            x[i] = H->nPost > 0 ? x[i] + al * e[i] + 0.8 * (r[i] - al * Ax[i]) / safe_diag_d(A->diag[i])
                                : x[i] + al * e[i];
        free(e);
        for (int s = 1; s < H->nPost; ++s)
        {
            br_amul_d(A, x, Ax);
            for (int i = 0; i < n; ++i) x[i] += 0.8 * (b[i] - Ax[i]) / safe_diag_d(A->diag[i]);
        }
        return;
    }
    for (int i = 0; i < n; ++i) x[i] += xc[L->map[i]];
    for (int s = 0; s < H->nPost; ++s)
    {
        br_amul_d(A, x, Ax);
        for (int i = 0; i < n; ++i) x[i] += 0.8 * (b[i] - Ax[i]) / safe_diag_d(A->diag[i]);
    }
}

static void br_vcycle_f(br_hier* H, int g)
{
    br_mat* A = &H->g[g];
    const int n = A->n;
    float *b = H->bf[g], *x = H->xf[g], *r = H->rf[g], *Ax = H->Axf[g];
    if (g == H->nG - 1)
    {
        double* bd = H->b[g];
        double* xd = H->x[g];
        for (int i = 0; i < n; ++i) bd[i] = (double)b[i];
        dense_apply(n, H->dense, bd, xd);
        for (int i = 0; i < n; ++i) x[i] = (float)xd[i];
        return;
    }
    const float om = (float)0.8;
    for (int s = 0; s < H->nPre; ++s)
    {
        if (s == 0)
        {
            for (int i = 0; i < n; ++i) x[i] = 0.0f + om * (b[i] - 0.0f) / safe_diag_f(A->fdiag[i]);
            continue;
        }
        br_amul_f(A, x, Ax);
        for (int i = 0; i < n; ++i) x[i] += om * (b[i] - Ax[i]) / safe_diag_f(A->fdiag[i]);
    }
    if (H->nPre == 0)
        for (int i = 0; i < n; ++i) x[i] = 0.0f;
    br_amul_f(A, x, Ax);
    for (int i = 0; i < n; ++i) r[i] = b[i] - Ax[i];
    const orc_level* L = &H->a->lv[g];
    float* bc = H->bf[g + 1];
    for (int c = 0; c < L->nCoarse; ++c) bc[c] = 0.0f;
    for (int i = 0; i < n; ++i) bc[L->map[i]] += r[i];
    br_vcycle_f(H, g + 1);
    float* xc = H->xf[g + 1];
    if (H->scale)
    {
        float* e = (float*)xcalloc((size_t)n, sizeof(float));
        for (int i = 0; i < n; ++i) e[i] = xc[L->map[i]];
        br_amul_f(A, e, Ax);
        double num = 0.0, den = 0.0;
        for (int i = 0; i < n; ++i) num += (double)e[i] * (double)r[i];
        for (int i = 0; i < n; ++i) den += (double)e[i] * (double)Ax[i];
        const float al = (float)(den > 1e-30 ? num / den : 1.0);
        for (int i = 0; i < n; ++i)
            x[i] = H->nPost > 0 ? x[i] + al * e[i] + om * (r[i] - al * Ax[i]) / safe_diag_f(A->fdiag[i])
                                : x[i] + al * e[i];
        free(e);
        for (int s = 1; s < H->nPost; ++s)
        {
            br_amul_f(A, x, Ax);
            for (int i = 0; i < n; ++i) x[i] += om * (b[i] - Ax[i]) / safe_diag_f(A->fdiag[i]);
        }
        return;
    }
    for (int i = 0; i < n; ++i) x[i] += xc[L->map[i]];
    for (int s = 0; s < H->nPost; ++s)
    {
        br_amul_f(A, x, Ax);
        for (int i = 0; i < n; ++i) x[i] += om * (b[i] - Ax[i]) / safe_diag_f(A->fdiag[i]);
    }
}

/* hierarchy values (brae Galerkin), float copies and work vectors for the cycles below */
static void br_hier_build(br_hier* Hp, const orc_aggl* a, const double* diag, const double* upper, int nPre,
                          int nPost, int single, int scale, double** cDp, double** cUp)
{
    const int *own, *nei;
    aggl_fine_addressing(a, &own, &nei);
    br_hier H;
    memset(&H, 0, sizeof(H));
    H.a = a;
    H.nG = a->nLevels + 1;
    H.nPre = nPre;
    H.nPost = nPost;
    H.single = single;
    H.scale = scale;
    H.g = (br_mat*)xcalloc((size_t)H.nG, sizeof(br_mat));
    H.b = (double**)xcalloc((size_t)H.nG, sizeof(double*));
    H.x = (double**)xcalloc((size_t)H.nG, sizeof(double*));
    H.r = (double**)xcalloc((size_t)H.nG, sizeof(double*));
    H.Ax = (double**)xcalloc((size_t)H.nG, sizeof(double*));
    H.bf = (float**)xcalloc((size_t)H.nG, sizeof(float*));
    H.xf = (float**)xcalloc((size_t)H.nG, sizeof(float*));
    H.rf = (float**)xcalloc((size_t)H.nG, sizeof(float*));
    H.Axf = (float**)xcalloc((size_t)H.nG, sizeof(float*));
    size_t totC = 0, totF = 0;
    for (int i = 0; i < a->nLevels; ++i) { totC += (size_t)a->lv[i].nCoarse; totF += (size_t)a->lv[i].nCoarseFaces; }
    double* cD = (double*)xcalloc(totC, sizeof(double));
    double* cU = (double*)xcalloc(totF, sizeof(double));
    orc_galerkin_brae(a, diag, upper, cD, cU);
    size_t oc = 0, of = 0;
    for (int g = 0; g < H.nG; ++g)
    {
        br_mat* A = &H.g[g];
        if (g == 0)
        {
            A->n = a->fineCells; A->nf = a->fineFaces; A->own = own; A->nei = nei;
            A->diag = (double*)diag; A->upper = (double*)upper;
        }
        else
        {
            const orc_level* L = &a->lv[g - 1];
            A->n = L->nCoarse; A->nf = L->nCoarseFaces; A->own = L->cOwn; A->nei = L->cNei;
            A->diag = cD + oc; A->upper = cU + of;
            oc += (size_t)L->nCoarse; of += (size_t)L->nCoarseFaces;
        }
        A->fdiag = (float*)xcalloc((size_t)A->n, sizeof(float));
        A->fupper = (float*)xcalloc((size_t)A->nf, sizeof(float));
        for (int i = 0; i < A->n; ++i) A->fdiag[i] = (float)A->diag[i];
        for (int i = 0; i < A->nf; ++i) A->fupper[i] = (float)A->upper[i];
        H.b[g] = (double*)xcalloc((size_t)A->n, sizeof(double));
        H.x[g] = (double*)xcalloc((size_t)A->n, sizeof(double));
        H.r[g] = (double*)xcalloc((size_t)A->n, sizeof(double));
        H.Ax[g] = (double*)xcalloc((size_t)A->n, sizeof(double));
        H.bf[g] = (float*)xcalloc((size_t)A->n, sizeof(float));
        H.xf[g] = (float*)xcalloc((size_t)A->n, sizeof(float));
        H.rf[g] = (float*)xcalloc((size_t)A->n, sizeof(float));
        H.Axf[g] = (float*)xcalloc((size_t)A->n, sizeof(float));
    }
    {
        br_mat* Ac = &H.g[H.nG - 1];
        const int n = Ac->n;
        double* dense = (double*)xcalloc((size_t)n * n, sizeof(double));
        for (int c = 0; c < n; ++c) dense[(size_t)c * n + c] = Ac->diag[c];
        for (int f = 0; f < Ac->nf; ++f)
        {
            dense[(size_t)Ac->own[f] * n + Ac->nei[f]] += Ac->upper[f];
            dense[(size_t)Ac->nei[f] * n + Ac->own[f]] += Ac->upper[f];
        }
        H.dense = (double*)xcalloc((size_t)n * n, sizeof(double));
        gauss_jordan_inverse(n, dense, Ac->diag, H.dense);
        free(dense);
    }
    *Hp = H;
    *cDp = cD;
    *cUp = cU;
}

static void br_hier_free(br_hier* Hp, double* cD, double* cU)
{
    br_hier H = *Hp;
    for (int g = 0; g < H.nG; ++g)
    {
        free(H.g[g].fdiag); free(H.g[g].fupper);
        free(H.b[g]); free(H.x[g]); free(H.r[g]); free(H.Ax[g]);
        free(H.bf[g]); free(H.xf[g]); free(H.rf[g]); free(H.Axf[g]);
    }
    free(H.g); free(H.b); free(H.x); free(H.r); free(H.Ax); free(H.bf); free(H.xf); free(H.rf); free(H.Axf);
    free(H.dense);
    free(cD);
    free(cU);
}

void orc_brae_vcycle(const orc_aggl* a, const double* diag, const double* upper, const double* r, double* z,
                     int nPre, int nPost, int single, int scale)
{
    br_hier H;
    double *cD, *cU;
    br_hier_build(&H, a, diag, upper, nPre, nPost, single, scale, &cD, &cU);
    const int nC = a->fineCells;
    if (single)
    {
        for (int i = 0; i < nC; ++i) H.bf[0][i] = (float)r[i];
        br_vcycle_f(&H, 0);
        for (int i = 0; i < nC; ++i) z[i] = (double)H.xf[0][i];
    }
    else
    {
        memcpy(H.b[0], r, (size_t)nC * sizeof(double));
        br_vcycle_d(&H, 0);
        memcpy(z, H.x[0], (size_t)nC * sizeof(double));
    }
    br_hier_free(&H, cD, cU);
}

orc_perf orc_brae_amgpcg(const orc_aggl* a, const double* diag, const double* upper, const double* b,
                         double* psi, double tol, double relTol, int maxIter, int minIter,
                         int nPre, int nPost, int flexible, int single, int scale, double normFactorIn,
                         double* psiTrace)
{
    const int *own, *nei;
    aggl_fine_addressing(a, &own, &nei);
    br_hier H;
    double *cD, *cU;
    br_hier_build(&H, a, diag, upper, nPre, nPost, single, scale, &cD, &cU);
    const int nC = a->fineCells;
    br_mat* A0 = &H.g[0];
    double* pA = (double*)xcalloc((size_t)nC, sizeof(double));
    double* wA = (double*)xcalloc((size_t)nC, sizeof(double));
    double* rA = (double*)xcalloc((size_t)nC, sizeof(double));
    double* rOld = (double*)xcalloc((size_t)nC, sizeof(double));
    br_amul_d(A0, psi, wA);
    for (int i = 0; i < nC; ++i) rA[i] = b[i] - wA[i];
    const double nf = normFactorIn > 0.0 ? normFactorIn
                                         : norm_factor_with(nC, A0->nf, own, nei, diag, upper, upper, psi, b, wA);
    orc_perf perf;
    perf.initialResidual = orc_sum_mag(nC, rA) / nf;
    perf.finalResidual = perf.initialResidual;
    perf.nIterations = 0;
    if (psiTrace) memcpy(psiTrace, psi, (size_t)nC * sizeof(double));
#define BR_CONV(fr) ((fr) < tol || (relTol > 0.0 && (fr) < relTol * perf.initialResidual))
    double wArA = 1e300, wArAold = 0.0, rhoOld = 0.0;
    if (minIter > 0 || !BR_CONV(perf.finalResidual))
    {
        int it = 0;
        for (;;)
        {
            /* w = M^-1 r */
            if (single)
            {
                for (int i = 0; i < nC; ++i) H.bf[0][i] = (float)rA[i];
                br_vcycle_f(&H, 0);
                for (int i = 0; i < nC; ++i) wA[i] = (double)H.xf[0][i];
            }
            else
            {
                memcpy(H.b[0], rA, (size_t)nC * sizeof(double));
                br_vcycle_d(&H, 0);
                memcpy(wA, H.x[0], (size_t)nC * sizeof(double));
            }
            wArAold = wArA;
            wArA = orc_sum_prod(nC, wA, rA);
            if (it == 0)
                memcpy(pA, wA, (size_t)nC * sizeof(double));
            else
            {
                double beta;
                if (flexible)
                {
                    const double zrOld = orc_sum_prod(nC, wA, rOld);
                    const double d = rhoOld;
                    beta = (d > 1e-300 || d < -1e-300) ? fmax(0.0, (wArA - zrOld) / d) : 0.0;
                }
                else
                {
                    const double d = wArAold;
                    beta = fabs(d) > 1e-300 ? wArA / d : 0.0;
                }
                for (int i = 0; i < nC; ++i) pA[i] = beta * pA[i] + wA[i];
            }
            if (flexible) memcpy(rOld, rA, (size_t)nC * sizeof(double));
            rhoOld = wArA;
            br_amul_d(A0, pA, wA);
            const double wApA = orc_sum_prod(nC, wA, pA);
            const double alpha = fabs(wApA) > 1e-300 ? wArA / wApA : 0.0;
            for (int i = 0; i < nC; ++i)
            {
                psi[i] += alpha * pA[i];
                rA[i] -= alpha * wA[i];
            }
            perf.finalResidual = orc_sum_mag(nC, rA) / nf;
            ++it;
            if (psiTrace) memcpy(psiTrace + (size_t)it * nC, psi, (size_t)nC * sizeof(double));
            const int conv = BR_CONV(perf.finalResidual);
            if ((conv || it >= maxIter) && it >= minIter) break;
        }
        perf.nIterations = it;
    }
#undef BR_CONV
    br_hier_free(&H, cD, cU);
    free(pA); free(wA); free(rA); free(rOld);
    return perf;
}

/* ---------------------------------------------------------------------------------------------
 * The SIMPLE pressure predictor's momentum glue: the host loops of pEqn_cpp.cu:40-133 and
 * fv_matrix_ops.cuh:211-340, as written (no contraction).
 * ------------------------------------------------------------------------------------------- */

void orc_fold_boundary_diag(int nCells, int nBFaces, const int* bFaceCell, const double* icX, const double* icY,
                            const double* icZ, const double* diag, double* D)
{
    memcpy(D, diag, sizeof(double) * (size_t)nCells);
    for (int i = 0; i < nBFaces; ++i) D[bFaceCell[i]] += (icX[i] + icY[i] + icZ[i]) / 3.0;
}

void orc_reciprocal_v(int nCells, const double* D, const double* V, double* rAU)
{
    for (int c = 0; c < nCells; ++c)
    {
        const double A = D[c] / V[c];
        rAU[c] = 1.0 / A;
    }
}

void orc_boundary_diag_term(int nBFaces, const double* icX, const double* icY, const double* icZ, int k,
                            double* out)
{
    const double* ic = k == 0 ? icX : (k == 1 ? icY : icZ);
    for (int i = 0; i < nBFaces; ++i)
    {
        const double av = (icX[i] + icY[i] + icZ[i]) / 3.0;
        out[i] = av - ic[i];
    }
}

void orc_matrix_h(int nCells, int nFaces, const int* owner, const int* neighbour, const double* upper,
                  const double* lower, const double* psi, const double* source, int nBFaces, const int* bFaceCell,
                  const double* bdDiag, const double* bdSrc, const double* V, double* H)
{
    memset(H, 0, sizeof(double) * (size_t)nCells);
    for (int i = 0; i < nBFaces; ++i)
    {
        const int c = bFaceCell[i];
        H[c] += bdDiag[i] * psi[c];
    }
    for (int f = 0; f < nFaces; ++f)
    {
        const int o = owner[f], n = neighbour[f];
        H[n] = H[n] - lower[f] * psi[o];
        H[o] = H[o] - upper[f] * psi[n];
    }
    for (int c = 0; c < nCells; ++c) H[c] += source[c];
    for (int i = 0; i < nBFaces; ++i) H[bFaceCell[i]] += bdSrc[i];
    for (int c = 0; c < nCells; ++c) H[c] = H[c] / V[c];
}

/* fvMatrix::relax(alpha) for uncoupled patches (fv_matrix_ops.cuh:158-207): sumOff from |upper| (owner) and
 * |lower| (neighbour) in face order; diag += cmptMagMax(ic) face by face; diag = max(|diag|, sumOff)/alpha;
 * diag -= cmptMin(ic) face by face; source_k += (diag - D0)*psi_k for each of nComp components. icMag/icMin:
 * per boundary face cmptMagMax / cmptMin of the internalCoeffs. cycSumOff (may be NULL) adds each cell's
 * coupled-interface magnitude after the internal faces. alpha <= 0: nothing changes. diag and the
 * sources in place; delta (may be NULL) = relaxed - D0 */
void orc_relax_matrix(int nCells, int nFaces, const int* owner, const int* neighbour, const double* upper,
                      const double* lower, int nBFaces, const int* bFaceCell, const double* icMag,
                      const double* icMin, const double* cycSumOff, double alpha, double* diag, int nComp,
                      const double* const* psi, double* const* source, double* delta)
{
    if (alpha <= 0.0)
    {
        if (delta) memset(delta, 0, sizeof(double) * (size_t)nCells);
        return;
    }
    double* D0 = (double*)malloc(sizeof(double) * (size_t)nCells);
    double* sumOff = (double*)calloc((size_t)nCells, sizeof(double));
    memcpy(D0, diag, sizeof(double) * (size_t)nCells);
    for (int f = 0; f < nFaces; ++f)
    {
        sumOff[owner[f]] += fabs(upper[f]);
        sumOff[neighbour[f]] += fabs(lower[f]);
    }
    if (cycSumOff)
        for (int c = 0; c < nCells; ++c) sumOff[c] += cycSumOff[c];
    for (int i = 0; i < nBFaces; ++i) diag[bFaceCell[i]] += icMag[i];
    for (int c = 0; c < nCells; ++c) diag[c] = fmax(fabs(diag[c]), sumOff[c]) / alpha;
    for (int i = 0; i < nBFaces; ++i) diag[bFaceCell[i]] -= icMin[i];
    for (int c = 0; c < nCells; ++c)
    {
        const double d = diag[c] - D0[c];
        for (int k = 0; k < nComp; ++k) source[k][c] += d * psi[k][c];
        if (delta) delta[c] = d;
    }
    free(D0);
    free(sumOff);
}

/* linearUpwind's deferred correction of div(phi, psi) for one component (fvm.cuh:395-433): per internal
 * face, up = owner if phi > 0 else neighbour, fc = phi * ((Cf - C_up) . grad_up) with the dot product
 * x + y + z as written; corr[owner] += fc, corr[neighbour] -= fc in face order. Boundary faces of uncoupled
 * patches contribute nothing. Cf [3*nFaces], C and grad [3*nCells], xyz interleaved. The caller subtracts
 * corr from the source. */
void orc_linear_upwind_corr(int nCells, int nFaces, const int* owner, const int* neighbour, const double* phi,
                            const double* Cf, const double* C, const double* grad, double* corr)
{
    memset(corr, 0, sizeof(double) * (size_t)nCells);
    for (int f = 0; f < nFaces; ++f)
    {
        const double pf = phi[f];
        const int up = (pf > 0.0) ? owner[f] : neighbour[f];
        const double dx = Cf[3 * (size_t)f] - C[3 * (size_t)up];
        const double dy = Cf[3 * (size_t)f + 1] - C[3 * (size_t)up + 1];
        const double dz = Cf[3 * (size_t)f + 2] - C[3 * (size_t)up + 2];
        const double* g = grad + 3 * (size_t)up;
        const double fc = pf * (dx * g[0] + dy * g[1] + dz * g[2]);
        corr[owner[f]] += fc;
        corr[neighbour[f]] -= fc;
    }
}

/* nuEff*dev2(T(grad(U))) per cell (linearViscousStress_cpp.cu:75-79, cf_types.cuh:174-179): tensors
 * component-major, t[(i*3 + j)*n + c]; gradU_ij = d(U_j)/d(x_i); T(gradU)_ij = gradU_ji; dev2 subtracts
 * (2/3)*tr on the diagonal, tr = xx + yy + zz */
void orc_scaled_dev2t(int n, const double* gradU, const double* nu, double* sigma)
{
    for (int c = 0; c < n; ++c)
    {
        const double tr = gradU[0 * (size_t)n + c] + gradU[4 * (size_t)n + c] + gradU[8 * (size_t)n + c];
        const double k = (2.0 / 3.0) * tr;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
            {
                const double t = gradU[(j * 3 + i) * (size_t)n + c];
                sigma[(i * 3 + j) * (size_t)n + c] = nu[c] * (i == j ? t - k : t);
            }
    }
}

/* V*fvc::div of a tensor field (fvc.cu:867-905 before its division by V): per internal face
 * T_f = w*T_own + (1-w)*T_nei, (Sf & T_f)_j = Sx*T_xj + Sy*T_yj + Sz*T_zj, added to the owner and
 * subtracted from the neighbour in face order; then per boundary face not marked empty, Sf_b & T_b added
 * in boundary order. Tcell [9*nC] and Tbnd [9*nB] component-major; Sf xyz per face, internal then boundary
 * (the boundary's at 3*(nFaces + i)). isEmpty may be NULL. */
void orc_tensor_div(int nCells, int nFaces, const int* owner, const int* neighbour, const double* w,
                    const double* Sf, int nBFaces, const int* bFaceCell, const int* isEmpty, const double* Tcell,
                    const double* Tbnd, double* dX, double* dY, double* dZ)
{
    double* d[3] = {dX, dY, dZ};
    for (int j = 0; j < 3; ++j) memset(d[j], 0, sizeof(double) * (size_t)nCells);
    const size_t nC = (size_t)nCells, nB = (size_t)nBFaces;
    for (int f = 0; f < nFaces; ++f)
    {
        const int o = owner[f], ne = neighbour[f];
        const double wf = w[f];
        double Tf[9];
        for (int q = 0; q < 9; ++q) Tf[q] = wf * Tcell[q * nC + o] + (1.0 - wf) * Tcell[q * nC + ne];
        const double* s = Sf + 3 * (size_t)f;
        for (int j = 0; j < 3; ++j)
        {
            const double r = s[0] * Tf[0 * 3 + j] + s[1] * Tf[1 * 3 + j] + s[2] * Tf[2 * 3 + j];
            d[j][o] += r;
            d[j][ne] = d[j][ne] - r;
        }
    }
    for (int i = 0; i < nBFaces; ++i)
    {
        if (isEmpty && isEmpty[i]) continue;
        const double* s = Sf + 3 * ((size_t)nFaces + i);
        for (int j = 0; j < 3; ++j)
            d[j][bFaceCell[i]] +=
                s[0] * Tbnd[(0 * 3 + j) * nB + i] + s[1] * Tbnd[(1 * 3 + j) * nB + i] + s[2] * Tbnd[(2 * 3 + j) * nB + i];
    }
}

/* cmptMax(cmptMag(v)) and cmptMin(v) of the internalCoeffs per boundary face (cf_types.cuh:199-202) */
void orc_cmpt_mag_max_min(int n, const double* x, const double* y, const double* z, double* mag, double* mn)
{
    for (int i = 0; i < n; ++i)
    {
        mag[i] = fmax(fmax(fabs(x[i]), fabs(y[i])), fabs(z[i]));
        mn[i] = fmin(fmin(x[i], y[i]), z[i]);
    }
}

/* GeometricField::relax (pEqn_cpp.cu:295-300): p = pPrev + alpha*(p - pPrev); nothing unless 0 < alpha < 1 */
void orc_relax_field(int n, double* p, const double* pPrev, double alpha)
{
    if (alpha >= 1.0 || alpha <= 0.0) return;
    for (int c = 0; c < n; ++c) p[c] = pPrev[c] + alpha * (p[c] - pPrev[c]);
}

void orc_vector_flux(int nFaces, const int* owner, const int* neighbour, const double* weights, const double* Sf,
                     const double* Ux, const double* Uy, const double* Uz, double* phi)
{
    for (int f = 0; f < nFaces; ++f)
    {
        const int o = owner[f], n = neighbour[f];
        const double w = weights[f];
        const double ux = w * Ux[o] + (1.0 - w) * Ux[n];
        const double uy = w * Uy[o] + (1.0 - w) * Uy[n];
        const double uz = w * Uz[o] + (1.0 - w) * Uz[n];
        phi[f] = ux * Sf[3 * (size_t)f] + uy * Sf[3 * (size_t)f + 1] + uz * Sf[3 * (size_t)f + 2];
    }
}

void orc_boundary_flux(int nBFaces, const double* SfB, const double* uxb, const double* uyb, const double* uzb,
                       const int* isEmpty, double* phiB)
{
    for (int i = 0; i < nBFaces; ++i)
        phiB[i] = (isEmpty && isEmpty[i]) ? 0.0
                                          : uxb[i] * SfB[3 * (size_t)i] + uyb[i] * SfB[3 * (size_t)i + 1] +
                                                uzb[i] * SfB[3 * (size_t)i + 2];
}

int orc_adjust_phi(int nFaces, const double* phiInt, int nBFaces, const int* adjustable, double* phiB,
                   double* massCorr, int* closedVolume)
{
    const double vsmall = 1e-300, small = 1e-15;
    double massIn = 0.0, fixedOut = 0.0, adjOut = 0.0, totalFlux = vsmall;
    for (int f = 0; f < nFaces; ++f) totalFlux += fabs(phiInt[f]);
    for (int i = 0; i < nBFaces; ++i)
    {
        const double v = phiB[i];
        if (v < 0.0)
            massIn -= v;
        else if (!adjustable[i])
            fixedOut += v;
        else
            adjOut += v;
    }
    double mc = 1.0;
    const double magAdj = fabs(adjOut);
    if (closedVolume)
        *closedVolume = fabs(massIn) / totalFlux < small && fabs(fixedOut) / totalFlux < small &&
                        fabs(adjOut) / totalFlux < small;
    if (magAdj > vsmall && magAdj / totalFlux > small)
        mc = (massIn - fixedOut) / adjOut;
    else if (fabs(fixedOut - massIn) / totalFlux > 1e-8)
    {
        if (massCorr) *massCorr = 1.0;
        return 1;
    }
    for (int i = 0; i < nBFaces; ++i)
        if (adjustable[i] && phiB[i] > 0.0) phiB[i] *= mc;
    if (massCorr) *massCorr = mc;
    return 0;
}

