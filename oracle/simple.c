/*
 * simple.c -- TEST INFRASTRUCTURE ONLY: the FP64 CPU restatement of one SIMPLE iteration of the
 * reference's host path (src/applications/solvers/simpleFoam/simpleFoam_cpp.cu:15-237), with the two
 * linear solvers it calls: PBiCGStab with DILU for the segregated momentum solve
 * (src/OpenFOAM/matrices/pbicgstab.cu:6-150, via solve_vector.cuh:27-116) and the GAMG of
 * src/OpenFOAM/matrices/gamg.cu:267-400 for the pressure. The parity tests, smoke() and bench.py's
 * CPU-baseline leg check the HIP library against it; the product never links or loads it.
 *
 * Every loop keeps the reference's traversal order, because that order fixes the floating-point bits.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static void* zalloc(size_t n, size_t sz)
{
    void* p = calloc(n ? n : 1, sz);
    if (!p) abort();
    return p;
}
#define DALLOC(n) ((double*)zalloc((size_t)(n), sizeof(double)))
#define IALLOC(n) ((int*)zalloc((size_t)(n), sizeof(int)))

/* ------------------------------------------------------------------------------------------------
 * PBiCGStab with the DILU preconditioner on an asymmetric LDU matrix whose uncoupled boundary arrives as
 * internalCoeffs/boundaryCoeffs per boundary face: pbicgstab.cu:6-150. The fold (diag += iC, source += bC
 * face by face), the Amul (diag*x, then lower/upper in face order), the normFactor, the DILU reciprocal
 * diagonal (rD[nei] -= upper*lower/rD[own] in face order, then 1/rD), its forward (face order) and backward
 * (reverse face order) substitutions, and the recurrence with its 1e-300 breakdown guards.
 * ---------------------------------------------------------------------------------------------- */
static void amul_ldu(int nC, int nF, const int* own, const int* nei, const double* d, const double* u,
                     const double* l, const double* x, double* y)
{
    for (int c = 0; c < nC; ++c) y[c] = d[c] * x[c];
    for (int f = 0; f < nF; ++f)
    {
        y[nei[f]] += l[f] * x[own[f]];
        y[own[f]] += u[f] * x[nei[f]];
    }
}

static double sum_mag(int n, const double* x)
{
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += fabs(x[i]);
    return s;
}

static double sum_prod(int n, const double* a, const double* b)
{
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

orc_perf orc_pbicgstab(int nCells, int nFaces, const int* owner, const int* neighbour, const double* diag,
                       const double* upper, const double* lower, const double* source, int nBFaces,
                       const int* bFaceCell, const double* iC, const double* bC, double* psi, double tol,
                       double relTol, int maxIter, int minIter)
{
    const int nC = nCells, nIf = nFaces;
    const int* own = owner;
    const int* nei = neighbour;
    double* diagC = DALLOC(nC);
    double* b = DALLOC(nC);
    memcpy(diagC, diag, sizeof(double) * (size_t)nC);
    memcpy(b, source, sizeof(double) * (size_t)nC);
    for (int i = 0; i < nBFaces; ++i)
    {
        diagC[bFaceCell[i]] += iC[i];
        b[bFaceCell[i]] += bC[i];
    }
    double* rD = DALLOC(nC);
    memcpy(rD, diagC, sizeof(double) * (size_t)nC);
    for (int f = 0; f < nIf; ++f) rD[nei[f]] -= upper[f] * lower[f] / rD[own[f]];
    for (int c = 0; c < nC; ++c) rD[c] = 1.0 / rD[c];

    double* yA = DALLOC(nC);
    double* rA = DALLOC(nC);
    amul_ldu(nC, nIf, own, nei, diagC, upper, lower, psi, yA);
    for (int c = 0; c < nC; ++c) rA[c] = b[c] - yA[c];

    double* sumA = DALLOC(nC);
    for (int c = 0; c < nC; ++c) sumA[c] = diagC[c];
    for (int f = 0; f < nIf; ++f)
    {
        sumA[nei[f]] += lower[f];
        sumA[own[f]] += upper[f];
    }
    double xRef = 0.0;
    for (int c = 0; c < nC; ++c) xRef += psi[c];
    xRef /= nC;
    double normFactor = 0.0;
    for (int c = 0; c < nC; ++c)
    {
        const double t = sumA[c] * xRef;
        normFactor += fabs(yA[c] - t) + fabs(b[c] - t);
    }
    normFactor += 1e-20;

    orc_perf perf;
    perf.initialResidual = sum_mag(nC, rA) / normFactor;
    perf.finalResidual = perf.initialResidual;
    perf.nIterations = 0;
#define CONVERGED(fr) (((fr) < tol) || (relTol > 0.0 && (fr) < relTol * perf.initialResidual))
    if (minIter > 0 || !CONVERGED(perf.finalResidual))
    {
        double* pA = DALLOC(nC);
        double* AyA = DALLOC(nC);
        double* sA = DALLOC(nC);
        double* zA = DALLOC(nC);
        double* tA = DALLOC(nC);
        double* rA0 = DALLOC(nC);
        memcpy(rA0, rA, sizeof(double) * (size_t)nC);
        double rA0rA = 0.0, alpha = 0.0, omega = 0.0;
        int nIter = 0;
        int done = 0;
        do
        {
            const double rA0rAold = rA0rA;
            rA0rA = sum_prod(nC, rA0, rA);
            if (fabs(rA0rA) < 1e-300) break;
            if (nIter == 0)
                memcpy(pA, rA, sizeof(double) * (size_t)nC);
            else
            {
                if (fabs(omega) < 1e-300) break;
                const double beta = (rA0rA / rA0rAold) * (alpha / omega);
                for (int c = 0; c < nC; ++c) pA[c] = rA[c] + beta * (pA[c] - omega * AyA[c]);
            }
            /* yA = M^-1 pA */
            for (int c = 0; c < nC; ++c) yA[c] = rD[c] * pA[c];
            for (int f = 0; f < nIf; ++f) yA[nei[f]] -= rD[nei[f]] * lower[f] * yA[own[f]];
            for (int f = nIf - 1; f >= 0; --f) yA[own[f]] -= rD[own[f]] * upper[f] * yA[nei[f]];
            amul_ldu(nC, nIf, own, nei, diagC, upper, lower, yA, AyA);
            alpha = rA0rA / sum_prod(nC, rA0, AyA);
            for (int c = 0; c < nC; ++c) sA[c] = rA[c] - alpha * AyA[c];
            perf.finalResidual = sum_mag(nC, sA) / normFactor;
            if (nIter >= minIter && CONVERGED(perf.finalResidual))
            {
                for (int c = 0; c < nC; ++c) psi[c] += alpha * yA[c];
                ++nIter;
                perf.nIterations = nIter;
                done = 1;
                break;
            }
            /* zA = M^-1 sA */
            for (int c = 0; c < nC; ++c) zA[c] = rD[c] * sA[c];
            for (int f = 0; f < nIf; ++f) zA[nei[f]] -= rD[nei[f]] * lower[f] * zA[own[f]];
            for (int f = nIf - 1; f >= 0; --f) zA[own[f]] -= rD[own[f]] * upper[f] * zA[nei[f]];
            amul_ldu(nC, nIf, own, nei, diagC, upper, lower, zA, tA);
            omega = sum_prod(nC, tA, sA) / sum_prod(nC, tA, tA);
            for (int c = 0; c < nC; ++c)
            {
                psi[c] += alpha * yA[c] + omega * zA[c];
                rA[c] = sA[c] - omega * tA[c];
            }
            perf.finalResidual = sum_mag(nC, rA) / normFactor;
            ++nIter;
        } while ((nIter < maxIter && !CONVERGED(perf.finalResidual)) || nIter < minIter);
        if (!done) perf.nIterations = nIter;
        free(pA); free(AyA); free(sA); free(zA); free(tA); free(rA0);
    }
#undef CONVERGED
    free(diagC); free(b); free(rD); free(yA); free(rA); free(sumA);
    return perf;
}

/* ------------------------------------------------------------------------------------------------
 * The GAMG of gamg.cu:267-400 (the solver cpu::simpleStep calls for p, simpleFoam_cpp.cu:156).
 * ---------------------------------------------------------------------------------------------- */
typedef struct
{
    int nCells, nFaces;
    int *lowerAddr, *upperAddr, *ownerStart;
    double *diag, *upper, *lower, *faceWeights;
    int* restrictMap;     /* this level's cell -> next-coarser cell (NULL on the coarsest) */
    int nCoarse;
} glevel;

static void glevel_free(glevel* L)
{
    free(L->lowerAddr); free(L->upperAddr); free(L->ownerStart);
    free(L->diag); free(L->upper); free(L->lower); free(L->faceWeights); free(L->restrictMap);
}

static int* build_owner_start(const int* lowerAddr, int nFaces, int nCells)
{
    int* os = IALLOC(nCells + 1);
    for (int f = 0; f < nFaces; ++f) os[lowerAddr[f] + 1]++;
    for (int c = 0; c < nCells; ++c) os[c + 1] += os[c];
    return os;
}

static void gamg_amul(const glevel* L, const double* x, double* Ax)
{
    for (int c = 0; c < L->nCells; ++c) Ax[c] = L->diag[c] * x[c];
    for (int f = 0; f < L->nFaces; ++f)
    {
        Ax[L->upperAddr[f]] += L->lower[f] * x[L->lowerAddr[f]];
        Ax[L->lowerAddr[f]] += L->upper[f] * x[L->upperAddr[f]];
    }
}

/* gamg.cu:41-107 agglomeratePair: each cell's faces, the upper-side ones (faces it neighbours) first in
 * face order, then its lower-side ones; greedy pairing with the heaviest face to an unmatched
 * neighbour (strict >), else joining the heaviest face's cluster, singletons numbered last; reverse
 * sweeps renumber n - 1 - id */
static int* agglomerate_pair(const glevel* L, int forward, int* nCoarseOut)
{
    const int nF = L->nCells;
    int* nNbr = IALLOC(nF);
    for (int f = 0; f < L->nFaces; ++f)
    {
        nNbr[L->upperAddr[f]]++;
        nNbr[L->lowerAddr[f]]++;
    }
    int* off = IALLOC(nF + 1);
    for (int c = 0; c < nF; ++c) off[c + 1] = off[c] + nNbr[c];
    int* cellFaces = IALLOC(off[nF]);
    memset(nNbr, 0, sizeof(int) * (size_t)nF);
    for (int f = 0; f < L->nFaces; ++f) cellFaces[off[L->upperAddr[f]] + nNbr[L->upperAddr[f]]++] = f;
    for (int f = 0; f < L->nFaces; ++f) cellFaces[off[L->lowerAddr[f]] + nNbr[L->lowerAddr[f]]++] = f;
    int* map = IALLOC(nF);
    for (int c = 0; c < nF; ++c) map[c] = -1;
    int nCoarse = 0;
    for (int cf = 0; cf < nF; ++cf)
    {
        const int c = forward ? cf : nF - cf - 1;
        if (map[c] >= 0) continue;
        int matchFace = -1;
        double maxW = -1e300;
        for (int o = off[c]; o < off[c + 1]; ++o)
        {
            const int f = cellFaces[o];
            if (map[L->upperAddr[f]] < 0 && map[L->lowerAddr[f]] < 0 && L->faceWeights[f] > maxW)
            {
                matchFace = f;
                maxW = L->faceWeights[f];
            }
        }
        if (matchFace >= 0)
        {
            map[L->upperAddr[matchFace]] = nCoarse;
            map[L->lowerAddr[matchFace]] = nCoarse;
            ++nCoarse;
        }
        else
        {
            int cm = -1;
            double cMax = -1e300;
            for (int o = off[c]; o < off[c + 1]; ++o)
            {
                const int f = cellFaces[o];
                if (L->faceWeights[f] > cMax)
                {
                    cm = f;
                    cMax = L->faceWeights[f];
                }
            }
            if (cm >= 0)
            {
                const int a = map[L->upperAddr[cm]], b = map[L->lowerAddr[cm]];
                map[c] = a > b ? a : b;
            }
        }
    }
    for (int cf = 0; cf < nF; ++cf)
    {
        const int c = forward ? cf : nF - cf - 1;
        if (map[c] < 0) map[c] = nCoarse++;
    }
    if (!forward)
    {
        --nCoarse;
        for (int c = 0; c < nF; ++c) map[c] = nCoarse - map[c];
        ++nCoarse;
    }
    free(nNbr); free(off); free(cellFaces);
    *nCoarseOut = nCoarse;
    return map;
}

typedef struct { long long key; int face; } keyed;

static int keyed_cmp(const void* a, const void* b)
{
    const keyed* x = (const keyed*)a;
    const keyed* y = (const keyed*)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return x->face < y->face ? -1 : (x->face > y->face);
}

/* gamg.cu:110-172 coarsen: coarse faces are the distinct (min, max) coarse-cell pairs numbered in sorted
 * order; diag summed over the fine cells in cell order, then per fine face in face order either the
 * flip-aware off-diagonals (and weights) onto its coarse face or upper + lower onto the coarse diagonal */
static void coarsen(const glevel* F, glevel* C)
{
    const int* map = F->restrictMap;
    const int nCC = F->nCoarse;
    memset(C, 0, sizeof(*C));
    C->nCells = nCC;
    keyed* ks = (keyed*)zalloc((size_t)F->nFaces, sizeof(keyed));
    int nk = 0;
    for (int f = 0; f < F->nFaces; ++f)
    {
        const int co = map[F->lowerAddr[f]], cn = map[F->upperAddr[f]];
        if (co == cn) continue;
        const int a = co < cn ? co : cn, b = co < cn ? cn : co;
        ks[nk].key = (long long)a * nCC + b;
        ks[nk].face = f;
        ++nk;
    }
    qsort(ks, (size_t)nk, sizeof(keyed), keyed_cmp);
    int* faceRestrict = IALLOC(F->nFaces);
    char* faceFlip = (char*)zalloc((size_t)F->nFaces, 1);
    int nCF = 0;
    for (int i = 0; i < nk; ++i)
        if (i == 0 || ks[i].key != ks[i - 1].key) ++nCF;
    C->nFaces = nCF;
    C->lowerAddr = IALLOC(nCF);
    C->upperAddr = IALLOC(nCF);
    int idx = -1;
    for (int i = 0; i < nk; ++i)
    {
        if (i == 0 || ks[i].key != ks[i - 1].key)
        {
            ++idx;
            C->lowerAddr[idx] = (int)(ks[i].key / nCC);
            C->upperAddr[idx] = (int)(ks[i].key % nCC);
        }
        faceRestrict[ks[i].face] = idx;
    }
    for (int f = 0; f < F->nFaces; ++f)
    {
        const int co = map[F->lowerAddr[f]], cn = map[F->upperAddr[f]];
        if (co == cn)
            faceRestrict[f] = -1 - co;
        else
            faceFlip[f] = (co > cn) ? 1 : 0;
    }
    C->diag = DALLOC(nCC);
    C->upper = DALLOC(nCF);
    C->lower = DALLOC(nCF);
    C->faceWeights = DALLOC(nCF);
    for (int c = 0; c < F->nCells; ++c) C->diag[map[c]] += F->diag[c];
    for (int f = 0; f < F->nFaces; ++f)
    {
        const int cf = faceRestrict[f];
        if (cf >= 0)
        {
            if (!faceFlip[f])
            {
                C->upper[cf] += F->upper[f];
                C->lower[cf] += F->lower[f];
            }
            else
            {
                C->upper[cf] += F->lower[f];
                C->lower[cf] += F->upper[f];
            }
            C->faceWeights[cf] += F->faceWeights[f];
        }
        else
            C->diag[-1 - cf] += F->upper[f] + F->lower[f];
    }
    C->ownerStart = build_owner_start(C->lowerAddr, nCF, nCC);
    free(ks); free(faceRestrict); free(faceFlip);
}

/* gamg.cu:187-205: GaussSeidel with bPrime (owned faces in order, then the lower neighbours' update) */
static void gamg_gauss_seidel(const glevel* L, double* psi, const double* source, int nSweeps, double* bPrime)
{
    for (int s = 0; s < nSweeps; ++s)
    {
        memcpy(bPrime, source, sizeof(double) * (size_t)L->nCells);
        int fEnd = L->ownerStart[0];
        for (int c = 0; c < L->nCells; ++c)
        {
            const int fStart = fEnd;
            fEnd = L->ownerStart[c + 1];
            double psii = bPrime[c];
            for (int fp = fStart; fp < fEnd; ++fp) psii -= L->upper[fp] * psi[L->upperAddr[fp]];
            psii /= L->diag[c];
            for (int fp = fStart; fp < fEnd; ++fp) bPrime[L->upperAddr[fp]] -= L->lower[fp] * psii;
            psi[c] = psii;
        }
    }
}

/* gamg.cu:207-219 */
static void gamg_scale(const glevel* L, double* field, const double* source, double* Acf)
{
    gamg_amul(L, field, Acf);
    double num = 0.0, den = 0.0;
    for (int c = 0; c < L->nCells; ++c)
    {
        num += field[c] * source[c];
        den += field[c] * Acf[c];
    }
    const double sf = num / ((den >= 0 ? fmax(den, 1e-300) : fmin(den, -1e-300)));
    for (int c = 0; c < L->nCells; ++c) field[c] = sf * field[c] + (source[c] - sf * Acf[c]) / L->diag[c];
}

/* gamg.cu:222-263: dense Gaussian elimination with partial pivoting, rows swapped */
static void gamg_dense_solve(const glevel* L, double* x, const double* b)
{
    const int n = L->nCells;
    double* A = DALLOC((size_t)n * n);
    double* rhs = DALLOC(n);
    memcpy(rhs, b, sizeof(double) * (size_t)n);
    for (int c = 0; c < n; ++c) A[(size_t)c * n + c] = L->diag[c];
    for (int f = 0; f < L->nFaces; ++f)
    {
        A[(size_t)L->lowerAddr[f] * n + L->upperAddr[f]] += L->upper[f];
        A[(size_t)L->upperAddr[f] * n + L->lowerAddr[f]] += L->lower[f];
    }
    for (int k = 0; k < n; ++k)
    {
        int piv = k;
        double best = fabs(A[(size_t)k * n + k]);
        for (int i = k + 1; i < n; ++i)
            if (fabs(A[(size_t)i * n + k]) > best)
            {
                best = fabs(A[(size_t)i * n + k]);
                piv = i;
            }
        if (piv != k)
        {
            for (int j = 0; j < n; ++j)
            {
                const double t = A[(size_t)k * n + j];
                A[(size_t)k * n + j] = A[(size_t)piv * n + j];
                A[(size_t)piv * n + j] = t;
            }
            const double t = rhs[k];
            rhs[k] = rhs[piv];
            rhs[piv] = t;
        }
        const double dkk = A[(size_t)k * n + k];
        for (int i = k + 1; i < n; ++i)
        {
            const double fct = A[(size_t)i * n + k] / dkk;
            if (fct == 0.0) continue;
            for (int j = k; j < n; ++j) A[(size_t)i * n + j] -= fct * A[(size_t)k * n + j];
            rhs[i] -= fct * rhs[k];
        }
    }
    for (int i = n - 1; i >= 0; --i)
    {
        double s = rhs[i];
        for (int j = i + 1; j < n; ++j) s -= A[(size_t)i * n + j] * x[j];
        x[i] = s / A[(size_t)i * n + i];
    }
    free(A); free(rhs);
}

orc_perf orc_gamg_ref(int nCells, int nFaces, const int* owner, const int* neighbour, const double* diag,
                      const double* upper, const double* lower, const double* source, int nBFaces,
                      const int* bFaceCell, const double* iC, const double* bC, const double* magSf,
                      double* psi, double tol, double relTol, int maxIter, int minIter, int nCellsInCoarsest,
                      int nPostSweeps, int nFinestSweeps, int* nLevelsOut)
{
    const int nC = nCells;
    int cap = 64, nLev = 1;
    glevel* lv = (glevel*)zalloc((size_t)cap, sizeof(glevel));
    glevel* L0 = &lv[0];
    L0->nCells = nC;
    L0->nFaces = nFaces;
    L0->lowerAddr = IALLOC(nFaces);
    L0->upperAddr = IALLOC(nFaces);
    memcpy(L0->lowerAddr, owner, sizeof(int) * (size_t)nFaces);
    memcpy(L0->upperAddr, neighbour, sizeof(int) * (size_t)nFaces);
    L0->diag = DALLOC(nC);
    L0->upper = DALLOC(nFaces);
    L0->lower = DALLOC(nFaces);
    memcpy(L0->diag, diag, sizeof(double) * (size_t)nC);
    memcpy(L0->upper, upper, sizeof(double) * (size_t)nFaces);
    memcpy(L0->lower, lower ? lower : upper, sizeof(double) * (size_t)nFaces);
    double* b = DALLOC(nC);
    memcpy(b, source, sizeof(double) * (size_t)nC);
    for (int i = 0; i < nBFaces; ++i)
    {
        L0->diag[bFaceCell[i]] += iC[i];
        b[bFaceCell[i]] += bC[i];
    }
    L0->ownerStart = build_owner_start(L0->lowerAddr, nFaces, nC);
    L0->faceWeights = DALLOC(nFaces);
    memcpy(L0->faceWeights, magSf, sizeof(double) * (size_t)nFaces);

    int forward = 1;
    for (int lev = 0; lev < 50; ++lev)
    {
        int nCoarse = 0;
        int* map = agglomerate_pair(&lv[lev], forward, &nCoarse);
        forward = !forward;
        if (nCoarse <= nCellsInCoarsest || nCoarse >= lv[lev].nCells || nCoarse < 2)
        {
            free(map);
            break;
        }
        lv[lev].restrictMap = map;
        lv[lev].nCoarse = nCoarse;
        if (nLev == cap) abort();
        coarsen(&lv[lev], &lv[nLev]);
        ++nLev;
    }
    const int coarsest = nLev - 1;
    if (nLevelsOut) *nLevelsOut = nLev;

    double* Apsi = DALLOC(nC);
    gamg_amul(L0, psi, Apsi);
    double* sumA = DALLOC(nC);
    for (int c = 0; c < nC; ++c) sumA[c] = L0->diag[c];
    for (int f = 0; f < L0->nFaces; ++f)
    {
        sumA[L0->upperAddr[f]] += L0->lower[f];
        sumA[L0->lowerAddr[f]] += L0->upper[f];
    }
    double xRef = 0.0;
    for (int c = 0; c < nC; ++c) xRef += psi[c];
    xRef /= nC;
    double normFactor = 0.0;
    for (int c = 0; c < nC; ++c)
    {
        const double t = sumA[c] * xRef;
        normFactor += fabs(Apsi[c] - t) + fabs(b[c] - t);
    }
    normFactor += 1e-20;
    double* finestResidual = DALLOC(nC);
    for (int c = 0; c < nC; ++c) finestResidual[c] = b[c] - Apsi[c];

    orc_perf perf;
    perf.initialResidual = sum_mag(nC, finestResidual) / normFactor;
    perf.finalResidual = perf.initialResidual;
    perf.nIterations = 0;
#define CONVERGED(fr) (((fr) < tol) || (relTol > 0.0 && (fr) < relTol * perf.initialResidual))
    if ((minIter > 0 || !CONVERGED(perf.finalResidual)) && coarsest >= 1)
    {
        double** cSource = (double**)zalloc((size_t)nLev, sizeof(double*));
        double** cCorr = (double**)zalloc((size_t)nLev, sizeof(double*));
        for (int k = 1; k < nLev; ++k)
        {
            cSource[k] = DALLOC(lv[k].nCells);
            cCorr[k] = DALLOC(lv[k].nCells);
        }
        double* work = DALLOC(nC);
        double* bPrime = DALLOC(nC);
        double* finestCorr = DALLOC(nC);
        int nIter = 0;
        do
        {
            /* restrict the residual straight down, coarse += fine in fine-cell order */
            memset(cSource[1], 0, sizeof(double) * (size_t)lv[1].nCells);
            for (int c = 0; c < nC; ++c) cSource[1][lv[0].restrictMap[c]] += finestResidual[c];
            for (int k = 1; k < coarsest; ++k)
            {
                memset(cSource[k + 1], 0, sizeof(double) * (size_t)lv[k + 1].nCells);
                for (int c = 0; c < lv[k].nCells; ++c) cSource[k + 1][lv[k].restrictMap[c]] += cSource[k][c];
            }
            gamg_dense_solve(&lv[coarsest], cCorr[coarsest], cSource[coarsest]);
            for (int k = coarsest - 1; k >= 1; --k)
            {
                for (int c = 0; c < lv[k].nCells; ++c) cCorr[k][c] = cCorr[k + 1][lv[k].restrictMap[c]];
                gamg_scale(&lv[k], cCorr[k], cSource[k], work);
                gamg_gauss_seidel(&lv[k], cCorr[k], cSource[k], nPostSweeps, bPrime);
            }
            for (int c = 0; c < nC; ++c) finestCorr[c] = cCorr[1][lv[0].restrictMap[c]];
            gamg_scale(L0, finestCorr, finestResidual, work);
            for (int c = 0; c < nC; ++c) psi[c] += finestCorr[c];
            gamg_gauss_seidel(L0, psi, b, nFinestSweeps, bPrime);
            gamg_amul(L0, psi, Apsi);
            for (int c = 0; c < nC; ++c) finestResidual[c] = b[c] - Apsi[c];
            perf.finalResidual = sum_mag(nC, finestResidual) / normFactor;
            ++nIter;
        } while ((nIter < maxIter && !CONVERGED(perf.finalResidual)) || nIter < minIter);
        perf.nIterations = nIter;
        for (int k = 1; k < nLev; ++k)
        {
            free(cSource[k]);
            free(cCorr[k]);
        }
        free(cSource); free(cCorr); free(work); free(bPrime); free(finestCorr);
    }
#undef CONVERGED
    for (int k = 0; k < nLev; ++k) glevel_free(&lv[k]);
    free(lv); free(b); free(Apsi); free(sumA); free(finestResidual);
    return perf;
}

/* one agglomeratePair pass of gamg.cu:41-107 on an LDU addressing with face weights (the restatement's
 * pass, exposed so the tests can hold it against the reference's compiled faceAreaPair pass and against
 * its invariants) */
int orc_gamg_ref_pair(int nCells, int nFaces, const int* lowerAddr, const int* upperAddr, const double* weights,
                      int forward, int* map)
{
    glevel L;
    memset(&L, 0, sizeof(L));
    L.nCells = nCells;
    L.nFaces = nFaces;
    L.lowerAddr = (int*)lowerAddr;
    L.upperAddr = (int*)upperAddr;
    L.faceWeights = (double*)weights;
    int n = 0;
    int* m = agglomerate_pair(&L, forward, &n);
    memcpy(map, m, sizeof(int) * (size_t)nCells);
    free(m);
    return n;
}

/* ------------------------------------------------------------------------------------------------
 * One SIMPLE iteration: simpleFoam_cpp.cu:15-237 for the laminar (or frozen-nuEff) path with uncoupled
 * fixedValue/zeroGradient/empty patches, no MRF/fvOptions/SIMPLEC.
 * ---------------------------------------------------------------------------------------------- */

/* patch values after evaluate(): fixedValue keeps its value, zeroGradient takes the owner cell's, empty 0 */
static void evaluate(int nB, const int* bfc, const int* type, const double* value, int stride, int k,
                     const double* cell, double* out)
{
    for (int i = 0; i < nB; ++i)
    {
        if (type[i] == ORC_BC_FIXED_VALUE)
            out[i] = value[(size_t)stride * i + k];
        else if (type[i] == ORC_BC_ZERO_GRADIENT)
            out[i] = cell[bfc[i]];
        else
            out[i] = 0.0;
    }
}

/* the Gauss gradient of a field from its current boundary values (every non-empty face takes its
 * patch value): fvc.cu:440-504 per component */
static void grad_of(const orc_simple_case* cs, const double* psi, const double* psiB, const int* gType,
                    double* grad)
{
    orc_grad_gauss(cs->nCells, cs->nFaces, cs->owner, cs->neighbour, cs->nBFaces, cs->bFaceCell, gType, psiB,
                   cs->Sf, cs->weights, cs->V, psi, grad);
}

/* fvc::gradUBoundary (fvc.cu:781-865) for uncoupled fixedValue/zeroGradient/empty patches:
 * gb = gc + n (sn - n.gc), n = (1/magSf) Sf, sn = deltaCoeffs*(Ub - Uc) on a fixedValue patch, 0 on a
 * zeroGradient one; zero on empty faces. gradU and gb component-major, gradU[(i*3 + j)*n + c] = dU_j/dx_i. */
void orc_grad_u_boundary(int nCells, int nFaces, int nBFaces, const int* bFaceCell, const int* uType,
                         const double* Sf, const double* magSf, const double* bDeltaCoeffs, const double* Ux,
                         const double* Uy, const double* Uz, const double* Ubx, const double* Uby,
                         const double* Ubz, const double* gradU, double* gb)
{
    const size_t n = (size_t)nCells, nB = (size_t)nBFaces;
    const double* U[3] = {Ux, Uy, Uz};
    const double* Ub[3] = {Ubx, Uby, Ubz};
    for (int i = 0; i < nBFaces; ++i)
    {
        if (uType[i] == ORC_BC_EMPTY)
        {
            for (int q = 0; q < 9; ++q) gb[q * nB + i] = 0.0;
            continue;
        }
        const int c = bFaceCell[i];
        const size_t gf = (size_t)nFaces + i;
        const double r = 1.0 / magSf[gf];
        const double nv[3] = {r * Sf[3 * gf], r * Sf[3 * gf + 1], r * Sf[3 * gf + 2]};
        double gc[9];
        for (int q = 0; q < 9; ++q) gc[q] = gradU[q * n + c];
        double sn[3];
        for (int j = 0; j < 3; ++j)
            sn[j] = uType[i] == ORC_BC_FIXED_VALUE ? bDeltaCoeffs[i] * (Ub[j][i] - U[j][c]) : 0.0;
        double ndg[3];
        for (int j = 0; j < 3; ++j) ndg[j] = nv[0] * gc[0 * 3 + j] + nv[1] * gc[1 * 3 + j] + nv[2] * gc[2 * 3 + j];
        for (int a = 0; a < 3; ++a)
            for (int j = 0; j < 3; ++j) gb[(a * 3 + j) * nB + i] = gc[a * 3 + j] + nv[a] * (sn[j] - ndg[j]);
    }
}

int orc_simple_step(const orc_simple_case* cs, orc_simple_fields* fl, orc_simple_result* res)
{
    const int nC = cs->nCells, nF = cs->nFaces, nB = cs->nBFaces;
    const int* own = cs->owner;
    const int* nei = cs->neighbour;
    const int* bfc = cs->bFaceCell;
    const double* V = cs->V;
    const double* w = cs->weights;
    const double* bMagSf = cs->magSf + nF;
    double* U[3] = {fl->Ux, fl->Uy, fl->Uz};
    double* Ub[3] = {fl->Ubx, fl->Uby, fl->Ubz};
    memset(res, 0, sizeof(*res));

    /* the gradient's boundary: every non-empty face takes its patch value */
    int* gType = IALLOC(nB);
    int* isEmpty = IALLOC(nB);
    for (int i = 0; i < nB; ++i)
    {
        gType[i] = cs->uType[i] == ORC_BC_EMPTY ? ORC_BC_EMPTY : ORC_BC_FIXED_VALUE;
        isEmpty[i] = cs->uType[i] == ORC_BC_EMPTY;
    }
    int* pgType = IALLOC(nB);
    for (int i = 0; i < nB; ++i) pgType[i] = cs->pType[i] == ORC_BC_EMPTY ? ORC_BC_EMPTY : ORC_BC_FIXED_VALUE;

    /* nuEff: laminar nu everywhere, or the frozen turbulent field (simpleFoam_cpp.cu:27-46) */
    double* nuC = DALLOC(nC);
    double* nuB = DALLOC(nB);
    for (int c = 0; c < nC; ++c) nuC[c] = cs->nuEff ? cs->nuEff[c] : cs->nu;
    for (int i = 0; i < nB; ++i) nuB[i] = cs->nuEffB ? cs->nuEffB[i] : cs->nu;

    double* pPrev = DALLOC(nC);
    memcpy(pPrev, fl->p, sizeof(double) * (size_t)nC);
    /* p's patch values at the top of the iteration (simpleFoam_cpp.cu:54-69) */
    evaluate(nB, bfc, cs->pType, cs->pValue, 1, 0, fl->p, fl->pb);

    /* ---- UEqn (UEqn_cpp.cu:156-264) ---- */
    double* D = DALLOC(nC);
    double* Up = DALLOC(nF);
    double* Lo = DALLOC(nF);
    orc_div_upwind(nC, nF, own, nei, fl->phi, D, Up, Lo);
    double* iC[3];
    double* bC[3];
    double* src[3];
    for (int k = 0; k < 3; ++k)
    {
        iC[k] = DALLOC(nB);
        bC[k] = DALLOC(nB);
        src[k] = DALLOC(nC);
        double* val = DALLOC(nB);
        for (int i = 0; i < nB; ++i) val[i] = cs->uValue[3 * (size_t)i + k];
        orc_div_bcoeffs(nB, cs->uType, val, fl->phiB, iC[k], bC[k]);
        free(val);
    }
    double* gradU = DALLOC(9 * (size_t)nC);   /* component-major, (i*3 + j): dU_j/dx_i */
    double* g3 = DALLOC(3 * (size_t)nC);
    for (int k = 0; k < 3; ++k)
    {
        grad_of(cs, U[k], Ub[k], gType, g3);
        for (int c = 0; c < nC; ++c)
            for (int i = 0; i < 3; ++i) gradU[(size_t)(i * 3 + k) * nC + c] = g3[3 * (size_t)c + i];
    }
    if (cs->linearUpwind)
    {
        /* linearUpwind's deferred correction, subtracted from the source (UEqn_cpp.cu:194-209) */
        double* corr = DALLOC(nC);
        for (int k = 0; k < 3; ++k)
        {
            for (int c = 0; c < nC; ++c)
                for (int i = 0; i < 3; ++i) g3[3 * (size_t)c + i] = gradU[(size_t)(i * 3 + k) * nC + c];
            orc_linear_upwind_corr(nC, nF, own, nei, fl->phi, cs->Cf, cs->C, g3, corr);
            for (int c = 0; c < nC; ++c) src[k][c] -= corr[c];
        }
        free(corr);
    }
    /* - fvm::laplacian(nuEff, U): face nuEff interpolated, boundary nuEff the patch's own
     * (linearViscousStress_cpp.cu:12-33, 105-137) */
    const int nonOrthDc = cs->correctedLaplacian;
    const double* dc = nonOrthDc ? cs->nonOrthDeltaCoeffs : cs->deltaCoeffs;
    double* nuF = DALLOC(nF);
    for (int f = 0; f < nF; ++f) nuF[f] = w[f] * nuC[own[f]] + (1.0 - w[f]) * nuC[nei[f]];
    {
        double* lD = DALLOC(nC);
        double* lU = DALLOC(nF);
        orc_laplacian(nC, nF, own, nei, dc, nuF, cs->magSf, lD, lU, NULL);
        for (int c = 0; c < nC; ++c) D[c] += -1.0 * lD[c];
        for (int f = 0; f < nF; ++f)
        {
            Up[f] += -1.0 * lU[f];
            Lo[f] += -1.0 * lU[f];
        }
        double* liC = DALLOC(nB);
        double* lbC = DALLOC(nB);
        double* val = DALLOC(nB);
        for (int k = 0; k < 3; ++k)
        {
            for (int i = 0; i < nB; ++i) val[i] = cs->uValue[3 * (size_t)i + k];
            orc_laplacian_bcoeffs(nB, cs->uType, val, nuB, bMagSf, cs->bDeltaCoeffs, liC, lbC);
            for (int i = 0; i < nB; ++i)
            {
                iC[k][i] += -1.0 * liC[i];
                bC[k][i] += -1.0 * lbC[i];
            }
        }
        free(lD); free(lU); free(liC); free(lbC); free(val);
    }
    if (cs->correctedLaplacian)
    {
        /* its explicit correction: source += V*div(ffc) per component (linearViscousStress_cpp.cu:138-161) */
        double* ffc = DALLOC(nF);
        for (int k = 0; k < 3; ++k)
        {
            for (int c = 0; c < nC; ++c)
                for (int i = 0; i < 3; ++i) g3[3 * (size_t)c + i] = gradU[(size_t)(i * 3 + k) * nC + c];
            orc_laplacian_corr(nC, nF, own, nei, g3, w, cs->corrVec, nuF, cs->magSf, ffc, NULL);
            double* s = DALLOC(nC);
            for (int f = 0; f < nF; ++f)
            {
                s[own[f]] += ffc[f];
                s[nei[f]] -= ffc[f];
            }
            for (int c = 0; c < nC; ++c) src[k][c] += s[c];
            free(s);
        }
        free(ffc);
    }
    {
        /* the explicit stress: source -= (-div(nuEff dev2(T(grad U))))*V (linearViscousStress_cpp.cu:163-184) */
        double* sig = DALLOC(9 * (size_t)nC);
        double* gb = DALLOC(9 * (size_t)nB);
        double* sigB = DALLOC(9 * (size_t)nB);
        orc_scaled_dev2t(nC, gradU, nuC, sig);
        orc_grad_u_boundary(nC, nF, nB, bfc, cs->uType, cs->Sf, cs->magSf, cs->bDeltaCoeffs, U[0], U[1], U[2],
                            Ub[0], Ub[1], Ub[2], gradU, gb);
        orc_scaled_dev2t(nB, gb, nuB, sigB);
        double* d[3] = {DALLOC(nC), DALLOC(nC), DALLOC(nC)};
        orc_tensor_div(nC, nF, own, nei, w, cs->Sf, nB, bfc, isEmpty, sig, sigB, d[0], d[1], d[2]);
        for (int k = 0; k < 3; ++k)
        {
            for (int c = 0; c < nC; ++c)
            {
                const double e = -(d[k][c] / V[c]);
                src[k][c] -= e * V[c];
            }
            free(d[k]);
        }
        free(sig); free(gb); free(sigB);
    }
    /* UEqn.relax() (fv_matrix_ops.cuh:161-209) */
    if (cs->relaxU > 0.0 && cs->relaxU < 1.0)
    {
        double* mag = DALLOC(nB);
        double* mn = DALLOC(nB);
        orc_cmpt_mag_max_min(nB, iC[0], iC[1], iC[2], mag, mn);
        orc_relax_matrix(nC, nF, own, nei, Up, Lo, nB, bfc, mag, mn, NULL, cs->relaxU, D, 3,
                         (const double* const*)U, src, NULL);
        free(mag); free(mn);
    }

    /* ---- momentum predictor: solve(UEqn == -grad(p)) on a copy (simpleFoam_cpp.cu:94-127) ---- */
    double* gradP = DALLOC(3 * (size_t)nC);
    if (cs->momentumPredictor)
    {
        grad_of(cs, fl->p, fl->pb, pgType, gradP);
        double* s = DALLOC(nC);
        for (int k = 0; k < 3; ++k)
        {
            for (int c = 0; c < nC; ++c) s[c] = src[k][c] - gradP[3 * (size_t)c + k] * V[c];
            const orc_perf pf = orc_pbicgstab(nC, nF, own, nei, D, Up, Lo, s, nB, bfc, iC[k], bC[k], U[k],
                                              cs->tolU, cs->relTolU, cs->maxIterU, cs->minIterU);
            res->UInitial[k] = pf.initialResidual;
            res->UFinal[k] = pf.finalResidual;
            res->UIters[k] = pf.nIterations;
        }
        free(s);
        for (int k = 0; k < 3; ++k) evaluate(nB, bfc, cs->uType, cs->uValue, 3, k, U[k], Ub[k]);
    }

    /* ---- pressurePredictor (pEqn_cpp.cu:31-198) ---- */
    double* rAU = DALLOC(nC);
    {
        double* Dm = DALLOC(nC);
        orc_fold_boundary_diag(nC, nB, bfc, iC[0], iC[1], iC[2], D, Dm);
        orc_reciprocal_v(nC, Dm, V, rAU);
        free(Dm);
    }
    double* HbyA[3];
    double* HbyAb[3];
    {
        double* bd = DALLOC(nB);
        for (int k = 0; k < 3; ++k)
        {
            HbyA[k] = DALLOC(nC);
            HbyAb[k] = DALLOC(nB);
            orc_boundary_diag_term(nB, iC[0], iC[1], iC[2], k, bd);
            orc_matrix_h(nC, nF, own, nei, Up, Lo, U[k], src[k], nB, bfc, bd, bC[k], V, HbyA[k]);
            for (int c = 0; c < nC; ++c) HbyA[k][c] = rAU[c] * HbyA[k][c];
            /* constrainHbyA: U's value where U is not assignable (fixedValue), else the owner cell's */
            for (int i = 0; i < nB; ++i)
                HbyAb[k][i] = cs->uType[i] == ORC_BC_FIXED_VALUE ? Ub[k][i] : HbyA[k][bfc[i]];
        }
        free(bd);
    }
    double* phiHbyA = DALLOC(nF);
    double* phiHbyAB = DALLOC(nB);
    orc_vector_flux(nF, own, nei, w, cs->Sf, HbyA[0], HbyA[1], HbyA[2], phiHbyA);
    orc_boundary_flux(nB, cs->Sf + 3 * (size_t)nF, HbyAb[0], HbyAb[1], HbyAb[2], isEmpty, phiHbyAB);
    res->adjustStatus = 0;
    if (cs->pRefCell >= 0)
    {
        int* adjustable = IALLOC(nB);
        for (int i = 0; i < nB; ++i) adjustable[i] = cs->uType[i] != ORC_BC_FIXED_VALUE;
        res->adjustStatus = orc_adjust_phi(nF, phiHbyA, nB, adjustable, phiHbyAB, &res->massCorr, NULL);
        free(adjustable);
        if (res->adjustStatus) goto cleanup;
    }

    /* ---- the pressure equation, nNonOrthogonalCorrectors + 1 times (simpleFoam_cpp.cu:148-174) ---- */
    {
        double* rAUf = DALLOC(nF);
        double* rAUb = DALLOC(nB);
        for (int f = 0; f < nF; ++f) rAUf[f] = w[f] * rAU[own[f]] + (1.0 - w[f]) * rAU[nei[f]];
        for (int i = 0; i < nB; ++i) rAUb[i] = rAU[bfc[i]];
        double* pD = DALLOC(nC);
        double* pU = DALLOC(nF);
        double* piC = DALLOC(nB);
        double* pbC = DALLOC(nB);
        double* psrc = DALLOC(nC);
        double* pDr = DALLOC(nC);
        double* ffc = DALLOC(nF);
        double* divP = DALLOC(nC);
        /* fvc::div(phiHbyA) (fvc.cu:697-733): internal faces, then non-empty boundary faces, over V */
        for (int f = 0; f < nF; ++f)
        {
            divP[own[f]] += phiHbyA[f];
            divP[nei[f]] -= phiHbyA[f];
        }
        for (int i = 0; i < nB; ++i)
            if (!isEmpty[i]) divP[bfc[i]] += phiHbyAB[i];
        for (int c = 0; c < nC; ++c) divP[c] /= V[c];
        for (int corr = 0; corr <= cs->nNonOrthCorr; ++corr)
        {
            orc_laplacian(nC, nF, own, nei, dc, rAUf, cs->magSf, pD, pU, NULL);
            orc_laplacian_bcoeffs(nB, cs->pType, cs->pValue, rAUb, bMagSf, cs->bDeltaCoeffs, piC, pbC);
            memset(psrc, 0, sizeof(double) * (size_t)nC);
            if (cs->correctedLaplacian)
            {
                grad_of(cs, fl->p, fl->pb, pgType, gradP);
                orc_laplacian_corr(nC, nF, own, nei, gradP, w, cs->corrVec, rAUf, cs->magSf, ffc, psrc);
            }
            for (int c = 0; c < nC; ++c) psrc[c] += divP[c] * V[c];
            memcpy(pDr, pD, sizeof(double) * (size_t)nC);
            orc_set_reference(pDr, psrc, cs->pRefCell, cs->pRefValue);
            orc_perf pf;
            if (cs->pSolver == ORC_P_GAMG_REF)
                pf = orc_gamg_ref(nC, nF, own, nei, pDr, pU, pU, psrc, nB, bfc, piC, pbC, cs->magSf, fl->p, cs->tolP,
                                  cs->relTolP, cs->maxIterP, cs->minIterP, 10, 2, 2, NULL);
            else
            {
                /* the folded system for the hierarchy's solvers */
                double* diagC = DALLOC(nC);
                double* b = DALLOC(nC);
                orc_fold(nC, pDr, psrc, nB, bfc, cs->pType, piC, pbC, diagC, b);
                if (cs->pSolver == ORC_P_BRAE_AMGPCG)
                    pf = orc_brae_amgpcg(cs->pAggl, diagC, pU, b, fl->p, cs->tolP, cs->relTolP, cs->maxIterP,
                                         cs->minIterP, cs->amgPre, cs->amgPost, cs->amgFlexible, cs->amgSingle,
                                         cs->amgScale, 0.0, NULL);
                else
                {
                    orc_gamg_controls gc;
                    orc_gamg_default_controls(&gc);
                    gc.smoother = 0;
                    pf = orc_pcg_gamg_solve(cs->pAggl, diagC, pU, b, fl->p, cs->tolP, cs->relTolP, cs->maxIterP,
                                            cs->minIterP, &gc);
                }
                free(diagC);
                free(b);
            }
            if (corr == 0) res->pInitial = pf.initialResidual;
            res->pFinal = pf.finalResidual;
            if (corr < 8) res->pIters[corr] = pf.nIterations;
            evaluate(nB, bfc, cs->pType, cs->pValue, 1, 0, fl->p, fl->pb);
            if (corr == cs->nNonOrthCorr)
            {
                /* phi = phiHbyA - pEqn.flux() (pEqn_cpp.cu:261-275, fv_matrix_ops.cuh:33-80) */
                for (int f = 0; f < nF; ++f)
                {
                    double flux = pU[f] * fl->p[nei[f]] - pU[f] * fl->p[own[f]];
                    if (cs->correctedLaplacian) flux += ffc[f];
                    fl->phi[f] = phiHbyA[f] - flux;
                }
                for (int i = 0; i < nB; ++i)
                    fl->phiB[i] = phiHbyAB[i] - (piC[i] * fl->p[bfc[i]] - pbC[i]);
            }
        }
        free(rAUf); free(rAUb); free(pD); free(pU); free(piC); free(pbC); free(psrc); free(pDr); free(ffc);
        free(divP);
    }

    /* p.relax(), then U = HbyA - rAU*grad(p) with the relaxed p (simpleFoam_cpp.cu:176-184) */
    orc_relax_field(nC, fl->p, pPrev, cs->relaxP);
    evaluate(nB, bfc, cs->pType, cs->pValue, 1, 0, fl->p, fl->pb);
    grad_of(cs, fl->p, fl->pb, pgType, gradP);
    for (int k = 0; k < 3; ++k)
    {
        for (int c = 0; c < nC; ++c) U[k][c] = HbyA[k][c] - rAU[c] * gradP[3 * (size_t)c + k];
        evaluate(nB, bfc, cs->uType, cs->uValue, 3, k, U[k], Ub[k]);
    }

cleanup:
    for (int k = 0; k < 3; ++k)
    {
        free(iC[k]); free(bC[k]); free(src[k]); free(HbyA[k]); free(HbyAb[k]);
    }
    free(gType); free(isEmpty); free(pgType); free(nuC); free(nuB); free(pPrev); free(D); free(Up); free(Lo);
    free(gradU); free(g3); free(nuF); free(gradP); free(rAU); free(phiHbyA); free(phiHbyAB);
    return res->adjustStatus;
}

/* ---- the GAMG's parts, exposed so the tests can hold each against what the mathematics fixes ---- */

/* coarsen one level (gamg.cu:110-172): fine addressing, values and restriction map in; the coarse level's
 * face count returned and, when the arrays are given (sized by a first call with them NULL), its
 * addressing (sorted (min, max) pairs) and Galerkin values written */
int orc_gamg_ref_coarsen(int nCells, int nFaces, const int* lowerAddr, const int* upperAddr, const double* diag,
                         const double* upper, const double* lower, const double* weights, const int* map,
                         int nCoarse, int* cLower, int* cUpper, double* cDiag, double* cUpperV, double* cLowerV,
                         double* cWeights)
{
    glevel F, C;
    memset(&F, 0, sizeof(F));
    F.nCells = nCells;
    F.nFaces = nFaces;
    F.lowerAddr = (int*)lowerAddr;
    F.upperAddr = (int*)upperAddr;
    F.diag = (double*)diag;
    F.upper = (double*)upper;
    F.lower = (double*)lower;
    F.faceWeights = (double*)weights;
    F.restrictMap = (int*)map;
    F.nCoarse = nCoarse;
    coarsen(&F, &C);
    const int nCF = C.nFaces;
    if (cLower)
    {
        memcpy(cLower, C.lowerAddr, sizeof(int) * (size_t)nCF);
        memcpy(cUpper, C.upperAddr, sizeof(int) * (size_t)nCF);
        memcpy(cDiag, C.diag, sizeof(double) * (size_t)nCoarse);
        memcpy(cUpperV, C.upper, sizeof(double) * (size_t)nCF);
        memcpy(cLowerV, C.lower, sizeof(double) * (size_t)nCF);
        memcpy(cWeights, C.faceWeights, sizeof(double) * (size_t)nCF);
    }
    glevel_free(&C);
    return nCF;
}

/* nSweeps of gamg.cu's GaussSeidel on an asymmetric LDU level (faces in upper-triangular order) */
void orc_gamg_ref_gauss_seidel(int nCells, int nFaces, const int* lowerAddr, const int* upperAddr,
                               const double* diag, const double* upper, const double* lower, const double* source,
                               double* psi, int nSweeps)
{
    glevel L;
    memset(&L, 0, sizeof(L));
    L.nCells = nCells;
    L.nFaces = nFaces;
    L.lowerAddr = (int*)lowerAddr;
    L.upperAddr = (int*)upperAddr;
    L.diag = (double*)diag;
    L.upper = (double*)upper;
    L.lower = (double*)lower;
    L.ownerStart = build_owner_start(lowerAddr, nFaces, nCells);
    double* bPrime = DALLOC(nCells);
    gamg_gauss_seidel(&L, psi, source, nSweeps, bPrime);
    free(bPrime);
    free(L.ownerStart);
}

/* gamg.cu's dense coarsest solve and its correction scaling */
void orc_gamg_ref_dense(int nCells, int nFaces, const int* lowerAddr, const int* upperAddr, const double* diag,
                        const double* upper, const double* lower, const double* b, double* x)
{
    glevel L;
    memset(&L, 0, sizeof(L));
    L.nCells = nCells;
    L.nFaces = nFaces;
    L.lowerAddr = (int*)lowerAddr;
    L.upperAddr = (int*)upperAddr;
    L.diag = (double*)diag;
    L.upper = (double*)upper;
    L.lower = (double*)lower;
    gamg_dense_solve(&L, x, b);
}

void orc_gamg_ref_scale(int nCells, int nFaces, const int* lowerAddr, const int* upperAddr, const double* diag,
                        const double* upper, const double* lower, const double* source, double* field)
{
    glevel L;
    memset(&L, 0, sizeof(L));
    L.nCells = nCells;
    L.nFaces = nFaces;
    L.lowerAddr = (int*)lowerAddr;
    L.upperAddr = (int*)upperAddr;
    L.diag = (double*)diag;
    L.upper = (double*)upper;
    L.lower = (double*)lower;
    double* Acf = DALLOC(nCells);
    gamg_scale(&L, field, source, Acf);
    free(Acf);
}
