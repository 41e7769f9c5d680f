// OpenFOAM's GAMGSolver on the device: the V-cycle, its smoothers, its correction scaling and its
// coarsest-level PCG-DIC, as a solver and as the PCG's preconditioner
// (src/matrices/lduMatrix/solvers/GAMG/gamg_solver_cpp.cu:627-735 vCycle, :979-1023 gamgSolveFolded,
// :1027-1143 pcgGamgSolve).
//
// The cycle runs on the hierarchy's own grids: the coarse matrices are the FP64 Galerkin sums, each
// coarse grid gets OpenFOAM's LDU addressing (owner < neighbour, faces by owner), and the smoothers
// are the level-scheduled sweeps, which equal the host loops to the bit. Restriction sums each coarse
// cell's fine cells in ascending order (restrictField); prolongation injects (prolongField). With the
// context's sequential reductions every sum runs in the host loop's order, so a solve is the host
// solve to the bit; with the tree reductions it differs only in the rounding of those sums.
//
// The coarsest level (tens of cells: it stops at nCellsInCoarsestLevel) is solved by ONE workgroup
// that holds the level in LDS and runs the whole PCG-DIC loop: per-cell work across the lanes, the
// sums and the two substitutions by lane 0 in the host's order. The reference brings this level to
// the host instead; here nothing leaves the device. A coarsest level too big for LDS runs the
// multi-launch guarded PCG-DIC.
#include "amg.hpp"

#include <algorithm>
#include <cmath>

namespace mfv {

namespace {

constexpr int kTpb = 256;
inline int blocks(int n) { return (n + kTpb - 1) / kTpb; }

constexpr double kVSmall = 1e-300;   // VSMALL
constexpr double kSmall = 1e-20;     // SMALL
constexpr double kGreat = 1e20;      // GREAT
constexpr int kCoarsestMaxIter = 1000;
constexpr int kCoarsestLanes = 64;
constexpr size_t kCoarsestLds = 60 * 1024;

// device scalars of a solve
enum : int { S_NF = 0, S_SUMMAG, S_RHO, S_RHOOLD, S_PQ, S_NUM, S_DEN, S_CINIT, S_CFINAL, S_CITER, S_COUNT = 16 };

// restrictField: coarse cell = its fine cells' values summed in ascending fine index
__global__ __launch_bounds__(kTpb) void restrictK(int nc, const int* __restrict__ cellStart,
                                                  const int* __restrict__ cellList, const double* __restrict__ ff,
                                                  double* __restrict__ cf)
{
    const int C = blockIdx.x * kTpb + threadIdx.x;
    if (C >= nc) return;
    double s = 0.0;
    for (int k = cellStart[C]; k < cellStart[C + 1]; ++k) s += ff[cellList[k]];
    cf[C] = s;
}

// prolongField: fine cell = its coarse cell's value
__global__ __launch_bounds__(kTpb) void prolongK(int nf, const int* __restrict__ map, const double* __restrict__ cf,
                                                 double* __restrict__ ff)
{
    const int i = blockIdx.x * kTpb + threadIdx.x;
    if (i < nf) ff[i] = cf[map[i]];
}

// GAMGSolver::scale's update (gamg_solver_cpp.cu:386-407): sf = (x.b)/(x.Ax +- VSMALL),
// x = sf*x + (b - sf*Ax)/D
__global__ __launch_bounds__(kTpb) void scaleApplyK(int n, double* __restrict__ field, const double* __restrict__ Acf,
                                                    const double* __restrict__ source,
                                                    const double* __restrict__ diag, const double* __restrict__ st)
{
    const int i = blockIdx.x * kTpb + threadIdx.x;
    if (i >= n) return;
    const double num = st[S_NUM], den = st[S_DEN];
    const double sf = num / (den >= 0 ? den + kVSmall : den - kVSmall);
    field[i] = sf * field[i] + (source[i] - sf * Acf[i]) / diag[i];
}

__global__ __launch_bounds__(kTpb) void fillK(int n, double v, double* __restrict__ x)
{
    const int i = blockIdx.x * kTpb + threadIdx.x;
    if (i < n) x[i] = v;
}

__global__ __launch_bounds__(kTpb) void addK(int n, const double* __restrict__ a, double* __restrict__ y)
{
    const int i = blockIdx.x * kTpb + threadIdx.x;
    if (i < n) y[i] += a[i];
}

__global__ __launch_bounds__(kTpb) void subK(int n, const double* __restrict__ b, const double* __restrict__ ax,
                                             double* __restrict__ r)
{
    const int i = blockIdx.x * kTpb + threadIdx.x;
    if (i < n) r[i] = b[i] - ax[i];
}

__global__ void shiftRhoK(double* st)
{
    if (threadIdx.x || blockIdx.x) return;
    st[S_RHOOLD] = st[S_RHO];
}

// PCG direction (PCG.C): p = w on the first iteration, else p = w + (rho/rhoOld) p
__global__ __launch_bounds__(kTpb) void pcgDirK(int n, int first, const double* __restrict__ w, double* __restrict__ p,
                                                const double* __restrict__ st)
{
    const int i = blockIdx.x * kTpb + threadIdx.x;
    if (i >= n) return;
    if (first)
        p[i] = w[i];
    else
    {
        const double beta = st[S_RHO] / st[S_RHOOLD];
        p[i] = w[i] + beta * p[i];
    }
}

// psi += alpha p, r -= alpha q with alpha = rho / (p.q); nothing when |p.q|/normFactor < VSMALL
// (the singularity test the host then reads)
__global__ __launch_bounds__(kTpb) void pcgUpdK(int n, const double* __restrict__ p, const double* __restrict__ q,
                                                double* __restrict__ psi, double* __restrict__ r,
                                                const double* __restrict__ st)
{
    const int i = blockIdx.x * kTpb + threadIdx.x;
    if (i >= n) return;
    const double pq = st[S_PQ];
    if (fabs(pq) / st[S_NF] < kVSmall) return;
    const double alpha = st[S_RHO] / pq;
    psi[i] += alpha * p[i];
    r[i] -= alpha * q[i];
}

// ---- the coarsest level: PCG with DIC from zero in one workgroup (gamg_solver_cpp.cu:321-383) ----
struct CoarseLdu
{
    int n, nF;
    const int *own, *nei, *ownerStart, *losortStart, *losort, *losortOwner;
    const double *diag, *upper;
};

size_t coarseLdsBytes(int n, int nF)
{
    return sizeof(double) * (7 * (size_t)n + nF + 8) + sizeof(int) * (4 * (size_t)nF + 2 * ((size_t)n + 1));
}

__global__ __launch_bounds__(kCoarsestLanes) void coarsestPcgDicK(CoarseLdu A, const double* __restrict__ src,
                                                                  double* __restrict__ psiOut, double tol,
                                                                  double relTol, double* __restrict__ st)
{
    extern __shared__ double lds[];
    const int n = A.n, nF = A.nF, t = threadIdx.x;
    double* dg = lds;
    double* up = dg + n;
    double* rD = up + nF;
    double* pA = rD + n;
    double* wA = pA + n;
    double* rA = wA + n;
    double* psi = rA + n;
    double* b = psi + n;
    double* sc = b + n;   // 8 scalars
    int* own = reinterpret_cast<int*>(sc + 8);
    int* nei = own + nF;
    int* os = nei + nF;
    int* ls = os + n + 1;
    int* lo = ls + n + 1;
    int* lOwn = lo + nF;

    for (int i = t; i < n; i += kCoarsestLanes)
    {
        dg[i] = A.diag[i];
        b[i] = src[i];
        psi[i] = 0.0;
    }
    for (int i = t; i < nF; i += kCoarsestLanes)
    {
        up[i] = A.upper[i];
        own[i] = A.own[i];
        nei[i] = A.nei[i];
        lo[i] = A.losort[i];
        lOwn[i] = A.losortOwner[i];
    }
    for (int i = t; i <= n; i += kCoarsestLanes)
    {
        os[i] = A.ownerStart[i];
        ls[i] = A.losortStart[i];
    }
    __syncthreads();

    // y = A x, each cell's terms in ascending face order (its lower faces precede the faces it owns)
    auto amul = [&](const double* x, double* y) {
        for (int c = t; c < n; c += kCoarsestLanes)
        {
            double s = dg[c] * x[c];
            for (int k = ls[c]; k < ls[c + 1]; ++k) s += up[lo[k]] * x[lOwn[k]];
            for (int f = os[c]; f < os[c + 1]; ++f) s += up[f] * x[nei[f]];
            y[c] = s;
        }
    };

    amul(psi, wA);
    __syncthreads();
    for (int c = t; c < n; c += kCoarsestLanes)
    {
        rA[c] = b[c] - wA[c];
        double s = dg[c];   // sumA
        for (int k = ls[c]; k < ls[c + 1]; ++k) s += up[lo[k]];
        for (int f = os[c]; f < os[c + 1]; ++f) s += up[f];
        pA[c] = s;
    }
    __syncthreads();
    if (t == 0)
    {
        // normFactor (lduMatrixSolver.C:235-274) and the initial residual, host order
        double average = 0.0;
        for (int c = 0; c < n; ++c) average += psi[c];
        average /= (double)n;
        double sum = 0.0;
        for (int c = 0; c < n; ++c)
        {
            const double tt = pA[c] * average;
            sum += fabs(wA[c] - tt) + fabs(b[c] - tt);
        }
        const double nf = sum + kSmall;
        double sm = 0.0;
        for (int c = 0; c < n; ++c) sm += fabs(rA[c]);
        sc[0] = nf;
        sc[1] = sm / nf;
    }
    __syncthreads();
    const double nf = sc[0], init = sc[1];
    double fin = init;
    int it = 0;
    auto conv = [&](double r) { return r < tol || (relTol > kSmall && r < relTol * init); };
    if (!conv(fin))
    {
        for (int c = t; c < n; c += kCoarsestLanes) rD[c] = dg[c];
        __syncthreads();
        if (t == 0)
            for (int f = 0; f < nF; ++f) rD[nei[f]] -= up[f] * up[f] / rD[own[f]];
        __syncthreads();
        for (int c = t; c < n; c += kCoarsestLanes) rD[c] = 1.0 / rD[c];
        __syncthreads();
        double wArA = kGreat, wArAold;
        do
        {
            wArAold = wArA;
            for (int c = t; c < n; c += kCoarsestLanes) wA[c] = rD[c] * rA[c];
            __syncthreads();
            if (t == 0)
            {
                for (int f = 0; f < nF; ++f) wA[nei[f]] -= rD[nei[f]] * up[f] * wA[own[f]];
                for (int f = nF - 1; f >= 0; --f) wA[own[f]] -= rD[own[f]] * up[f] * wA[nei[f]];
                double s = 0.0;
                for (int c = 0; c < n; ++c) s += wA[c] * rA[c];
                sc[2] = s;
            }
            __syncthreads();
            wArA = sc[2];
            if (it == 0)
                for (int c = t; c < n; c += kCoarsestLanes) pA[c] = wA[c];
            else
            {
                const double beta = wArA / wArAold;
                for (int c = t; c < n; c += kCoarsestLanes) pA[c] = wA[c] + beta * pA[c];
            }
            __syncthreads();
            amul(pA, wA);
            __syncthreads();
            if (t == 0)
            {
                double s = 0.0;
                for (int c = 0; c < n; ++c) s += wA[c] * pA[c];
                sc[3] = s;
            }
            __syncthreads();
            const double wApA = sc[3];
            if (fabs(wApA) / nf < kVSmall) break;
            const double alpha = wArA / wApA;
            for (int c = t; c < n; c += kCoarsestLanes)
            {
                psi[c] += alpha * pA[c];
                rA[c] -= alpha * wA[c];
            }
            __syncthreads();
            if (t == 0)
            {
                double s = 0.0;
                for (int c = 0; c < n; ++c) s += fabs(rA[c]);
                sc[4] = s / nf;
            }
            __syncthreads();
            fin = sc[4];
        } while (++it < kCoarsestMaxIter && !conv(fin));
    }
    for (int c = t; c < n; c += kCoarsestLanes) psiOut[c] = psi[c];
    if (t == 0)
    {
        st[S_CINIT] = init;
        st[S_CFINAL] = fin;
        st[S_CITER] = (double)it;
    }
}

// ---- the cycle ----
// gamg.cu's scaleCorrection update (gamg.cu:207-219): sf = num/(den >= 0 ? max(den, VSMALL) : min(den, -VSMALL)),
// x = sf*x + (b - sf*Ax)/D
__global__ __launch_bounds__(kTpb) void scaleRefApplyK(int n, double* __restrict__ field, const double* __restrict__ Acf,
                                                       const double* __restrict__ source,
                                                       const double* __restrict__ diag, const double* __restrict__ st)
{
    const int i = blockIdx.x * kTpb + threadIdx.x;
    if (i >= n) return;
    const double num = st[S_NUM], den = st[S_DEN];
    const double sf = num / (den >= 0 ? fmax(den, kVSmall) : fmin(den, -kVSmall));
    field[i] = sf * field[i] + (source[i] - sf * Acf[i]) / diag[i];
}

// gamg.cu's denseSolve (gamg.cu:222-263) of the coarsest level, one lane: the LDU matrix scattered into a
// dense row-major n x n (diag, then per face upper at (own, nei) and lower at (nei, own)), Gaussian elimination
// with partial pivoting (rows swapped, the first largest |pivot| kept, zero factors skipped), back substitution
__global__ void denseRefSolveK(int n, int nF, const int* __restrict__ own, const int* __restrict__ nei,
                               const double* __restrict__ diag, const double* __restrict__ upper,
                               const double* __restrict__ lower, const double* __restrict__ b, double* __restrict__ x,
                               double* __restrict__ A)
{
    if (threadIdx.x || blockIdx.x) return;
    double* rhs = A + (size_t)n * n;
    for (size_t i = 0; i < (size_t)n * n; ++i) A[i] = 0.0;
    for (int i = 0; i < n; ++i)
    {
        rhs[i] = b[i];
        A[(size_t)i * n + i] = diag[i];
    }
    for (int f = 0; f < nF; ++f)
    {
        A[(size_t)own[f] * n + nei[f]] += upper[f];
        A[(size_t)nei[f] * n + own[f]] += lower[f];
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
}

struct Level
{
    Mesh* m;
    LduView A;
    double* rD;
};

class Cycle
{
public:
    Cycle(Ctx& c, Mesh& fine, Amg& a, const double* diag, const double* upper, const GamgControls& g)
        : c_(c), a_(a), g_(g), w_(*a.of)
    {
        const int nG = (int)a.grids.size();
        lv_.resize(nG);
        lv_[0] = Level{&fine, LduView{fine.nCells, fine.nFaces, diag, upper, upper}, w_.rD[0].p};
        for (int gi = 1; gi < nG; ++gi)
        {
            Grid& G = a.grids[gi];
            lv_[gi] = Level{w_.mesh[gi].get(), LduView{G.n, G.nFaces, G.diag64.p, G.upper64.p, G.upper64.p},
                            w_.rD[gi].p};
        }
        if (dic())
            for (Level& L : lv_) launchDicFactor(c_, *L.m, L.A, L.rD);
    }

    // GAMGSolver::Vcycle (gamg_solver_cpp.cu:627-735), on psi with its source; Apsi, fCorr scratch,
    // fRes the finest residual
    void vcycle(double* psi, const double* source, double* Apsi, double* fCorr, const double* fRes)
    {
        const int coarsest = (int)lv_.size() - 1;   // grid index of the coarsest level
        restrictTo(1, fRes);
        for (int gi = 1; gi < coarsest; ++gi)
        {
            Level& L = lv_[gi];
            double* corr = w_.corr[gi].p;
            double* src = w_.src[gi].p;
            const int li = gi - 1;
            if (g_.nPreSweeps)
            {
                fill(L.A.nCells, corr);
                smooth(L, corr, src, std::min(g_.nPreSweeps + g_.preSweepsLevelMultiplier * li, g_.maxPreSweeps));
                if (g_.scaleCorrection && gi < coarsest - 1) scale(L, corr, w_.ACf.p, src);
                // coarseSources -= A corr (the source is replaced by the residual)
                launchResidual(c_, *L.m, L.A, corr, src, w_.ACf.p);
                launchCopy(c_, L.A.nCells, w_.ACf.p, src);
            }
            restrictTo(gi + 1, src);
        }
        coarsestSolve();
        for (int gi = coarsest - 1; gi >= 1; --gi)
        {
            Level& L = lv_[gi];
            double* corr = w_.corr[gi].p;
            double* src = w_.src[gi].p;
            const int li = gi - 1;
            const int n = L.A.nCells;
            if (g_.nPreSweeps) launchCopy(c_, n, corr, w_.preSmoothed.p);
            prolongK<<<blocks(n), kTpb, 0, c_.stream>>>(n, a_.grids[gi].map.p, w_.corr[gi + 1].p, corr);
            if (g_.scaleCorrection && gi < coarsest - 1) scale(L, corr, w_.ACf.p, src);
            if (g_.nPreSweeps) addK<<<blocks(n), kTpb, 0, c_.stream>>>(n, w_.preSmoothed.p, corr);
            smooth(L, corr, src, std::min(g_.nPostSweeps + g_.postSweepsLevelMultiplier * li, g_.maxPostSweeps));
        }
        Level& F = lv_[0];
        const int n = F.A.nCells;
        prolongK<<<blocks(n), kTpb, 0, c_.stream>>>(n, a_.grids[0].map.p, w_.corr[1].p, fCorr);
        if (g_.scaleCorrection) scale(F, fCorr, Apsi, fRes);
        addK<<<blocks(n), kTpb, 0, c_.stream>>>(n, fCorr, psi);
        smooth(F, psi, source, g_.nFinestSweeps);
        hipCheck(hipGetLastError(), "gamg v-cycle");
    }

    const Level& fine() const { return lv_[0]; }

    // gamg.cu's cycle (gamg.cu:356-386) on psi with b, fRes the finest residual, fCorr and Apsi scratch
    void vcycleRef(double* psi, const double* b, double* Apsi, double* fCorr, const double* fRes, int nPost,
                   int nFinest)
    {
        const int coarsest = (int)lv_.size() - 1;
        restrictTo(1, fRes);
        for (int gi = 1; gi < coarsest; ++gi) restrictTo(gi + 1, w_.src[gi].p);
        {
            Level& L = lv_[coarsest];
            const Mesh& m = *L.m;
            denseRefSolveK<<<1, 64, 0, c_.stream>>>(L.A.nCells, L.A.nFaces, m.owner.p, m.neighbour.p, L.A.diag,
                                                    L.A.upper, L.A.lower, w_.src[coarsest].p, w_.corr[coarsest].p,
                                                    w_.dense.p);
        }
        for (int gi = coarsest - 1; gi >= 1; --gi)
        {
            Level& L = lv_[gi];
            double* corr = w_.corr[gi].p;
            double* src = w_.src[gi].p;
            const int n = L.A.nCells;
            prolongK<<<blocks(n), kTpb, 0, c_.stream>>>(n, a_.grids[gi].map.p, w_.corr[gi + 1].p, corr);
            scaleRef(L, corr, w_.ACf.p, src);
            if (nPost > 0) launchGaussSeidel(c_, *L.m, L.A, src, corr, nPost);
        }
        Level& F = lv_[0];
        const int n = F.A.nCells;
        prolongK<<<blocks(n), kTpb, 0, c_.stream>>>(n, a_.grids[0].map.p, w_.corr[1].p, fCorr);
        scaleRef(F, fCorr, Apsi, fRes);
        addK<<<blocks(n), kTpb, 0, c_.stream>>>(n, fCorr, psi);
        if (nFinest > 0) launchGaussSeidel(c_, *F.m, F.A, b, psi, nFinest);
        hipCheck(hipGetLastError(), "gamg.cu v-cycle");
    }

private:
    bool dic() const { return g_.smoother == OfSmoother::DIC || g_.smoother == OfSmoother::DICGaussSeidel; }

    void fill(int n, double* x) { fillK<<<blocks(n), kTpb, 0, c_.stream>>>(n, 0.0, x); }

    // src of grid gi = restriction of the finer grid's field
    void restrictTo(int gi, const double* finer)
    {
        const Grid& G = a_.grids[gi - 1];
        const int nc = a_.grids[gi].n;
        restrictK<<<blocks(nc), kTpb, 0, c_.stream>>>(nc, G.cellStart.p, G.cellList.p, finer, w_.src[gi].p);
    }

    void smooth(Level& L, double* psi, const double* source, int nSweeps)
    {
        if (nSweeps <= 0) return;
        if (dic()) launchDicSmooth(c_, *L.m, L.A, L.rD, source, psi, nSweeps, w_.tR.p, w_.tW.p);
        if (g_.smoother == OfSmoother::GaussSeidel || g_.smoother == OfSmoother::DICGaussSeidel)
            launchGaussSeidel(c_, *L.m, L.A, source, psi, nSweeps);
        if (g_.smoother == OfSmoother::SymGaussSeidel) launchGaussSeidel(c_, *L.m, L.A, source, psi, nSweeps, w_.bPrime.p);
    }

    // gamg.cu's scaleCorrection: the two sums each in ascending cell order (sequential mode), then the update
    void scaleRef(Level& L, double* field, double* Acf, const double* source)
    {
        const int n = L.A.nCells;
        launchAmul(c_, *L.m, L.A, field, Acf);
        launchDot(c_, n, field, source, w_.st.p + S_NUM, 0);
        launchDot(c_, n, field, Acf, w_.st.p + S_DEN, 1);
        scaleRefApplyK<<<blocks(n), kTpb, 0, c_.stream>>>(n, field, Acf, source, L.A.diag, w_.st.p);
    }

    void scale(Level& L, double* field, double* Acf, const double* source)
    {
        const int n = L.A.nCells;
        launchAmul(c_, *L.m, L.A, field, Acf);
        launchDot(c_, n, field, source, w_.st.p + S_NUM, 0);
        launchDot(c_, n, field, Acf, w_.st.p + S_DEN, 1);
        scaleApplyK<<<blocks(n), kTpb, 0, c_.stream>>>(n, field, Acf, source, L.A.diag, w_.st.p);
    }

    void coarsestSolve()
    {
        const int gi = (int)lv_.size() - 1;
        Level& L = lv_[gi];
        const Mesh& m = *L.m;
        const int n = L.A.nCells, nF = L.A.nFaces;
        double* corr = w_.corr[gi].p;
        const size_t bytes = coarseLdsBytes(n, nF);
        if (bytes <= kCoarsestLds)
        {
            CoarseLdu A{n, nF, m.owner.p, m.neighbour.p, m.ownerStart.p, m.losortStart.p, m.losort.p, m.losortOwner.p,
                        L.A.diag, L.A.upper};
            coarsestPcgDicK<<<1, kCoarsestLanes, bytes, c_.stream>>>(A, w_.src[gi].p, corr, g_.tolerance, g_.relTol,
                                                                     w_.st.p);
            hipCheck(hipGetLastError(), "gamg coarsest");
            return;
        }
        fill(n, corr);
        PcgControls pc;
        pc.tolerance = g_.tolerance;
        pc.relTol = g_.relTol;
        pc.maxIter = kCoarsestMaxIter;
        pc.minIter = 0;
        pc.checkEvery = 4;
        pcgDicSolve(c_, *L.m, L.A, w_.src[gi].p, corr, nullptr, pc, true);
    }

    Ctx& c_;
    Amg& a_;
    const GamgControls& g_;
    OfWork& w_;
    std::vector<Level> lv_;
};

void checkControls(const GamgControls& g)
{
    const int s = (int)g.smoother;
    if (s < 0 || s > 3) invalid("gamg: unknown smoother");
    if (g.maxIter < 0 || g.minIter < 0) invalid("gamg: negative iteration bound");
    if (g.nPreSweeps < 0 || g.nPostSweeps < 0 || g.nFinestSweeps < 0 || g.maxPreSweeps < 0 || g.maxPostSweeps < 0)
        invalid("gamg: negative sweep count");
    if (g.nVcycles < 1) invalid("gamg: nVcycles must be at least 1");
}

// the per-hierarchy state: coarse LDU addressing (once), vectors sized to the grids
void prepare(Ctx& c, const Mesh& fine, Amg& a)
{
    const int nG = (int)a.grids.size();
    if (nG < 2) invalid("gamg: the mesh is too small for a coarse grid");
    if (a.nFine != fine.nCells) invalid("gamg: hierarchy of another mesh");
    if (a.of && (int)a.of->mesh.size() == nG) return;
    auto w = std::make_unique<OfWork>();
    w->mesh.resize(nG);
    w->corr.resize(nG);
    w->src.resize(nG);
    w->rD.resize(nG);
    const int nf = fine.nCells;
    for (int gi = 1; gi < nG; ++gi)
    {
        const Grid& G = a.grids[gi];
        w->mesh[gi] = std::make_unique<Mesh>();
        buildMesh(c, *w->mesh[gi], G.n, G.nFaces, G.hOwn.data(), G.hNei.data(), 0, nullptr);
        w->corr[gi].resize(G.n);
        w->src[gi].resize(G.n);
        w->rD[gi].resize(G.n);
    }
    w->rD[0].resize(nf);
    for (DBuf<double>* b : {&w->ACf, &w->preSmoothed, &w->tR, &w->tW, &w->bPrime, &w->Apsi, &w->fCorr, &w->fRes,
                            &w->pA, &w->wA, &w->rA, &w->AwA})
        b->resize(nf);
    w->st.resize(S_COUNT);
    a.of = std::move(w);
}

// read back device scalars [first, first + count)
void readScalars(Ctx& c, const double* st, int first, int count, double* out)
{
    hipCheck(hipMemcpyAsync(c.mailbox, st + first, sizeof(double) * count, hipMemcpyDeviceToHost, c.stream), "read");
    hipCheck(hipStreamSynchronize(c.stream), "read sync");
    std::copy(c.mailbox, c.mailbox + count, out);
}

bool converged(const PcgPerf& p, double tol, double relTol)
{
    return p.finalResidual < tol || (relTol > kSmall && p.finalResidual < relTol * p.initialResidual);
}

} // namespace

PcgPerf gamgSolve(Ctx& c, Mesh& m, Amg& a, const double* diag, const double* upper, const double* b, double* psi,
                  const GamgControls& g)
{
    checkControls(g);
    prepare(c, m, a);
    OfWork& w = *a.of;
    const int n = m.nCells;
    const LduView A{n, m.nFaces, diag, upper, upper};
    amgCoarseMatrices(c, a, diag, upper);
    Cycle cyc(c, m, a, diag, upper, g);
    double* st = w.st.p;
    launchAmul(c, m, A, psi, w.Apsi.p);
    launchNormFactor(c, m, A, psi, b, st + S_NF);
    subK<<<blocks(n), kTpb, 0, c.stream>>>(n, b, w.Apsi.p, w.fRes.p);
    launchSumMag(c, n, w.fRes.p, st + S_SUMMAG, 2);
    double h[2];
    readScalars(c, st, S_NF, 2, h);
    const double nf = h[0];
    PcgPerf perf;
    perf.initialResidual = h[1] / nf;
    perf.finalResidual = perf.initialResidual;
    perf.nIterations = 0;
    if (!(g.minIter <= 0 && converged(perf, g.tolerance, g.relTol)))
    {
        do
        {
            cyc.vcycle(psi, b, w.Apsi.p, w.fCorr.p, w.fRes.p);
            launchAmul(c, m, A, psi, w.Apsi.p);
            subK<<<blocks(n), kTpb, 0, c.stream>>>(n, b, w.Apsi.p, w.fRes.p);
            launchSumMag(c, n, w.fRes.p, st + S_SUMMAG, 2);
            readScalars(c, st, S_SUMMAG, 1, h);
            perf.finalResidual = h[0] / nf;
        } while ((++perf.nIterations < g.maxIter && !converged(perf, g.tolerance, g.relTol)) ||
                 perf.nIterations < g.minIter);
    }
    perf.converged = converged(perf, g.tolerance, g.relTol);
    return perf;
}

PcgPerf gamgRefSolve(Ctx& c, Mesh& m, Amg& a, const double* diag, const double* upper, const double* b, double* psi,
                     const GamgRefControls& g)
{
    if (g.maxIter < 0 || g.minIter < 0) invalid("gamg.cu: negative iteration bound");
    if (g.nPostSweeps < 0 || g.nFinestSweeps < 0) invalid("gamg.cu: negative sweep count");
    if (a.controls.kind != AggKind::Pair || a.controls.mergeLevels != 1)
        throw Error(kErrUnsupported, "mfv: gamg.cu: the hierarchy must be gamg.cu's agglomeratePair levels "
                                     "(MFV_AGG_PAIR, mergeLevels 1)");
    prepare(c, m, a);
    OfWork& w = *a.of;
    const int coarseN = a.grids.back().n;
    w.dense.resize((size_t)coarseN * coarseN + coarseN);
    const int n = m.nCells;
    const LduView A{n, m.nFaces, diag, upper, upper};
    amgCoarseMatrices(c, a, diag, upper);
    GamgControls gc;
    gc.smoother = OfSmoother::GaussSeidel;
    Cycle cyc(c, m, a, diag, upper, gc);
    double* st = w.st.p;
    launchAmul(c, m, A, psi, w.Apsi.p);
    launchNormFactor(c, m, A, psi, b, st + S_NF);
    subK<<<blocks(n), kTpb, 0, c.stream>>>(n, b, w.Apsi.p, w.fRes.p);
    launchSumMag(c, n, w.fRes.p, st + S_SUMMAG, 2);
    double h[2];
    readScalars(c, st, S_NF, 2, h);
    const double nf = h[0];
    PcgPerf perf;
    perf.initialResidual = h[1] / nf;
    perf.finalResidual = perf.initialResidual;
    perf.nIterations = 0;
    // gamg.cu's stopping rule: relTol only when positive
    auto conv = [&](double fr) { return fr < g.tolerance || (g.relTol > 0.0 && fr < g.relTol * perf.initialResidual); };
    if (g.minIter > 0 || !conv(perf.finalResidual))
    {
        int it = 0;
        do
        {
            cyc.vcycleRef(psi, b, w.Apsi.p, w.fCorr.p, w.fRes.p, g.nPostSweeps, g.nFinestSweeps);
            launchAmul(c, m, A, psi, w.Apsi.p);
            subK<<<blocks(n), kTpb, 0, c.stream>>>(n, b, w.Apsi.p, w.fRes.p);
            launchSumMag(c, n, w.fRes.p, st + S_SUMMAG, 2);
            readScalars(c, st, S_SUMMAG, 1, h);
            perf.finalResidual = h[0] / nf;
            ++it;
        } while ((it < g.maxIter && !conv(perf.finalResidual)) || it < g.minIter);
        perf.nIterations = it;
    }
    perf.converged = conv(perf.finalResidual);
    return perf;
}

PcgPerf pcgGamgSolve(Ctx& c, Mesh& m, Amg& a, const double* diag, const double* upper, const double* b, double* psi,
                     double tolerance, double relTol, int maxIter, int minIter, const GamgControls& g)
{
    checkControls(g);
    if (maxIter < 0 || minIter < 0) invalid("pcg-gamg: negative iteration bound");
    prepare(c, m, a);
    OfWork& w = *a.of;
    const int n = m.nCells;
    const LduView A{n, m.nFaces, diag, upper, upper};
    double* st = w.st.p;
    launchAmul(c, m, A, psi, w.wA.p);
    subK<<<blocks(n), kTpb, 0, c.stream>>>(n, b, w.wA.p, w.rA.p);
    launchNormFactor(c, m, A, psi, b, st + S_NF);
    launchSumMag(c, n, w.rA.p, st + S_SUMMAG, 2);
    double h[4];
    readScalars(c, st, S_NF, 2, h);
    const double nf = h[0];
    PcgPerf perf;
    perf.initialResidual = h[1] / nf;
    perf.finalResidual = perf.initialResidual;
    perf.nIterations = 0;
    if (minIter <= 0 && converged(perf, tolerance, relTol))
    {
        perf.converged = true;
        return perf;
    }
    // the preconditioner is built only now (PCG.C:111-118)
    amgCoarseMatrices(c, a, diag, upper);
    Cycle cyc(c, m, a, diag, upper, g);
    do
    {
        // GAMGPreconditioner::precondition: nVcycles cycles from zero, the residual recomputed between
        fillK<<<blocks(n), kTpb, 0, c.stream>>>(n, 0.0, w.wA.p);
        launchCopy(c, n, w.rA.p, w.fRes.p);
        for (int k = 0; k < g.nVcycles; ++k)
        {
            cyc.vcycle(w.wA.p, w.rA.p, w.AwA.p, w.fCorr.p, w.fRes.p);
            if (k < g.nVcycles - 1)
            {
                launchAmul(c, m, A, w.wA.p, w.AwA.p);
                subK<<<blocks(n), kTpb, 0, c.stream>>>(n, w.rA.p, w.AwA.p, w.fRes.p);
            }
        }
        shiftRhoK<<<1, 1, 0, c.stream>>>(st);
        launchDot(c, n, w.wA.p, w.rA.p, st + S_RHO, 0);
        pcgDirK<<<blocks(n), kTpb, 0, c.stream>>>(n, perf.nIterations == 0, w.wA.p, w.pA.p, st);
        launchAmul(c, m, A, w.pA.p, w.wA.p);
        launchDot(c, n, w.wA.p, w.pA.p, st + S_PQ, 1);
        pcgUpdK<<<blocks(n), kTpb, 0, c.stream>>>(n, w.pA.p, w.wA.p, psi, w.rA.p, st);
        launchSumMag(c, n, w.rA.p, st + S_SUMMAG, 2);
        hipCheck(hipGetLastError(), "pcg-gamg iteration");
        readScalars(c, st, S_SUMMAG, 4, h);   // sum|r|, rho, rhoOld, p.q
        if (std::fabs(h[3]) / nf < kVSmall) break;
        perf.finalResidual = h[0] / nf;
    } while ((++perf.nIterations < maxIter && !converged(perf, tolerance, relTol)) || perf.nIterations < minIter);
    perf.converged = converged(perf, tolerance, relTol);
    return perf;
}

} // namespace mfv
