// The sequential sweeps of OpenFOAM's solvers on the device, in the host loop's exact order: the DIC
// preconditioner (factor and forward/backward substitution) and Gauss-Seidel, and the PCG with DIC
// that the reference's `pcg` runs (src/OpenFOAM/matrices/pcg.cu:6-149).
//
// A sweep that walks the cells in index order is a chain of dependencies along the lower (forward) or
// upper (backward) neighbours. The mesh's level schedule groups cells whose dependencies are all in
// earlier groups; each group is one launch, one cell per thread, and every cell applies its own
// updates in the host loop's face order (the face loops of the reference visit a cell's lower faces in
// ascending face index, its owned faces ascending for Gauss-Seidel and descending for the backward
// substitution). The results are the host loop's to the bit. The reference's device path does the same
// with its own level schedule (device_sym_gauss_seidel.cuh, device_dilu.cuh).
#include "amg.hpp"
#include "reduce.hpp"

#include <algorithm>

namespace mfv {

namespace {

constexpr int kTpb = 256;
inline int blocks(int n) { return (n + kTpb - 1) / kTpb; }

void buildLevels(Ctx& c, Mesh& m)
{
    if (m.sched) return;
    auto s = std::make_unique<LevelSched>();
    const int n = m.nCells, nF = m.nFaces;
    std::vector<int> lf(n, 0), lb(n, 0);
    // faces ascending: a face's owner has all of its own lower faces earlier, so its level is final
    for (int f = 0; f < nF; ++f) lf[m.hNeighbour[f]] = std::max(lf[m.hNeighbour[f]], lf[m.hOwner[f]] + 1);
    for (int f = nF - 1; f >= 0; --f) lb[m.hOwner[f]] = std::max(lb[m.hOwner[f]], lb[m.hNeighbour[f]] + 1);
    auto bucket = [n](const std::vector<int>& lev, std::vector<int>& start, std::vector<int>& cells) {
        const int nL = n ? *std::max_element(lev.begin(), lev.end()) + 1 : 0;
        start.assign(nL + 1, 0);
        for (int c = 0; c < n; ++c) ++start[lev[c] + 1];
        for (int l = 0; l < nL; ++l) start[l + 1] += start[l];
        std::vector<int> at(start.begin(), start.end() - 1);
        cells.resize(n);
        for (int c = 0; c < n; ++c) cells[at[lev[c]]++] = c;
    };
    std::vector<int> fc, bc;
    bucket(lf, s->fwdStart, fc);
    bucket(lb, s->bwdStart, bc);
    s->fwdCells.upload(fc, c.stream);
    s->bwdCells.upload(bc, c.stream);
    hipCheck(hipStreamSynchronize(c.stream), "level schedule");
    m.sched = std::move(s);
}

// rD before inversion: diag - sum over lower faces (ascending) of upper^2 / rD[lower neighbour]
// (DICPreconditioner::calcReciprocalD; pcg.cu, gamg_solver_cpp.cu:194-210)
__global__ __launch_bounds__(kTpb) void dicFactorK(int cnt, const int* __restrict__ cells,
                                                   const label* __restrict__ losortStart,
                                                   const label* __restrict__ losort,
                                                   const label* __restrict__ losortOwner,
                                                   const double* __restrict__ diag, const double* __restrict__ upper,
                                                   double* __restrict__ rD)
{
    const int i = blockIdx.x * kTpb + threadIdx.x;
    if (i >= cnt) return;
    const int c = cells[i];
    double d = diag[c];
    for (int k = losortStart[c]; k < losortStart[c + 1]; ++k)
    {
        const double u = upper[losort[k]];
        d -= u * u / rD[losortOwner[k]];
    }
    rD[c] = d;
}

__global__ __launch_bounds__(kTpb) void invertK(int n, double* __restrict__ rD, const int* __restrict__ stop)
{
    if (stop && *stop) return;
    const int c = blockIdx.x * kTpb + threadIdx.x;
    if (c < n) rD[c] = 1.0 / rD[c];
}

// forward substitution: w = rD*r, then w[c] -= rD[c]*upper*w[lower] over lower faces ascending
__global__ __launch_bounds__(kTpb) void dicFwdK(int cnt, const int* __restrict__ cells,
                                                const label* __restrict__ losortStart, const label* __restrict__ losort,
                                                const label* __restrict__ losortOwner, const double* __restrict__ rD,
                                                const double* __restrict__ upper, const double* __restrict__ r,
                                                double* __restrict__ w, const int* __restrict__ stop)
{
    if (stop && *stop) return;
    const int i = blockIdx.x * kTpb + threadIdx.x;
    if (i >= cnt) return;
    const int c = cells[i];
    const double rd = rD[c];
    double v = rd * r[c];
    for (int k = losortStart[c]; k < losortStart[c + 1]; ++k) v -= rd * upper[losort[k]] * w[losortOwner[k]];
    w[c] = v;
}

// backward substitution: w[c] -= rD[c]*upper*w[upper neighbour] over owned faces descending
__global__ __launch_bounds__(kTpb) void dicBwdK(int cnt, const int* __restrict__ cells,
                                                const label* __restrict__ ownerStart,
                                                const label* __restrict__ neighbour, const double* __restrict__ rD,
                                                const double* __restrict__ upper, double* __restrict__ w,
                                                const int* __restrict__ stop)
{
    if (stop && *stop) return;
    const int i = blockIdx.x * kTpb + threadIdx.x;
    if (i >= cnt) return;
    const int c = cells[i];
    const double rd = rD[c];
    double v = w[c];
    for (int f = ownerStart[c + 1] - 1; f >= ownerStart[c]; --f) v -= rd * upper[f] * w[neighbour[f]];
    w[c] = v;
}

// Gauss-Seidel (GaussSeidelSmoother, smooth_solver_cpp.cu:87-155): bPrime[c] = b[c] less the lower
// neighbours' new values (ascending), then psi[c] = (bPrime[c] - sum upper*psi_old[upper]) / diag[c]
__global__ __launch_bounds__(kTpb) void gsFwdK(int cnt, const int* __restrict__ cells,
                                               const label* __restrict__ ownerStart,
                                               const label* __restrict__ neighbour,
                                               const label* __restrict__ losortStart, const label* __restrict__ losort,
                                               const label* __restrict__ losortOwner, const double* __restrict__ diag,
                                               const double* __restrict__ upper, const double* __restrict__ b,
                                               double* __restrict__ psi, double* __restrict__ bPrime)
{
    const int i = blockIdx.x * kTpb + threadIdx.x;
    if (i >= cnt) return;
    const int c = cells[i];
    double t = b[c];
    for (int k = losortStart[c]; k < losortStart[c + 1]; ++k) t -= upper[losort[k]] * psi[losortOwner[k]];
    if (bPrime) bPrime[c] = t;
    for (int f = ownerStart[c]; f < ownerStart[c + 1]; ++f) t -= upper[f] * psi[neighbour[f]];
    psi[c] = t / diag[c];
}

// symGaussSeidel's reverse walk (smooth_solver_cpp.cu:141-163), cells on the backward schedule: the
// forward walk's bPrime[c] (b less the lower neighbours' forward values), less the upper neighbours'
// new values in ascending face order, over diag[c]
__global__ __launch_bounds__(kTpb) void gsBwdK(int cnt, const int* __restrict__ cells,
                                               const label* __restrict__ ownerStart,
                                               const label* __restrict__ neighbour, const double* __restrict__ diag,
                                               const double* __restrict__ upper,
                                               const double* __restrict__ bPrime, double* __restrict__ psi)
{
    const int i = blockIdx.x * kTpb + threadIdx.x;
    if (i >= cnt) return;
    const int c = cells[i];
    double t = bPrime[c];
    for (int f = ownerStart[c]; f < ownerStart[c + 1]; ++f) t -= upper[f] * psi[neighbour[f]];
    psi[c] = t / diag[c];
}

__global__ __launch_bounds__(kTpb) void addToK(int n, const double* __restrict__ a, double* __restrict__ y)
{
    const int c = blockIdx.x * kTpb + threadIdx.x;
    if (c < n) y[c] += a[c];
}

__global__ __launch_bounds__(kTpb) void subK(int n, const double* __restrict__ b, const double* __restrict__ ax,
                                             double* __restrict__ r)
{
    const int c = blockIdx.x * kTpb + threadIdx.x;
    if (c < n) r[c] = b[c] - ax[c];
}

// ---- the PCG-DIC recurrence (pcg.cu:6-149); device state in st/fl ----
enum : int { D_RHO = 0, D_RHOOLD, D_PQ, D_NF, D_INIT, D_FINAL, D_TOL, D_RELTOL, D_T0, D_T1, D_T2, D_RELMIN,
             D_COUNT = 16 };
enum : int { G_ITER = 0, G_STOP, G_CONV, G_MAXIT, G_MINIT, G_COUNT = 8 };

__device__ __forceinline__ bool dicConverged(const double* st, double res)
{
    return res < st[D_TOL] || (st[D_RELTOL] > st[D_RELMIN] && res < st[D_RELTOL] * st[D_INIT]);
}

__global__ void dicSetupK(double* st, int* fl, double tol, double relTol, double relMin, int maxIter, int minIter,
                          const double* nf)
{
    if (threadIdx.x || blockIdx.x) return;
    st[D_TOL] = tol;
    st[D_RELTOL] = relTol;
    st[D_RELMIN] = relMin;
    if (nf) st[D_NF] = *nf;
    fl[G_MAXIT] = maxIter;
    fl[G_MINIT] = minIter;
    fl[G_ITER] = 0;
    st[D_RHO] = 1e300;
}

// initial residual; iterate when minIter > 0 or not converged
__global__ void dicInitK(double* st, int* fl)
{
    if (threadIdx.x || blockIdx.x) return;
    const double init = st[D_T2] / st[D_NF];
    st[D_INIT] = init;
    st[D_FINAL] = init;
    const bool conv = dicConverged(st, init);
    fl[G_CONV] = conv ? 1 : 0;
    fl[G_STOP] = (fl[G_MINIT] > 0 || !conv) ? 0 : 1;
}

__global__ void dicRhoK(double* st, const int* fl)
{
    if (threadIdx.x || blockIdx.x || fl[G_STOP]) return;
    st[D_RHOOLD] = st[D_RHO];
    st[D_RHO] = st[D_T0];
}

__global__ __launch_bounds__(kTpb) void dicDirK(int n, const double* __restrict__ w, double* __restrict__ p,
                                                const double* __restrict__ st, const int* __restrict__ fl)
{
    if (fl[G_STOP]) return;
    const int c = blockIdx.x * kTpb + threadIdx.x;
    if (c >= n) return;
    if (fl[G_ITER] == 0)
        p[c] = w[c];
    else
    {
        const double beta = st[D_RHO] / st[D_RHOOLD];
        p[c] = w[c] + beta * p[c];
    }
}

// the guarded solver's singularity test: no update, no count, stop
__global__ void dicGuardK(double* st, int* fl)
{
    if (threadIdx.x || blockIdx.x || fl[G_STOP]) return;
    if (fabs(st[D_T1]) / st[D_NF] < 1e-300) fl[G_STOP] = 1;
}

__global__ __launch_bounds__(kTpb) void dicUpdK(int n, const double* __restrict__ p, const double* __restrict__ q,
                                                double* __restrict__ psi, double* __restrict__ r,
                                                const double* __restrict__ st, const int* __restrict__ fl)
{
    if (fl[G_STOP]) return;
    const int c = blockIdx.x * kTpb + threadIdx.x;
    if (c >= n) return;
    const double alpha = st[D_RHO] / st[D_T1];
    psi[c] += alpha * p[c];
    r[c] -= alpha * q[c];
}

__global__ void dicResK(double* st, int* fl)
{
    if (threadIdx.x || blockIdx.x || fl[G_STOP]) return;
    const double res = st[D_T2] / st[D_NF];
    st[D_FINAL] = res;
    const int it = fl[G_ITER] + 1;
    fl[G_ITER] = it;
    const bool conv = dicConverged(st, res);
    fl[G_CONV] = conv ? 1 : 0;
    fl[G_STOP] = ((it < fl[G_MAXIT] && !conv) || it < fl[G_MINIT]) ? 0 : 1;
}

void dicApply(Ctx& c, Mesh& m, const LduView& A, const double* rD, const double* r, double* w, const int* stop)
{
    const LevelSched& s = *m.sched;
    for (size_t l = 0; l + 1 < s.fwdStart.size(); ++l)
    {
        const int a = s.fwdStart[l], cnt = s.fwdStart[l + 1] - a;
        dicFwdK<<<blocks(cnt), kTpb, 0, c.stream>>>(cnt, s.fwdCells.p + a, m.losortStart.p, m.losort.p,
                                                    m.losortOwner.p, rD, A.upper, r, w, stop);
    }
    for (size_t l = 0; l + 1 < s.bwdStart.size(); ++l)
    {
        const int a = s.bwdStart[l], cnt = s.bwdStart[l + 1] - a;
        dicBwdK<<<blocks(cnt), kTpb, 0, c.stream>>>(cnt, s.bwdCells.p + a, m.ownerStart.p, m.neighbour.p, rD, A.upper,
                                                    w, stop);
    }
}

void dicFactor(Ctx& c, Mesh& m, const LduView& A, double* rD, const int* stop)
{
    const LevelSched& s = *m.sched;
    for (size_t l = 0; l + 1 < s.fwdStart.size(); ++l)
    {
        const int a = s.fwdStart[l], cnt = s.fwdStart[l + 1] - a;
        dicFactorK<<<blocks(cnt), kTpb, 0, c.stream>>>(cnt, s.fwdCells.p + a, m.losortStart.p, m.losort.p,
                                                       m.losortOwner.p, A.diag, A.upper, rD);
    }
    invertK<<<blocks(m.nCells), kTpb, 0, c.stream>>>(m.nCells, rD, stop);
}

} // namespace

void buildLevelSchedule(Ctx& c, Mesh& m) { buildLevels(c, m); }

void launchDicPrecondition(Ctx& c, Mesh& m, const LduView& A, const double* r, double* w, double* rD)
{
    if (A.lower != A.upper) throw Error(kErrUnsupported, "mfv: DIC needs a symmetric matrix (lower == upper)");
    if (!m.nCells) return;
    buildLevels(c, m);
    dicFactor(c, m, A, rD, nullptr);
    dicApply(c, m, A, rD, r, w, nullptr);
    hipCheck(hipGetLastError(), "dic");
}

void launchGaussSeidel(Ctx& c, Mesh& m, const LduView& A, const double* b, double* psi, int nSweeps, double* bPrime)
{
    if (A.lower != A.upper) throw Error(kErrUnsupported, "mfv: Gauss-Seidel here takes a symmetric matrix");
    if (!m.nCells) return;
    buildLevels(c, m);
    const LevelSched& s = *m.sched;
    for (int sw = 0; sw < nSweeps; ++sw)
    {
        for (size_t l = 0; l + 1 < s.fwdStart.size(); ++l)
        {
            const int a = s.fwdStart[l], cnt = s.fwdStart[l + 1] - a;
            gsFwdK<<<blocks(cnt), kTpb, 0, c.stream>>>(cnt, s.fwdCells.p + a, m.ownerStart.p, m.neighbour.p,
                                                       m.losortStart.p, m.losort.p, m.losortOwner.p, A.diag, A.upper,
                                                       b, psi, bPrime);
        }
        if (!bPrime) continue;
        for (size_t l = 0; l + 1 < s.bwdStart.size(); ++l)
        {
            const int a = s.bwdStart[l], cnt = s.bwdStart[l + 1] - a;
            gsBwdK<<<blocks(cnt), kTpb, 0, c.stream>>>(cnt, s.bwdCells.p + a, m.ownerStart.p, m.neighbour.p, A.diag,
                                                       A.upper, bPrime, psi);
        }
    }
    hipCheck(hipGetLastError(), "gauss-seidel");
}

void launchDicFactor(Ctx& c, Mesh& m, const LduView& A, double* rD)
{
    if (A.lower != A.upper) throw Error(kErrUnsupported, "mfv: DIC needs a symmetric matrix (lower == upper)");
    if (!m.nCells) return;
    buildLevels(c, m);
    dicFactor(c, m, A, rD, nullptr);
    hipCheck(hipGetLastError(), "dic factor");
}

void launchDicSmooth(Ctx& c, Mesh& m, const LduView& A, const double* rD, const double* b, double* psi, int nSweeps,
                     double* rA, double* w)
{
    if (A.lower != A.upper) throw Error(kErrUnsupported, "mfv: the DIC smoother needs a symmetric matrix");
    if (!m.nCells) return;
    buildLevels(c, m);
    for (int sw = 0; sw < nSweeps; ++sw)
    {
        launchResidual(c, m, A, psi, b, rA);
        dicApply(c, m, A, rD, rA, w, nullptr);
        addToK<<<blocks(m.nCells), kTpb, 0, c.stream>>>(m.nCells, w, psi);
    }
    hipCheck(hipGetLastError(), "dic smoother");
}

PcgPerf pcgDicSolve(Ctx& c, Mesh& m, const LduView& A, const double* b, double* psi, const double* normFactor,
                    const PcgControls& pc, bool guarded)
{
    if (A.lower != A.upper) throw Error(kErrUnsupported, "mfv: PCG-DIC solves symmetric matrices (lower == upper)");
    const int n = m.nCells;
    buildLevels(c, m);
    DBuf<double> rD(n), r(n), w(n), p(n), q(n), st(D_COUNT);
    DBuf<int> fl(G_COUNT);
    double* S = st.p;
    int* F = fl.p;
    const int* stop = F + G_STOP;
    dicSetupK<<<1, 1, 0, c.stream>>>(S, F, pc.tolerance, pc.relTol, guarded ? 1e-20 : 0.0, pc.maxIter,
                                     guarded ? 0 : pc.minIter, normFactor);
    // r = b - A psi (the host loop's Amul, then the subtraction)
    launchAmul(c, m, A, psi, q.p);
    subK<<<blocks(n), kTpb, 0, c.stream>>>(n, b, q.p, r.p);
    if (!normFactor) launchNormFactor(c, m, A, psi, b, S + D_NF);
    launchSumMag(c, n, r.p, S + D_T2, 0);
    dicInitK<<<1, 1, 0, c.stream>>>(S, F);
    dicFactor(c, m, A, rD.p, nullptr);
    hipCheck(hipGetLastError(), "pcg-dic init");
    for (int it = 0; it < std::max(pc.maxIter, pc.minIter); ++it)
    {
        dicApply(c, m, A, rD.p, r.p, w.p, stop);
        launchDot(c, n, w.p, r.p, S + D_T0, 0);
        dicRhoK<<<1, 1, 0, c.stream>>>(S, F);
        dicDirK<<<blocks(n), kTpb, 0, c.stream>>>(n, w.p, p.p, S, F);
        launchAmul(c, m, A, p.p, q.p);
        launchDot(c, n, q.p, p.p, S + D_T1, 1);
        if (guarded) dicGuardK<<<1, 1, 0, c.stream>>>(S, F);
        dicUpdK<<<blocks(n), kTpb, 0, c.stream>>>(n, p.p, q.p, psi, r.p, S, F);
        launchSumMag(c, n, r.p, S + D_T2, 2);
        dicResK<<<1, 1, 0, c.stream>>>(S, F);
        hipCheck(hipGetLastError(), "pcg-dic iteration");
        if ((it + 1) % std::max(1, pc.checkEvery) == 0)
        {
            hipCheck(hipMemcpyAsync(c.mailbox, F, sizeof(int) * G_COUNT, hipMemcpyDeviceToHost, c.stream), "poll");
            hipCheck(hipStreamSynchronize(c.stream), "poll sync");
            if (reinterpret_cast<const int*>(c.mailbox)[G_STOP]) break;
        }
    }
    hipCheck(hipMemcpyAsync(c.mailbox, S, sizeof(double) * D_COUNT, hipMemcpyDeviceToHost, c.stream), "report");
    hipCheck(hipMemcpyAsync(c.mailbox + D_COUNT, F, sizeof(int) * G_COUNT, hipMemcpyDeviceToHost, c.stream),
             "report flags");
    hipCheck(hipStreamSynchronize(c.stream), "report sync");
    const double* st_h = c.mailbox;
    const int* fl_h = reinterpret_cast<const int*>(c.mailbox + D_COUNT);
    PcgPerf perf;
    perf.initialResidual = st_h[D_INIT];
    perf.finalResidual = st_h[D_FINAL];
    perf.nIterations = fl_h[G_ITER];
    perf.converged = fl_h[G_CONV] != 0;
    return perf;
}

} // namespace mfv
