// The serving loop: a stream of pressure-Poisson requests on hex boxes of uneven sizes, admitted by an
// HBM budget into a fixed number of slots, each slot a host thread with its own context (HIP stream),
// so admitted requests run concurrently on the device. A request keeps its mesh, matrix and AMG
// hierarchy resident across its sequence of right-hand sides; with resident systems a slot keeps them
// after the request for the next one on the same box. Small requests can hand their solves to a
// batcher that runs every waiting solve whole, one workgroup each, in one launch. There is no
// counterpart in the reference (its solvers run one case per process); the request shape follows its
// pressure corrector: 7-point FP64 laplacian on the box, zeroGradient walls with a reference cell,
// right-hand side = -div of a seeded random velocity, mixed-precision AMG-PCG.
//
// Results do not depend on what else is running: every kernel's arithmetic order is fixed by its
// inputs (no atomics, per-context scratch, one workgroup per batched solve), so a request's solution is
// the same bits alone or in a mix.
#include "mfv.h"
#include "mfv_internal.hpp"
#include "reduce.hpp"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <list>
#include <map>
#include <memory>
#include <tuple>
#include <mutex>
#include <thread>
#include <vector>

namespace mfv {

namespace {

constexpr int kTpb = 256;
inline int blocks(long long n) { return (int)((n + kTpb - 1) / kTpb); }

__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// seeded uniform(-1, 1) velocity per cell component (53-bit mantissa mapping of a counter hash)
__global__ void velocityK(int n, uint64_t seed, double* __restrict__ U)
{
    const int i = blockIdx.x * kTpb + threadIdx.x;
    if (i >= 3 * n) return;
    const uint64_t h = mix64(seed * 0x100000001b3ull + (uint64_t)i);
    U[i] = 2.0 * ((double)(h >> 11) * (1.0 / 9007199254740992.0)) - 1.0;
}

// phi = interpolate(U) . Sf on internal faces (linear weights, xyz-interleaved U and Sf)
__global__ void faceFluxUK(int nF, const int* __restrict__ own, const int* __restrict__ nei,
                           const double* __restrict__ w, const double* __restrict__ U, const double* __restrict__ Sf,
                           double* __restrict__ phi)
{
    const int f = blockIdx.x * kTpb + threadIdx.x;
    if (f >= nF) return;
    const int o = own[f], n = nei[f];
    const double wf = w[f], wn = 1.0 - wf;
    double s = 0.0;
    for (int k = 0; k < 3; ++k) s += (wf * U[3 * o + k] + wn * U[3 * n + k]) * Sf[3 * f + k];
    phi[f] = s;
}

// order-independent fingerprint of an array's bits: per workgroup the wrapping sum of mixed words
__global__ __launch_bounds__(kTpb) void bitsSumK(int n, const double* __restrict__ x, uint64_t* __restrict__ parts)
{
    __shared__ uint64_t lds[kTpb];
    uint64_t s = 0;
    for (int i = blockIdx.x * kTpb + threadIdx.x; i < n; i += gridDim.x * kTpb)
    {
        uint64_t b;
        memcpy(&b, &x[i], 8);
        s += mix64(b ^ ((uint64_t)i << 1));
    }
    lds[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        uint64_t t = 0;
        for (int k = 0; k < kTpb; ++k) t += lds[k];
        parts[blockIdx.x] = t;
    }
}

struct Box
{
    int nC = 0;
    std::vector<int> owner, neighbour, faceOffsets, faceVerts;
    std::vector<double> points;
};

uint64_t splitmix(uint64_t& s)
{
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// uniform box: internal faces by owner (+x, +y, +z neighbours), then the six walls; quads with their
// normal out of the owner; interior points jittered by +-jitter of the spacing. Built by threads over
// plane ranges: an interior point's jitter draws are the splitmix calls 3q, 3q+1, 3q+2 of the sequence
// (q its rank among the interior points, the state after call m being seed + (m+1)*gamma), and each
// thread's faces are its cells' in cell order, so the box is the same whatever the thread count.
Box makeBox(int nx, int ny, int nz, double lx, double ly, double lz, double jitter, uint64_t seed)
{
    Box b;
    b.nC = nx * ny * nz;
    const int px = nx + 1, py = ny + 1, pz = nz + 1;
    auto P = [&](int i, int j, int k) { return i + px * (j + py * k); };
    b.points.resize(3 * (size_t)px * py * pz);
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const int nT = (int)std::min<long long>(hw, std::max<long long>(1, (long long)b.nC / 65536));
    auto parallel = [&](int n, const auto& body) {   // body(lo, hi) over [0, n) in nT contiguous ranges
        if (nT <= 1)
        {
            body(0, n);
            return;
        }
        std::vector<std::thread> th;
        for (int t = 0; t < nT; ++t)
        {
            const int lo = (int)((long long)n * t / nT), hi = (int)((long long)n * (t + 1) / nT);
            th.emplace_back([&, lo, hi] { body(lo, hi); });
        }
        for (auto& x : th) x.join();
    };
    parallel(pz, [&](int k0, int k1) {
        for (int k = k0; k < k1; ++k)
            for (int j = 0; j < py; ++j)
                for (int i = 0; i < px; ++i)
                {
                    double x = lx * i / nx, y = ly * j / ny, z = lz * k / nz;
                    if (jitter > 0 && i > 0 && i < nx && j > 0 && j < ny && k > 0 && k < nz)
                    {
                        const uint64_t q = ((uint64_t)(k - 1) * (ny - 1) + (uint64_t)(j - 1)) * (nx - 1) + (i - 1);
                        uint64_t st = seed + 3 * q * 0x9e3779b97f4a7c15ull;
                        auto u = [&] { return 2.0 * ((double)(splitmix(st) >> 11) * (1.0 / 9007199254740992.0)) - 1.0; };
                        x += jitter * lx / nx * u();
                        y += jitter * ly / ny * u();
                        z += jitter * lz / nz * u();
                    }
                    const size_t p = (size_t)P(i, j, k);
                    b.points[3 * p] = x;
                    b.points[3 * p + 1] = y;
                    b.points[3 * p + 2] = z;
                }
    });
    // internal faces in cell order, each cell's +x, +y, +z face where it has one; a cell's first face
    // index in closed form (the faces of the cells before it, direction by direction), so the threads
    // write their cell ranges in place
    const size_t nI = (size_t)(nx - 1) * ny * nz + (size_t)nx * (ny - 1) * nz + (size_t)nx * ny * (nz - 1);
    const size_t nWall = 2 * ((size_t)nx * ny + (size_t)ny * nz + (size_t)nx * nz);
    b.owner.resize(nI + nWall);
    b.neighbour.resize(nI);
    b.faceVerts.resize(4 * (nI + nWall));
    auto put = [&](size_t f, int own, int a, int bb, int c, int d) {
        b.owner[f] = own;
        int* v = b.faceVerts.data() + 4 * f;
        v[0] = a;
        v[1] = bb;
        v[2] = c;
        v[3] = d;
    };
    parallel(b.nC, [&](int c0, int c1) {
        for (int c = c0; c < c1; ++c)
        {
            const int i = c % nx, j = (c / nx) % ny, k = c / (nx * ny);
            size_t f = ((size_t)j + (size_t)ny * k) * (nx - 1) + i;                          // +x faces before
            f += (size_t)k * nx * (ny - 1) + (size_t)j * nx + (j < ny - 1 ? i : 0);          // +y faces before
            f += (size_t)k * nx * ny + (k < nz - 1 ? (size_t)j * nx + i : 0);                // +z faces before
            if (i < nx - 1)
            {
                b.neighbour[f] = c + 1;
                put(f++, c, P(i + 1, j, k), P(i + 1, j + 1, k), P(i + 1, j + 1, k + 1), P(i + 1, j, k + 1));
            }
            if (j < ny - 1)
            {
                b.neighbour[f] = c + nx;
                put(f++, c, P(i, j + 1, k), P(i, j + 1, k + 1), P(i + 1, j + 1, k + 1), P(i + 1, j + 1, k));
            }
            if (k < nz - 1)
            {
                b.neighbour[f] = c + nx * ny;
                put(f++, c, P(i, j, k + 1), P(i + 1, j, k + 1), P(i + 1, j + 1, k + 1), P(i, j + 1, k + 1));
            }
        }
    });
    // the six walls, each in cell order
    auto cell = [&](int i, int j, int k) { return i + nx * (j + ny * k); };
    size_t f = nI;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            put(f++, cell(0, j, k), P(0, j, k), P(0, j, k + 1), P(0, j + 1, k + 1), P(0, j + 1, k));
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            put(f++, cell(nx - 1, j, k), P(nx, j, k), P(nx, j + 1, k), P(nx, j + 1, k + 1), P(nx, j, k + 1));
    for (int k = 0; k < nz; ++k)
        for (int i = 0; i < nx; ++i)
            put(f++, cell(i, 0, k), P(i, 0, k), P(i + 1, 0, k), P(i + 1, 0, k + 1), P(i, 0, k + 1));
    for (int k = 0; k < nz; ++k)
        for (int i = 0; i < nx; ++i)
            put(f++, cell(i, ny - 1, k), P(i, ny, k), P(i, ny, k + 1), P(i + 1, ny, k + 1), P(i + 1, ny, k));
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i)
            put(f++, cell(i, j, 0), P(i, j, 0), P(i, j + 1, 0), P(i + 1, j + 1, 0), P(i + 1, j, 0));
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i)
            put(f++, cell(i, j, nz - 1), P(i, j, nz), P(i + 1, j, nz), P(i + 1, j + 1, nz), P(i, j + 1, nz));
    b.faceOffsets.resize(b.owner.size() + 1);
    for (size_t f = 0; f <= b.owner.size(); ++f) b.faceOffsets[f] = (int)(4 * f);
    return b;
}

double now(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

void check(int rc, mfv_ctx ctx, const char* what)
{
    if (rc != MFV_OK) throw Error(rc, std::string("mfv: serve: ") + what + ": " + mfv_last_error(ctx));
}

} // namespace

// A request's system on the device: the box's mesh, geometry, matrix and hierarchy (Galerkin done), and
// its work vectors. Kept between requests when the server holds resident systems.
struct System
{
    mfv_mesh mesh = nullptr;
    mfv_amg amg = nullptr;
    int nC = 0, nI = 0, nB = 0;
    double bytes = 0;
    DBuf<double> Cf, Sf, magSf, C, V, w, dc, ndc, corr, bDelta, diag, upper, b, p, U, phi, zeroB;
    DBuf<uint64_t> parts;
    System() = default;
    System(const System&) = delete;
    System& operator=(const System&) = delete;
    ~System()
    {
        if (amg) mfv_amg_destroy(amg);
        if (mesh) mfv_mesh_destroy(mesh);
    }
};

// the box parameters a system is built from: equal keys, equal systems
struct BoxKey
{
    int nx, ny, nz;
    double lx, ly, lz, jitter;
    unsigned long long seed;
    bool operator<(const BoxKey& o) const
    {
        return std::tie(nx, ny, nz, lx, ly, lz, jitter, seed) < std::tie(o.nx, o.ny, o.nz, o.lx, o.ly, o.lz, o.jitter, o.seed);
    }
    bool operator==(const BoxKey& o) const { return !(*this < o) && !(o < *this); }
};
BoxKey keyOf(const mfv_serve_request& q) { return BoxKey{q.nx, q.ny, q.nz, q.lx, q.ly, q.lz, q.jitter, q.seed}; }

// the request's box on the device, its laplacian and its hierarchy, on the calling thread's context
std::unique_ptr<System> buildSystem(mfv_ctx ctx, const mfv_serve_request& q, int precision, double bytes)
{
    const Box box = makeBox(q.nx, q.ny, q.nz, q.lx, q.ly, q.lz, q.jitter, q.seed);
    auto S = std::make_unique<System>();
    const int nC = box.nC, nI = (int)box.neighbour.size(), nB = (int)box.owner.size() - nI;
    const int nF = (int)box.owner.size();
    S->nC = nC;
    S->nI = nI;
    S->nB = nB;
    S->bytes = bytes;
    hipStream_t st = static_cast<hipStream_t>(mfv_ctx_stream(ctx));
    const std::vector<int> bfc(box.owner.begin() + nI, box.owner.end());
    check(mfv_mesh_create(ctx, nC, nI, box.owner.data(), box.neighbour.data(), nB, bfc.data(), &S->mesh), ctx, "mesh");
    for (DBuf<double>* d : {&S->Cf, &S->Sf}) d->resize(3 * (size_t)nF);
    S->magSf.resize(nF);
    S->C.resize(3 * (size_t)nC);
    S->U.resize(3 * (size_t)nC);
    for (DBuf<double>* d : {&S->V, &S->diag, &S->b, &S->p}) d->resize(nC);
    for (DBuf<double>* d : {&S->w, &S->dc, &S->ndc, &S->upper, &S->phi}) d->resize(nI);
    S->corr.resize(3 * (size_t)nI);
    S->bDelta.resize(nB);
    S->zeroB.resize(nB);
    {
        DBuf<double> pts, gammaf;
        DBuf<int> off, verts;
        pts.upload(box.points, st);
        off.upload(box.faceOffsets, st);
        verts.upload(box.faceVerts, st);
        check(mfv_mesh_geometry(ctx, S->mesh, pts.p, off.p, verts.p, S->Cf.p, S->Sf.p, S->magSf.p, S->C.p, S->V.p,
                                S->w.p, S->dc.p, S->ndc.p, S->corr.p, S->bDelta.p),
              ctx, "geometry");
        // -laplacian(1, p) held positive (brae's sign), all walls zeroGradient, reference cell 0
        std::vector<double> minusOne(nI, -1.0);
        gammaf.upload(minusOne, st);
        check(mfv_laplacian_coeffs(ctx, S->mesh, S->dc.p, gammaf.p, S->magSf.p, S->diag.p, S->upper.p, nullptr), ctx,
              "laplacian");
        hipCheck(hipStreamSynchronize(st), "geometry sync");
    }
    hipCheck(hipMemsetAsync(S->b.p, 0, sizeof(double) * nC, st), "rhs clear");
    check(mfv_set_reference(ctx, nC, S->diag.p, S->b.p, 0, 0.0), ctx, "reference");
    std::vector<double> hw(nI);
    hipCheck(hipMemcpyAsync(hw.data(), S->magSf.p, sizeof(double) * nI, hipMemcpyDeviceToHost, st), "weights");
    hipCheck(hipStreamSynchronize(st), "weights sync");
    // levels of two composed pairwise passes and a V(2,1) cycle (the fastest measured cycle on the
    // 20 M-cell channel's pressure corrector)
    mfv_amg_controls ac{MFV_AGG_BRAE, 64, 2};
    check(mfv_amg_build(ctx, S->mesh, hw.data(), &ac, &S->amg), ctx, "amg");
    check(mfv_amg_galerkin(ctx, S->amg, S->diag.p, S->upper.p, nullptr, precision), ctx, "galerkin");
    hipCheck(hipMemsetAsync(S->zeroB.p, 0, sizeof(double) * nB, st), "phib clear");
    S->parts.resize(std::min(1024, blocks(nC)));
    return S;
}

// The serving loop's batcher: slots hand it the solves of their small requests (right-hand side ready,
// their stream idle), and it launches all that are waiting as one batch of whole solves on its own
// context (mfv_amg_pcg_solve_batch, one workgroup each); solves that arrive meanwhile go in the next
// launch. A solve's result is the same bits in any batch.
struct Batcher
{
    struct Job
    {
        mfv_batch_solve solve;
        mfv_perf perf{};
        int status = MFV_OK;
        std::string error;
        bool done = false;
    };
    std::mutex mu;
    std::condition_variable cv;
    std::vector<Job*> waiting;
    bool quit = false;
    mfv_ctx ctx = nullptr;
    long long launches = 0, solves = 0;

    // a slot's call: blocks until its solve is done
    void solve(Job& j)
    {
        std::unique_lock<std::mutex> lk(mu);
        waiting.push_back(&j);
        cv.notify_all();
        cv.wait(lk, [&] { return j.done; });
    }
    void run()
    {
        for (;;)
        {
            std::vector<Job*> batch;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [&] { return quit || !waiting.empty(); });
                if (waiting.empty()) return;
                batch.swap(waiting);
            }
            std::vector<mfv_batch_solve> in(batch.size());
            std::vector<mfv_perf> out(batch.size());
            for (size_t k = 0; k < batch.size(); ++k) in[k] = batch[k]->solve;
            const int rc = mfv_amg_pcg_solve_batch(ctx, (int)in.size(), in.data(), out.data());
            const std::string err = rc == MFV_OK ? std::string() : std::string(mfv_last_error(ctx));
            {
                std::lock_guard<std::mutex> g(mu);
                ++launches;
                solves += (long long)batch.size();
                for (size_t k = 0; k < batch.size(); ++k)
                {
                    batch[k]->perf = out[k];
                    batch[k]->status = rc;
                    batch[k]->error = err;
                    batch[k]->done = true;
                }
            }
            cv.notify_all();
        }
    }
};

// the request's steps on a built system: from p = 0, each step's right-hand side -div(U) of a seeded
// velocity, the solve warm-started from the previous step's p, the solution's fingerprint. With a
// batcher the solves go through it (the same cycle, one workgroup each), else through the per-grid
// kernels.
void solveSteps(mfv_ctx ctx, System& S, const mfv_serve_request& q, int precision, mfv_serve_result& res,
                Batcher* batcher = nullptr)
{
    const int nC = S.nC, nI = S.nI;
    hipStream_t st = static_cast<hipStream_t>(mfv_ctx_stream(ctx));
    mfv_pcg_controls pc;
    mfv_pcg_default_controls(&pc);
    pc.tolerance = q.tolerance;
    pc.precision = precision;
    pc.scaleCorrection = 1;   // scaled correction: a nonlinear cycle, so flexible CG
    pc.flexible = 1;
    pc.nPreSweeps = 2;
    pc.nPostSweeps = 1;
    const mfv_ldu A{S.diag.p, S.upper.p, S.upper.p};
    hipCheck(hipMemsetAsync(S.p.p, 0, sizeof(double) * nC, st), "p clear");
    const int *dOwn = nullptr, *dNei = nullptr;
    check(mfv_mesh_addressing(S.mesh, &dOwn, &dNei, nullptr, nullptr, nullptr), ctx, "addressing");
    const int nParts = (int)S.parts.n;
    std::vector<uint64_t> hp(nParts);
    uint64_t fp = 0;
    res.totalIterations = 0;
    for (int s = 0; s < q.nSteps; ++s)
    {
        velocityK<<<blocks(3LL * nC), kTpb, 0, st>>>(nC, q.seed * 1000003ull + (uint64_t)s, S.U.p);
        faceFluxUK<<<blocks(nI), kTpb, 0, st>>>(nI, dOwn, dNei, S.w.p, S.U.p, S.Sf.p, S.phi.p);
        hipCheck(hipMemsetAsync(S.b.p, 0, sizeof(double) * nC, st), "rhs clear");
        check(mfv_div_source(ctx, S.mesh, S.phi.p, S.zeroB.p, S.V.p, -1.0, S.b.p), ctx, "div");
        mfv_perf perf{};
        if (batcher)
        {
            hipCheck(hipStreamSynchronize(st), "rhs sync");
            Batcher::Job j;
            j.solve = mfv_batch_solve{S.mesh, S.amg, S.diag.p, S.b.p, S.p.p, pc};
            batcher->solve(j);
            if (j.status != MFV_OK) throw Error(j.status, "mfv: serve: batched solve: " + j.error);
            perf = j.perf;
        }
        else
            check(mfv_amg_pcg_solve(ctx, S.mesh, S.amg, &A, S.b.p, S.p.p, nullptr, &pc, &perf), ctx, "solve");
        res.totalIterations += perf.nIterations;
        res.finalResidual = perf.finalResidual;
        bitsSumK<<<nParts, kTpb, 0, st>>>(nC, S.p.p, S.parts.p);
        hipCheck(hipMemcpyAsync(hp.data(), S.parts.p, sizeof(uint64_t) * nParts, hipMemcpyDeviceToHost, st), "fp");
        hipCheck(hipStreamSynchronize(st), "fp sync");
        uint64_t t = 0;
        for (uint64_t v : hp) t += v;
        fp = fp * 0x100000001b3ull ^ t;
    }
    res.fingerprint = fp;
    res.nCells = nC;
    res.nSteps = q.nSteps;
}

} // namespace mfv

using namespace mfv;

struct mfv_server_s
{
    int device = 0, nSlots = 1, precision = MFV_PREC_FP32;
    double budget = 0;
    double bytesPerCell = 600.0;
    double resident = 0;   // bytes of the budget that hold systems between requests (0: none kept)
    int batchCells = 0;    // requests of at most this many cells solve in batches (0: none)
};

extern "C" {

int mfv_server_create(int device, int nSlots, double hbmBudgetBytes, int precision, mfv_server* out)
{
    if (!out || nSlots < 1 || nSlots > 64 || hbmBudgetBytes <= 0) return MFV_ERR_INVALID;
    if (precision != MFV_PREC_FP64 && precision != MFV_PREC_FP32 && precision != MFV_PREC_FP16)
        return MFV_ERR_INVALID;
    mfv_server s = new (std::nothrow) mfv_server_s();
    if (!s) return MFV_ERR_ALLOC;
    s->device = device;
    s->nSlots = nSlots;
    s->budget = hbmBudgetBytes;
    s->precision = precision;
    *out = s;
    return MFV_OK;
}

int mfv_server_set_resident(mfv_server s, double residentBytes)
{
    if (!s || !(residentBytes >= 0) || residentBytes >= s->budget) return MFV_ERR_INVALID;
    s->resident = residentBytes;
    return MFV_OK;
}

int mfv_server_set_batch(mfv_server s, int maxCells)
{
    if (!s || maxCells < 0) return MFV_ERR_INVALID;
    s->batchCells = maxCells;
    return MFV_OK;
}

int mfv_server_destroy(mfv_server s)
{
    delete s;
    return MFV_OK;
}

double mfv_server_request_bytes(mfv_server s, const mfv_serve_request* q)
{
    return s && q ? s->bytesPerCell * (double)q->nx * q->ny * q->nz : 0.0;
}

int mfv_server_run(mfv_server s, const mfv_serve_request* reqs, int nReq, mfv_serve_result* results)
{
    if (!s || nReq < 0 || (nReq && (!reqs || !results))) return MFV_ERR_INVALID;
    // the resident systems' share of the budget is held back from admission for the whole run
    const double admit = s->budget - s->resident;
    for (int i = 0; i < nReq; ++i)
    {
        const mfv_serve_request& q = reqs[i];
        if (q.nx < 2 || q.ny < 2 || q.nz < 2 || q.nSteps < 1 || q.tolerance <= 0 ||
            (double)q.nx * q.ny * q.nz > 1.0e8 || mfv_server_request_bytes(s, &q) > admit)
            return MFV_ERR_INVALID;
        if (i && q.arrival < reqs[i - 1].arrival) return MFV_ERR_INVALID;   // arrivals in order
    }
    const auto t0 = std::chrono::steady_clock::now();
    std::mutex mu;
    std::condition_variable cv;
    int next = 0;            // first request not yet admitted (FIFO)
    double inflight = 0;     // admitted bytes
    int firstError = MFV_OK;
    std::string errorText;
    // resident systems: each slot keeps its own (they live on its context); `owner` says which slot
    // holds a box, `idle` which slots wait for work, so a request whose box is resident on an idle slot
    // goes to that slot
    std::map<BoxKey, int> owner;
    std::vector<char> idle(s->nSlots, 0);
    const double perSlot = s->resident / s->nSlots;
    // the batcher of the small requests' solves, on a context of its own
    Batcher batcher;
    std::thread batchThread;
    if (s->batchCells > 0)
    {
        if (mfv_ctx_create(s->device, nullptr, &batcher.ctx) != MFV_OK) return MFV_ERR_HIP;
        batchThread = std::thread([&] { batcher.run(); });
    }
    auto worker = [&](int slot) {
        mfv_ctx ctx = nullptr;
        if (mfv_ctx_create(s->device, nullptr, &ctx) != MFV_OK)
        {
            std::lock_guard<std::mutex> g(mu);
            if (firstError == MFV_OK) firstError = MFV_ERR_HIP;
            next = nReq;
            cv.notify_all();
            return;
        }
        struct Entry
        {
            BoxKey key;
            std::unique_ptr<System> sys;
        };
        std::list<Entry> cache;   // most recently used first
        double cached = 0;
        for (;;)
        {
            int i;
            double bytes;
            {
                std::unique_lock<std::mutex> lk(mu);
                idle[slot] = 1;
                // admit the queue's head when it has arrived, fits in what the budget has left and is not
                // resident on another idle slot
                for (;;)
                {
                    if (next >= nReq) break;
                    const double wait = reqs[next].arrival - now(t0);
                    bytes = mfv_server_request_bytes(s, &reqs[next]);
                    if (wait <= 0 && inflight + bytes <= admit)
                    {
                        const auto o = owner.find(keyOf(reqs[next]));
                        if (o == owner.end() || o->second == slot || !idle[o->second]) break;
                        cv.notify_all();   // the slot that holds it takes it
                        cv.wait(lk);
                        continue;
                    }
                    if (wait > 0)
                        cv.wait_for(lk, std::chrono::duration<double>(wait));
                    else
                        cv.wait(lk);
                }
                if (next >= nReq) break;
                i = next++;
                inflight += bytes;
                idle[slot] = 0;
                results[i].admitted = now(t0);
                results[i].arrival = reqs[i].arrival;
            }
            cv.notify_all();
            int rc = MFV_OK;
            std::string err;
            try
            {
                const BoxKey key = keyOf(reqs[i]);
                std::unique_ptr<System> sys;
                for (auto it = cache.begin(); it != cache.end(); ++it)
                    if (it->key == key)
                    {
                        sys = std::move(it->sys);
                        cached -= sys->bytes;
                        cache.erase(it);
                        break;
                    }
                results[i].resident = sys ? 1 : 0;
                if (!sys) sys = buildSystem(ctx, reqs[i], s->precision, bytes);
                results[i].built = now(t0);
                const bool small = s->batchCells > 0 && (double)reqs[i].nx * reqs[i].ny * reqs[i].nz <= s->batchCells;
                results[i].batched = small ? 1 : 0;
                solveSteps(ctx, *sys, reqs[i], s->precision, results[i], small ? &batcher : nullptr);
                if (sys->bytes <= perSlot)
                {
                    // keep it, the least recently used going first when the slot's share is full
                    cached += sys->bytes;
                    cache.push_front(Entry{key, std::move(sys)});
                    std::lock_guard<std::mutex> g(mu);
                    owner[key] = slot;
                    while (cached > perSlot)
                    {
                        Entry& e = cache.back();
                        cached -= e.sys->bytes;
                        const auto o = owner.find(e.key);
                        if (o != owner.end() && o->second == slot) owner.erase(o);
                        cache.pop_back();
                    }
                }
            }
            catch (const Error& e)
            {
                rc = e.code;
                err = e.what();
            }
            catch (const std::exception& e)
            {
                rc = MFV_ERR_INVALID;
                err = e.what();
            }
            {
                std::lock_guard<std::mutex> g(mu);
                results[i].finished = now(t0);
                results[i].status = rc;
                inflight -= bytes;
                if (rc != MFV_OK && firstError == MFV_OK)
                {
                    firstError = rc;
                    errorText = err;
                }
            }
            cv.notify_all();
        }
        {
            std::lock_guard<std::mutex> g(mu);
            for (const Entry& e : cache)
            {
                const auto o = owner.find(e.key);
                if (o != owner.end() && o->second == slot) owner.erase(o);
            }
            idle[slot] = 0;
        }
        cv.notify_all();
        cache.clear();   // the systems go before their context
        mfv_ctx_destroy(ctx);
    };
    std::vector<std::thread> threads;
    for (int k = 0; k < s->nSlots; ++k) threads.emplace_back(worker, k);
    for (auto& t : threads) t.join();
    if (batchThread.joinable())
    {
        {
            std::lock_guard<std::mutex> g(batcher.mu);
            batcher.quit = true;
        }
        batcher.cv.notify_all();
        batchThread.join();
        mfv_ctx_destroy(batcher.ctx);
    }
    return firstError;
}

} // extern "C"
