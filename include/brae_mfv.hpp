/*
 * brae_mfv.hpp -- header-only C++ layer over the mfv C ABI (mfv.h) that re-creates the signatures
 * brae's callers use for the pressure assembly and solve path, in namespace brae, so a call site such
 * as the SIMPLE pressure block (src/applications/solvers/simpleFoam/simpleFoam.cu:343-381) compiles
 * against it unchanged apart from the include. Host code only; link with libmfv.so.
 *
 * Conventions kept from the reference:
 *   - DeviceBuffer<T> is RAII device memory; callees resize the caller's outputs
 *     (device_fvm.cu:313 `upper.resize(nIf)`), views are borrowed (src/cuda/device_buffer.cuh:716-760);
 *   - errors throw std::runtime_error("brae hip: ...") as cudaCheck does (device_buffer.cuh:25-28);
 *     a solve that does not converge is not an error (nIterations == maxIter);
 *   - one solve at a time per host thread, each host thread on its own stream (the reference runs on
 *     cudaStreamPerThread): threadContext() gives every thread its own mfv context;
 *   - a symmetric matrix passes lower == upper (device_amg.cu:1703-1710).
 * Deliberate differences: vector fields (gradients, Sf) are xyz-interleaved arrays of 3*n doubles; the
 * smoothed-aggregation switch of buildAMG and the V-cycle graph capture are refused / ignored by name.
 */
#pragma once

#include "mfv.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace brae {

using scalar = double;
using label = int;

inline void mfvCheck(int rc, mfv_ctx ctx, const char* what)
{
    if (rc != MFV_OK)
        throw std::runtime_error(std::string("brae hip: ") + what + ": " + mfv_last_error(ctx));
}

inline void hipCheckOrThrow(hipError_t e, const char* what)
{
    if (e != hipSuccess) throw std::runtime_error(std::string("brae hip: ") + what + ": " + hipGetErrorString(e));
}

// The calling thread's context: its own stream and reduction scratch, created on first use on the
// current device and released when the thread ends.
inline mfv_ctx threadContext()
{
    struct Holder
    {
        mfv_ctx c = nullptr;
        ~Holder()
        {
            if (c) mfv_ctx_destroy(c);
        }
    };
    thread_local Holder h;
    if (!h.c)
    {
        int dev = 0;
        hipCheckOrThrow(hipGetDevice(&dev), "hipGetDevice");
        mfvCheck(mfv_ctx_create(dev, nullptr, &h.c), nullptr, "context");
    }
    return h.c;
}

inline hipStream_t threadStream() { return static_cast<hipStream_t>(mfv_ctx_stream(threadContext())); }

// RAII device array (DeviceBuffer, src/cuda/device_buffer.cuh:716-760)
template <typename T>
class DeviceBuffer
{
public:
    DeviceBuffer() = default;
    explicit DeviceBuffer(std::size_t n) { resize(n); }
    explicit DeviceBuffer(const std::vector<T>& h) { upload(h); }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    DeviceBuffer(DeviceBuffer&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
    DeviceBuffer& operator=(DeviceBuffer&& o) noexcept
    {
        if (this != &o)
        {
            release();
            p_ = o.p_;
            n_ = o.n_;
            o.p_ = nullptr;
            o.n_ = 0;
        }
        return *this;
    }
    ~DeviceBuffer() { release(); }

    void resize(std::size_t n)
    {
        if (n == n_) return;
        release();
        if (n) hipCheckOrThrow(hipMalloc(&p_, n * sizeof(T)), "hipMalloc");
        n_ = n;
    }
    void upload(const std::vector<T>& h)
    {
        resize(h.size());
        if (n_)
        {
            hipCheckOrThrow(hipMemcpyAsync(p_, h.data(), n_ * sizeof(T), hipMemcpyHostToDevice, threadStream()),
                            "upload");
            hipCheckOrThrow(hipStreamSynchronize(threadStream()), "upload sync");
        }
    }
    std::vector<T> download() const
    {
        std::vector<T> h(n_);
        if (n_)
        {
            hipCheckOrThrow(hipMemcpyAsync(h.data(), p_, n_ * sizeof(T), hipMemcpyDeviceToHost, threadStream()),
                            "download");
            hipCheckOrThrow(hipStreamSynchronize(threadStream()), "download sync");
        }
        return h;
    }
    void zero()
    {
        if (n_) hipCheckOrThrow(hipMemsetAsync(p_, 0, n_ * sizeof(T), threadStream()), "zero");
    }
    T* data() { return p_; }
    const T* data() const { return p_; }
    std::size_t size() const { return n_; }

private:
    void release()
    {
        if (p_) (void)hipFree(p_);
        p_ = nullptr;
        n_ = 0;
    }
    T* p_ = nullptr;
    std::size_t n_ = 0;
};

// The mesh as the path sees it (DeviceMesh, src/cuda/device_mesh.cuh): the LDU addressing handle and
// the geometry the assembly reads. Built from host addressing and points+faces (geometry on the device).
struct DeviceMesh
{
    mfv_mesh mesh = nullptr;
    int nCells = 0, nInternalFaces = 0, nBoundaryFaces = 0;
    DeviceBuffer<scalar> Cf, Sf, magSf, C, V;               // per face (internal then boundary) / per cell
    DeviceBuffer<scalar> weights, deltaCoeffs, nonOrthDeltaCoeffs, corrVec;   // per internal face
    DeviceBuffer<scalar> bDeltaCoeffs;                        // per boundary face
    DeviceBuffer<scalar> magSfInternal, magSfBoundary;

    DeviceMesh() = default;
    DeviceMesh(const DeviceMesh&) = delete;
    DeviceMesh& operator=(const DeviceMesh&) = delete;
    ~DeviceMesh()
    {
        if (mesh) mfv_mesh_destroy(mesh);
    }

    // owner: every face (internal faces first, upper-triangular; then boundary faces by patch);
    // neighbour: internal faces; faces: polygons as offsets [nFaces+1] into vertex lists; points xyz.
    void build(int nC, const std::vector<label>& owner, const std::vector<label>& neighbour,
               const std::vector<scalar>& points, const std::vector<label>& faceOffsets,
               const std::vector<label>& faceVerts)
    {
        mfv_ctx ctx = threadContext();
        nCells = nC;
        nInternalFaces = static_cast<int>(neighbour.size());
        nBoundaryFaces = static_cast<int>(owner.size()) - nInternalFaces;
        const std::vector<label> bfc(owner.begin() + nInternalFaces, owner.end());
        mfvCheck(mfv_mesh_create(ctx, nCells, nInternalFaces, owner.data(), neighbour.data(), nBoundaryFaces,
                                 bfc.data(), &mesh),
                 ctx, "mesh");
        const std::size_t nF = owner.size();
        DeviceBuffer<scalar> dPts(points);
        DeviceBuffer<label> dOff(faceOffsets), dVerts(faceVerts);
        Cf.resize(3 * nF);
        Sf.resize(3 * nF);
        magSf.resize(nF);
        C.resize(3 * (std::size_t)nCells);
        V.resize(nCells);
        weights.resize(nInternalFaces);
        deltaCoeffs.resize(nInternalFaces);
        nonOrthDeltaCoeffs.resize(nInternalFaces);
        corrVec.resize(3 * (std::size_t)nInternalFaces);
        bDeltaCoeffs.resize(nBoundaryFaces);
        mfvCheck(mfv_mesh_geometry(ctx, mesh, dPts.data(), dOff.data(), dVerts.data(), Cf.data(), Sf.data(),
                                   magSf.data(), C.data(), V.data(), weights.data(), deltaCoeffs.data(),
                                   nonOrthDeltaCoeffs.data(), corrVec.data(), bDeltaCoeffs.data()),
                 ctx, "geometry");
        magSfInternal.resize(nInternalFaces);
        magSfBoundary.resize(nBoundaryFaces);
        if (nInternalFaces)
            hipCheckOrThrow(hipMemcpyAsync(magSfInternal.data(), magSf.data(), nInternalFaces * sizeof(scalar),
                                           hipMemcpyDeviceToDevice, threadStream()), "magSf");
        if (nBoundaryFaces)
            hipCheckOrThrow(hipMemcpyAsync(magSfBoundary.data(), magSf.data() + nInternalFaces,
                                           nBoundaryFaces * sizeof(scalar), hipMemcpyDeviceToDevice, threadStream()),
                            "magSf");
        mfvCheck(mfv_ctx_sync(ctx), ctx, "mesh sync");
    }
};

// The value half of DeviceLduView (src/matrices/lduMatrix/lduMatrix/device_ldu.cuh:20-66) plus the
// mesh it lives on.
struct DeviceLduView
{
    int nCells = 0, nInternalFaces = 0;
    const scalar* diag = nullptr;
    const scalar* upper = nullptr;
    const scalar* lower = nullptr;
    mfv_mesh mesh = nullptr;
    mfv_ldu ldu() const { return mfv_ldu{diag, upper, lower}; }
};

inline DeviceLduView lduView(const DeviceMesh& dm, const DeviceBuffer<scalar>& diag,
                             const DeviceBuffer<scalar>& upper, const DeviceBuffer<scalar>* lower = nullptr)
{
    return DeviceLduView{dm.nCells, dm.nInternalFaces, diag.data(), upper.data(),
                         lower ? lower->data() : upper.data(), dm.mesh};
}

struct DeviceSolverPerf
{
    scalar initialResidual = 0, finalResidual = 0;
    int nIterations = 0;
};

// ---------------- assembly (src/cuda/device_mesh.cuh:495-612, device_fvm.cu) ----------------

// deviceLaplacianCoeffs (device_fvm.cu:304): upper = lower = dc*gammaf*|Sf|, diag = -sum in face
// order; nonOrth selects nonOrthDeltaCoeffs
inline void deviceLaplacianCoeffs(const DeviceMesh& dm, const DeviceBuffer<scalar>& gammafInt,
                                  DeviceBuffer<scalar>& diag, DeviceBuffer<scalar>& upper,
                                  DeviceBuffer<scalar>& lower, bool nonOrth = false)
{
    mfv_ctx ctx = threadContext();
    diag.resize(dm.nCells);
    upper.resize(dm.nInternalFaces);
    lower.resize(dm.nInternalFaces);
    const DeviceBuffer<scalar>& dc = nonOrth ? dm.nonOrthDeltaCoeffs : dm.deltaCoeffs;
    mfvCheck(mfv_laplacian_coeffs(ctx, dm.mesh, dc.data(), gammafInt.data(), dm.magSfInternal.data(), diag.data(),
                                  upper.data(), lower.data()),
             ctx, "deviceLaplacianCoeffs");
}

// deviceLaplacianCoeffsSymmetric (device_fvm.cu:349): the off-diagonal held once
inline void deviceLaplacianCoeffsSymmetric(const DeviceMesh& dm, const DeviceBuffer<scalar>& gammafInt,
                                           DeviceBuffer<scalar>& diag, DeviceBuffer<scalar>& upper,
                                           bool nonOrth = false)
{
    mfv_ctx ctx = threadContext();
    diag.resize(dm.nCells);
    upper.resize(dm.nInternalFaces);
    const DeviceBuffer<scalar>& dc = nonOrth ? dm.nonOrthDeltaCoeffs : dm.deltaCoeffs;
    mfvCheck(mfv_laplacian_coeffs(ctx, dm.mesh, dc.data(), gammafInt.data(), dm.magSfInternal.data(), diag.data(),
                                  upper.data(), nullptr),
             ctx, "deviceLaplacianCoeffsSymmetric");
}

// deviceDivUpwindCoeffs (device_fvm.cu:646-662)
inline void deviceDivUpwindCoeffs(const DeviceMesh& dm, const DeviceBuffer<scalar>& phiInt,
                                  DeviceBuffer<scalar>& diag, DeviceBuffer<scalar>& upper,
                                  DeviceBuffer<scalar>& lower)
{
    mfv_ctx ctx = threadContext();
    diag.resize(dm.nCells);
    upper.resize(dm.nInternalFaces);
    lower.resize(dm.nInternalFaces);
    mfvCheck(mfv_div_upwind_coeffs(ctx, dm.mesh, phiInt.data(), diag.data(), upper.data(), lower.data()), ctx,
             "deviceDivUpwindCoeffs");
}

// deviceDivCentralCoeffs (device_fvm.cu:628-643): Gauss linear convection on the mesh's weights
inline void deviceDivCentralCoeffs(const DeviceMesh& dm, const DeviceBuffer<scalar>& phiInt,
                                   DeviceBuffer<scalar>& diag, DeviceBuffer<scalar>& upper,
                                   DeviceBuffer<scalar>& lower)
{
    mfv_ctx ctx = threadContext();
    diag.resize(dm.nCells);
    upper.resize(dm.nInternalFaces);
    lower.resize(dm.nInternalFaces);
    mfvCheck(mfv_div_central_coeffs(ctx, dm.mesh, phiInt.data(), dm.weights.data(), diag.data(), upper.data(),
                                    lower.data()),
             ctx, "deviceDivCentralCoeffs");
}

// deviceFold (device_fvm.cu:895-912): diagC = rawDiag + iC, b = source + bC (flattened boundary faces)
inline void deviceFold(const DeviceMesh& dm, const DeviceBuffer<scalar>& rawDiag, const DeviceBuffer<scalar>& source,
                       const DeviceBuffer<scalar>& iC, const DeviceBuffer<scalar>& bC, DeviceBuffer<scalar>& diagC,
                       DeviceBuffer<scalar>& b)
{
    mfv_ctx ctx = threadContext();
    diagC.resize(dm.nCells);
    b.resize(dm.nCells);
    mfvCheck(mfv_fold(ctx, dm.mesh, rawDiag.data(), source.data(), iC.data(), bC.data(), diagC.data(), b.data()),
             ctx, "deviceFold");
}

// deviceSetReference (src/applications/solvers/simpleFoam/device_simple.cuh:104-106)
inline void deviceSetReference(DeviceBuffer<scalar>& diag, DeviceBuffer<scalar>& b, label refCell, scalar refValue)
{
    mfv_ctx ctx = threadContext();
    if (refCell >= 0 && (static_cast<std::size_t>(refCell) >= diag.size() || static_cast<std::size_t>(refCell) >= b.size()))
        throw std::runtime_error("brae hip: deviceSetReference: refCell " + std::to_string(refCell) +
                                 " is outside the diagonal or the source");
    const int n = static_cast<int>(std::min(diag.size(), b.size()));
    mfvCheck(mfv_set_reference(ctx, n, diag.data(), b.data(), refCell, refValue), ctx, "deviceSetReference");
}

// deviceMatrixFluxInternal (device_simple.cuh:51-52): flux = upper*p[nei] - lower*p[own]
inline void deviceMatrixFluxInternal(const DeviceLduView& A, const DeviceBuffer<scalar>& p,
                                     DeviceBuffer<scalar>& fluxInt)
{
    mfv_ctx ctx = threadContext();
    fluxInt.resize(A.nInternalFaces);
    const mfv_ldu l = A.ldu();
    mfvCheck(mfv_face_flux(ctx, A.mesh, &l, p.data(), fluxInt.data()), ctx, "deviceMatrixFluxInternal");
}

// ---- the SIMPLE pressure predictor's momentum glue (pressurePredictor, pEqn.cu:94-335) ----

// A()'s numerator (pEqn.cu:117-138): D = diag + cmptAv(internalCoeffs), iC[3] the momentum matrix's
// internalCoeffs per boundary face and component
inline void deviceFoldBoundaryDiag(const DeviceMesh& dm, const DeviceBuffer<scalar>& diag,
                                   const DeviceBuffer<scalar>* iC, DeviceBuffer<scalar>& D)
{
    mfv_ctx ctx = threadContext();
    D.resize(dm.nCells);
    mfvCheck(mfv_fold_boundary_diag(ctx, dm.mesh, diag.data(), iC[0].data(), iC[1].data(), iC[2].data(), D.data()),
             ctx, "deviceFoldBoundaryDiag");
}

// deviceReciprocalV (src/cuda/device_simple.cuh:26): rAU = 1/(diagC/V), diagC already folded
inline void deviceReciprocalV(const DeviceMesh& dm, const DeviceBuffer<scalar>& diagC, DeviceBuffer<scalar>& rAU)
{
    mfv_ctx ctx = threadContext();
    rAU.resize(dm.nCells);
    mfvCheck(mfv_reciprocal_v(ctx, dm.mesh, diagC.data(), dm.V.data(), rAU.data()), ctx, "deviceReciprocalV");
}

// bdDiagK = cmptAv(iC) - iC_k (pEqn.cu:146-149)
inline void deviceBoundaryDiagTerm(const DeviceMesh& dm, const DeviceBuffer<scalar>* iC, int k,
                                   DeviceBuffer<scalar>& bdDiagK)
{
    mfv_ctx ctx = threadContext();
    bdDiagK.resize(dm.nBoundaryFaces);
    mfvCheck(mfv_boundary_diag_term(ctx, dm.mesh, iC[0].data(), iC[1].data(), iC[2].data(), k, bdDiagK.data()), ctx,
             "deviceBoundaryDiagTerm");
}

// deviceMatrixH (src/cuda/device_simple.cuh:14-23): H = (source + lduMatrix::H(psi) + bdDiag*psi + bdSrc)/V;
// valid = false: a component the mesh's empty patches knock out (zeros)
inline void deviceMatrixH(const DeviceLduView& A, const DeviceMesh& dm, const DeviceBuffer<scalar>& psiK,
                          const DeviceBuffer<scalar>& sourceK, const DeviceBuffer<scalar>& bdDiagK,
                          const DeviceBuffer<scalar>& bdSrcK, DeviceBuffer<scalar>& Hk, bool valid = true)
{
    mfv_ctx ctx = threadContext();
    Hk.resize(dm.nCells);
    const mfv_ldu l = A.ldu();
    mfvCheck(mfv_matrix_h(ctx, dm.mesh, &l, psiK.data(), sourceK.data(), bdDiagK.data(), bdSrcK.data(), dm.V.data(),
                          nullptr, valid ? 1 : 0, Hk.data()),
             ctx, "deviceMatrixH");
}

// HbyA_k = rAU*H_k in one pass (pEqn.cu:150-185: deviceMatrixH, then deviceHadamard)
inline void deviceMatrixHbyA(const DeviceLduView& A, const DeviceMesh& dm, const DeviceBuffer<scalar>& psiK,
                             const DeviceBuffer<scalar>& sourceK, const DeviceBuffer<scalar>& bdDiagK,
                             const DeviceBuffer<scalar>& bdSrcK, const DeviceBuffer<scalar>& rAU,
                             DeviceBuffer<scalar>& HbyAk, bool valid = true)
{
    mfv_ctx ctx = threadContext();
    HbyAk.resize(dm.nCells);
    const mfv_ldu l = A.ldu();
    mfvCheck(mfv_matrix_h(ctx, dm.mesh, &l, psiK.data(), sourceK.data(), bdDiagK.data(), bdSrcK.data(), dm.V.data(),
                          rAU.data(), valid ? 1 : 0, HbyAk.data()),
             ctx, "deviceMatrixHbyA");
}

// HbyA of the three components in one pass over the matrix (no counterpart of its own: pEqn.cu:143-185's three
// deviceMatrixH + deviceHadamard calls); U, source, bdDiag, bdSrc, HbyA: arrays of three; valid may be null
inline void deviceMatrixHbyA3(const DeviceLduView& A, const DeviceMesh& dm, const DeviceBuffer<scalar>* U,
                              const DeviceBuffer<scalar>* source, const DeviceBuffer<scalar>* bdDiag,
                              const DeviceBuffer<scalar>* bdSrc, const DeviceBuffer<scalar>& rAU,
                              DeviceBuffer<scalar>* HbyA, const bool* valid = nullptr)
{
    mfv_ctx ctx = threadContext();
    const double* psi[3];
    const double* src[3];
    const double* bd[3];
    const double* bs[3];
    double* out[3];
    int v[3];
    for (int k = 0; k < 3; ++k)
    {
        HbyA[k].resize(dm.nCells);
        psi[k] = U[k].data();
        src[k] = source[k].data();
        bd[k] = bdDiag[k].data();
        bs[k] = bdSrc[k].data();
        out[k] = HbyA[k].data();
        v[k] = valid ? (valid[k] ? 1 : 0) : 1;
    }
    const mfv_ldu l = A.ldu();
    mfvCheck(mfv_matrix_h3(ctx, dm.mesh, &l, psi, src, bd, bs, dm.V.data(), rAU.data(), v, out), ctx,
             "deviceMatrixHbyA3");
}

// constrainHbyA (pEqn.cu:187-203): HbyAb = Ub where takeU (U not assignable), the owner cell's HbyA elsewhere
inline void deviceConstrainHbyA(const DeviceMesh& dm, const DeviceBuffer<int>& takeU, const DeviceBuffer<scalar>& Ub,
                                const DeviceBuffer<scalar>& HbyA, DeviceBuffer<scalar>& HbyAb)
{
    mfv_ctx ctx = threadContext();
    HbyAb.resize(dm.nBoundaryFaces);
    mfvCheck(mfv_constrain_hbya(ctx, dm.mesh, takeU.data(), Ub.data(), HbyA.data(), HbyAb.data()), ctx,
             "deviceConstrainHbyA");
}

// deviceRelaxDiag (src/cuda/device_simple.cuh:118-120): fvMatrix::relax's diagonal for uncoupled patches,
// relaxedDiag = max(|diag + cmptMagMax(iC)|, sumOff)/alpha - cmptMin(iC) and delta = relaxedDiag - diag.
// iCmaxMag/iCmin nullptr: |iCbnd| and iCbnd.
inline void deviceRelaxDiag(const DeviceLduView& A, const DeviceMesh& dm, const DeviceBuffer<scalar>& iCbnd, scalar alpha,
                            DeviceBuffer<scalar>& relaxedDiag, DeviceBuffer<scalar>& delta,
                            const scalar* cycSumOff = nullptr, const scalar* iCmaxMag = nullptr,
                            const scalar* iCmin = nullptr)
{
    mfv_ctx ctx = threadContext();
    relaxedDiag.resize(dm.nCells);
    delta.resize(dm.nCells);
    const mfv_ldu l = A.ldu();
    mfvCheck(mfv_relax_diag(ctx, dm.mesh, &l, iCbnd.data(), alpha, cycSumOff, iCmaxMag, iCmin, relaxedDiag.data(),
                            delta.data(), 0, nullptr, nullptr),
             ctx, "deviceRelaxDiag");
}

// UEqn.relax() in one pass (UEqn.cu:724-760: deviceRelaxDiag, then source_k += delta*U_k per component)
inline void deviceRelaxMomentum(const DeviceLduView& A, const DeviceMesh& dm, const DeviceBuffer<scalar>* iC,
                                scalar alpha, const DeviceBuffer<scalar>* U, DeviceBuffer<scalar>* source,
                                DeviceBuffer<scalar>& relaxedDiag, DeviceBuffer<scalar>& delta)
{
    mfv_ctx ctx = threadContext();
    DeviceBuffer<scalar> iCmaxMag(dm.nBoundaryFaces), iCmin(dm.nBoundaryFaces);
    mfvCheck(mfv_cmpt_max_mag3(ctx, dm.nBoundaryFaces, iC[0].data(), iC[1].data(), iC[2].data(), iCmaxMag.data()),
             ctx, "deviceCmptMaxMag3");
    mfvCheck(mfv_cmpt_min3(ctx, dm.nBoundaryFaces, iC[0].data(), iC[1].data(), iC[2].data(), iCmin.data()), ctx,
             "deviceCmptMin3");
    relaxedDiag.resize(dm.nCells);
    delta.resize(dm.nCells);
    const mfv_ldu l = A.ldu();
    const scalar* psi[3] = {U[0].data(), U[1].data(), U[2].data()};
    scalar* src[3] = {source[0].data(), source[1].data(), source[2].data()};
    mfvCheck(mfv_relax_diag(ctx, dm.mesh, &l, iC[0].data(), alpha, nullptr, iCmaxMag.data(), iCmin.data(),
                            relaxedDiag.data(), delta.data(), 3, psi, src),
             ctx, "deviceRelaxMomentum");
}

// deviceLinearUpwindCorr (src/cuda/device_simple.cuh:34-36): linearUpwind's deferred correction of
// div(phi, U_k) per cell; the caller does source -= corrSource. grad is xyz per cell, the layout
// deviceGaussGrad writes here (the reference takes three arrays gx, gy, gz).
inline void deviceLinearUpwindCorr(const DeviceMesh& dm, const DeviceBuffer<scalar>& phiInt,
                                   const DeviceBuffer<scalar>& grad, DeviceBuffer<scalar>& corrSource)
{
    mfv_ctx ctx = threadContext();
    corrSource.resize(dm.nCells);
    const scalar* g[1] = {grad.data()};
    scalar* out[1] = {corrSource.data()};
    mfvCheck(mfv_linear_upwind_corr(ctx, dm.mesh, phiInt.data(), dm.Cf.data(), dm.C.data(), 1, g, out), ctx,
             "deviceLinearUpwindCorr");
}

// the three components' corrections in one pass over the faces (three deviceLinearUpwindCorr calls' values)
inline void deviceLinearUpwindCorr3(const DeviceMesh& dm, const DeviceBuffer<scalar>& phiInt,
                                    const DeviceBuffer<scalar>* grad, DeviceBuffer<scalar>* corrSource)
{
    mfv_ctx ctx = threadContext();
    const scalar* g[3] = {grad[0].data(), grad[1].data(), grad[2].data()};
    scalar* out[3];
    for (int k = 0; k < 3; ++k)
    {
        corrSource[k].resize(dm.nCells);
        out[k] = corrSource[k].data();
    }
    mfvCheck(mfv_linear_upwind_corr(ctx, dm.mesh, phiInt.data(), dm.Cf.data(), dm.C.data(), 3, g, out), ctx,
             "deviceLinearUpwindCorr3");
}

// deviceScaledDev2T (src/cuda/device_divdevreff.cuh:102-103): sigma = nu*dev2(T(gradU)), tensors packed
// [(i*3 + j)*n + c]
inline void deviceScaledDev2T(int n, const DeviceBuffer<scalar>& gradU, const DeviceBuffer<scalar>& nu,
                              DeviceBuffer<scalar>& sigma)
{
    mfv_ctx ctx = threadContext();
    sigma.resize(9 * static_cast<std::size_t>(n));
    mfvCheck(mfv_scaled_dev2t(ctx, n, gradU.data(), nu.data(), sigma.data()), ctx, "deviceScaledDev2T");
}

// deviceTensorDivSource (src/cuda/device_divdevreff.cuh:104-107): V*fvc::div(T) per cell from the cell
// tensors and the boundary faces' tensors (both packed [(i*3 + j)*n + e]); empty faces per isEmpty (may be
// nullptr). Cyclic and AMI interfaces are not carried.
inline void deviceTensorDivSource(const DeviceMesh& dm, const DeviceBuffer<scalar>& Tcell,
                                  const DeviceBuffer<scalar>& Tbnd, DeviceBuffer<scalar>& srcX,
                                  DeviceBuffer<scalar>& srcY, DeviceBuffer<scalar>& srcZ,
                                  const DeviceBuffer<int>* isEmpty = nullptr)
{
    mfv_ctx ctx = threadContext();
    srcX.resize(dm.nCells);
    srcY.resize(dm.nCells);
    srcZ.resize(dm.nCells);
    mfvCheck(mfv_tensor_div_source(ctx, dm.mesh, dm.weights.data(), dm.Sf.data(), isEmpty ? isEmpty->data() : nullptr,
                                   Tcell.data(), Tbnd.data(), srcX.data(), srcY.data(), srcZ.data()),
             ctx, "deviceTensorDivSource");
}

// deviceDivDevReff's sigma and tensor-divergence stages (src/cuda/device_divdevreff.cu:408-620) in one
// pass: V*div(nu*dev2(T(gradU))) without storing sigma; sigmaB is the boundary faces' stress
inline void deviceDivDevReffExplicit(const DeviceMesh& dm, const DeviceBuffer<scalar>& gradU,
                                     const DeviceBuffer<scalar>& nu, const DeviceBuffer<scalar>& sigmaB,
                                     DeviceBuffer<scalar>& srcX, DeviceBuffer<scalar>& srcY,
                                     DeviceBuffer<scalar>& srcZ, const DeviceBuffer<int>* isEmpty = nullptr)
{
    mfv_ctx ctx = threadContext();
    srcX.resize(dm.nCells);
    srcY.resize(dm.nCells);
    srcZ.resize(dm.nCells);
    mfvCheck(mfv_div_dev_reff_explicit(ctx, dm.mesh, dm.weights.data(), dm.Sf.data(),
                                       isEmpty ? isEmpty->data() : nullptr, gradU.data(), nu.data(), sigmaB.data(),
                                       srcX.data(), srcY.data(), srcZ.data()),
             ctx, "deviceDivDevReffExplicit");
}

// deviceCmptMaxMag3 / deviceCmptMin3 (src/cuda/device_simple.cuh:123-128)
inline void deviceCmptMaxMag3(const DeviceBuffer<scalar>& a, const DeviceBuffer<scalar>& b,
                              const DeviceBuffer<scalar>& c, DeviceBuffer<scalar>& out)
{
    mfv_ctx ctx = threadContext();
    out.resize(a.size());
    mfvCheck(mfv_cmpt_max_mag3(ctx, static_cast<int>(a.size()), a.data(), b.data(), c.data(), out.data()), ctx,
             "deviceCmptMaxMag3");
}

inline void deviceCmptMin3(const DeviceBuffer<scalar>& a, const DeviceBuffer<scalar>& b, const DeviceBuffer<scalar>& c,
                           DeviceBuffer<scalar>& out)
{
    mfv_ctx ctx = threadContext();
    out.resize(a.size());
    mfvCheck(mfv_cmpt_min3(ctx, static_cast<int>(a.size()), a.data(), b.data(), c.data(), out.data()), ctx,
             "deviceCmptMin3");
}

// relaxField (src/applications/solvers/simpleFoam/pEqn.cuh:154): p = pPrev + alpha*(p - pPrev)
inline void relaxField(DeviceBuffer<scalar>& p, const DeviceBuffer<scalar>& pPrev, scalar alpha)
{
    mfv_ctx ctx = threadContext();
    mfvCheck(mfv_relax_field(ctx, static_cast<int>(p.size()), p.data(), pPrev.data(), alpha), ctx, "relaxField");
}

// deviceVectorFlux (src/cuda/device_simple.cuh:48): phiInt = interpolate(U) . Sf
inline void deviceVectorFlux(const DeviceMesh& dm, const DeviceBuffer<scalar>& Ux, const DeviceBuffer<scalar>& Uy,
                             const DeviceBuffer<scalar>& Uz, DeviceBuffer<scalar>& phiInt)
{
    mfv_ctx ctx = threadContext();
    phiInt.resize(dm.nInternalFaces);
    mfvCheck(mfv_vector_flux(ctx, dm.mesh, dm.weights.data(), dm.Sf.data(), Ux.data(), Uy.data(), Uz.data(),
                             phiInt.data()),
             ctx, "deviceVectorFlux");
}

// deviceBoundaryFlux (src/cuda/device_simple.cuh:59): phiB = Ub . Sf_b (zero on the faces isEmpty marks)
inline void deviceBoundaryFlux(const DeviceMesh& dm, const DeviceBuffer<scalar>& uxb, const DeviceBuffer<scalar>& uyb,
                               const DeviceBuffer<scalar>& uzb, DeviceBuffer<scalar>& phiB,
                               const DeviceBuffer<int>* isEmpty = nullptr)
{
    mfv_ctx ctx = threadContext();
    phiB.resize(dm.nBoundaryFaces);
    mfvCheck(mfv_boundary_flux(ctx, dm.mesh, dm.Sf.data(), uxb.data(), uyb.data(), uzb.data(),
                               isEmpty ? isEmpty->data() : nullptr, phiB.data()),
             ctx, "deviceBoundaryFlux");
}

// deviceAdjustPhi (src/cuda/device_simple.cuh:99-102): the adjustable outflow of phiB scaled in place; returns
// massCorr; throws when the continuity error cannot be removed (adjustPhi.C's FatalError)
inline scalar deviceAdjustPhi(const DeviceBuffer<label>& adjustable, DeviceBuffer<scalar>& phiB,
                              const DeviceBuffer<scalar>* phiInt = nullptr, bool* closedVolume = nullptr)
{
    mfv_ctx ctx = threadContext();
    double mc = 1.0;
    int closed = 0;
    mfvCheck(mfv_adjust_phi(ctx, static_cast<int>(phiB.size()), adjustable.data(), phiB.data(),
                            phiInt ? static_cast<int>(phiInt->size()) : 0, phiInt ? phiInt->data() : nullptr, &mc,
                            &closed),
             ctx, "deviceAdjustPhi");
    if (closedVolume) *closedVolume = closed != 0;
    return mc;
}

// fvc::gaussGrad with the patch values of a field whose patches are fixedValue (bcType 1, value
// bcValue) or zeroGradient (0) or empty (2); grad is xyz-interleaved (3*nCells)
inline void deviceGaussGrad(const DeviceMesh& dm, const DeviceBuffer<scalar>& vol, const DeviceBuffer<int>& bcType,
                            const DeviceBuffer<scalar>& bcValue, DeviceBuffer<scalar>& grad)
{
    mfv_ctx ctx = threadContext();
    grad.resize(3 * (std::size_t)dm.nCells);
    mfvCheck(mfv_grad_gauss(ctx, dm.mesh, bcType.data(), bcValue.data(), dm.Sf.data(), dm.weights.data(),
                            dm.V.data(), vol.data(), grad.data()),
             ctx, "deviceGaussGrad");
}

// deviceLaplacianCorrFlux + deviceFaceDivSource (device_mesh.cuh:536-552): ffc per internal face and,
// when src is given, src -= div(ffc); grad xyz-interleaved
inline void deviceLaplacianCorrFlux(const DeviceMesh& dm, const DeviceBuffer<scalar>& gammafInt,
                                    const DeviceBuffer<scalar>& grad, DeviceBuffer<scalar>& ffc,
                                    DeviceBuffer<scalar>* src = nullptr)
{
    mfv_ctx ctx = threadContext();
    ffc.resize(dm.nInternalFaces);
    mfvCheck(mfv_laplacian_corr(ctx, dm.mesh, grad.data(), dm.weights.data(), dm.corrVec.data(), gammafInt.data(),
                                dm.magSfInternal.data(), ffc.data(), src ? src->data() : nullptr),
             ctx, "deviceLaplacianCorrFlux");
}

// deviceLaplacianCorrFluxLimited (device_fvm.cu:497-517): `limited psi corrected`, the correction capped per
// face against the orthogonal snGrad of phi; psi >= 1 is the unlimited correction
inline void deviceLaplacianCorrFluxLimited(const DeviceMesh& dm, const DeviceBuffer<scalar>& gammafInt,
                                           const DeviceBuffer<scalar>& phi, const DeviceBuffer<scalar>& grad,
                                           scalar psi, DeviceBuffer<scalar>& ffc, DeviceBuffer<scalar>* src = nullptr)
{
    mfv_ctx ctx = threadContext();
    ffc.resize(dm.nInternalFaces);
    mfvCheck(mfv_laplacian_corr_limited(ctx, dm.mesh, grad.data(), dm.weights.data(), dm.corrVec.data(),
                                        gammafInt.data(), dm.magSfInternal.data(), dm.nonOrthDeltaCoeffs.data(),
                                        phi.data(), psi, ffc.data(), src ? src->data() : nullptr),
             ctx, "deviceLaplacianCorrFluxLimited");
}

// ---------------- products and norms ----------------

// deviceAmul (device_ldu.cuh:241); onField has no effect without cyclic jumps, which the path refuses
inline void deviceAmul(const DeviceLduView& A, const DeviceBuffer<scalar>& psi, DeviceBuffer<scalar>& Apsi,
                       bool onField = false)
{
    (void)onField;
    mfv_ctx ctx = threadContext();
    Apsi.resize(A.nCells);
    const mfv_ldu l = A.ldu();
    mfvCheck(mfv_amul(ctx, A.mesh, &l, psi.data(), Apsi.data()), ctx, "deviceAmul");
}

// deviceNormFactorInto (device_pcg.cuh:44); `ones` is not needed (sumA is formed in place)
inline void deviceNormFactorInto(const DeviceLduView& A, const DeviceBuffer<scalar>& psi,
                                 const DeviceBuffer<scalar>& b, const DeviceBuffer<scalar>& ones,
                                 DeviceBuffer<scalar>& dNf)
{
    (void)ones;
    mfv_ctx ctx = threadContext();
    dNf.resize(1);
    const mfv_ldu l = A.ldu();
    mfvCheck(mfv_norm_factor(ctx, A.mesh, &l, psi.data(), b.data(), dNf.data()), ctx, "deviceNormFactorInto");
}

// ---------------- AMG hierarchy and solve ----------------

// AMGData (device_amg.cuh:308-388): the hierarchy, its work vectors, the V-cycle precision. Movable.
struct AMGData
{
    mfv_amg amg = nullptr;
    mfv_mesh addressing = nullptr;   // the fine addressing the hierarchy was built on
    int nFine = 0;
    int precision = MFV_PREC_FP32;   // the reference's FP32 V-cycle
    int nPreSweeps = 1, nPostSweeps = 1;

    AMGData() = default;
    AMGData(const AMGData&) = delete;
    AMGData& operator=(const AMGData&) = delete;
    AMGData(AMGData&& o) noexcept { *this = std::move(o); }
    AMGData& operator=(AMGData&& o) noexcept
    {
        if (this != &o)
        {
            reset();
            amg = o.amg;
            addressing = o.addressing;
            owner_ = std::move(o.owner_);
            nFine = o.nFine;
            precision = o.precision;
            nPreSweeps = o.nPreSweeps;
            nPostSweeps = o.nPostSweeps;
            o.amg = nullptr;
            o.addressing = nullptr;
        }
        return *this;
    }
    ~AMGData() { reset(); }
    int nLevels() const
    {
        int n = 0;
        if (amg) mfv_amg_num_grids(amg, &n);
        return n > 0 ? n - 1 : 0;
    }
    // takes ownership of a new addressing handle (destroyed with the last hierarchy that uses it)
    void adoptAddressing(mfv_mesh m)
    {
        addressing = m;
        owner_ = std::shared_ptr<mfv_mesh_s>(m, [](mfv_mesh_s* p) { mfv_mesh_destroy(p); });
    }
    // shares another hierarchy's addressing (a clone lives on its source's mesh)
    void shareAddressing(const AMGData& o)
    {
        addressing = o.addressing;
        owner_ = o.owner_;
    }

private:
    std::shared_ptr<mfv_mesh_s> owner_;
    void reset()
    {
        if (amg) mfv_amg_destroy(amg);
        amg = nullptr;
        owner_.reset();
        addressing = nullptr;
    }
};

// buildAMG (device_amg.cuh:393): brae's pairwise hierarchy from the fine internal addressing and face
// weights (|Sf|); smoothed aggregation is not carried and is refused
inline AMGData buildAMG(const std::vector<label>& fineOwner, const std::vector<label>& fineNei,
                        const std::vector<scalar>& faceWeights, int nFine, const bool* smoothedAggregation = nullptr)
{
    if (smoothedAggregation && *smoothedAggregation)
        throw std::runtime_error("brae hip: buildAMG: smoothed aggregation is not carried");
    mfv_ctx ctx = threadContext();
    AMGData a;
    a.nFine = nFine;
    mfv_mesh m = nullptr;
    mfvCheck(mfv_mesh_create(ctx, nFine, static_cast<int>(fineOwner.size()), fineOwner.data(), fineNei.data(), 0,
                             nullptr, &m),
             ctx, "buildAMG addressing");
    a.adoptAddressing(m);
    mfv_amg_controls c{MFV_AGG_BRAE, 64, 1};
    mfvCheck(mfv_amg_build(ctx, a.addressing, faceWeights.data(), &c, &a.amg), ctx, "buildAMG");
    return a;
}

// AMGCacheRead (device_amg.cuh:410-416): what the cache directory held
enum class AMGCacheRead
{
    loaded,
    absent,
    otherMeshOrBuild,
    unreadable
};

// buildOrLoadAMG (device_amg.cuh:448-451): the hierarchy from <cacheDir>'s file when it holds this mesh's
// and this build's signature, else built (and written there with writeCache); `how` says what the file held
inline AMGData buildOrLoadAMG(const std::vector<label>& fineOwner, const std::vector<label>& fineNei,
                              const std::vector<scalar>& faceWeights, int nFine, const std::string& cacheDir,
                              bool writeCache, const bool* smoothedAggregation = nullptr,
                              AMGCacheRead* how = nullptr)
{
    if (smoothedAggregation && *smoothedAggregation)
        throw std::runtime_error("brae hip: buildOrLoadAMG: smoothed aggregation is not carried");
    if (fineOwner.size() != fineNei.size() || faceWeights.size() != fineOwner.size())
        throw std::runtime_error("brae hip: buildOrLoadAMG: owner, neighbour and weights differ in length");
    mfv_ctx ctx = threadContext();
    AMGData a;
    a.nFine = nFine;
    mfv_mesh m = nullptr;
    mfvCheck(mfv_mesh_create(ctx, nFine, static_cast<int>(fineOwner.size()), fineOwner.data(), fineNei.data(), 0,
                             nullptr, &m),
             ctx, "buildOrLoadAMG addressing");
    a.adoptAddressing(m);
    mfv_amg_controls c{MFV_AGG_BRAE, 64, 1};
    int h = MFV_CACHE_ABSENT;
    mfvCheck(mfv_amg_build_or_load(ctx, a.addressing, faceWeights.data(), &c, cacheDir.c_str(), writeCache ? 1 : 0,
                                   &a.amg, &h),
             ctx, "buildOrLoadAMG");
    if (how)
        *how = h == MFV_CACHE_LOADED             ? AMGCacheRead::loaded
               : h == MFV_CACHE_OTHER_MESH_OR_BUILD ? AMGCacheRead::otherMeshOrBuild
               : h == MFV_CACHE_UNREADABLE          ? AMGCacheRead::unreadable
                                                    : AMGCacheRead::absent;
    return a;
}

// cloneAMG (device_amg_cache.cu:70): a second hierarchy of the same structure on the same fine addressing;
// its coarse values start unset (amgGalerkin), its precision and sweeps are the source's
inline AMGData cloneAMG(const AMGData& S)
{
    if (!S.amg || !S.addressing) throw std::runtime_error("brae hip: cloneAMG: empty hierarchy");
    mfv_ctx ctx = threadContext();
    AMGData a;
    a.nFine = S.nFine;
    a.precision = S.precision;
    a.nPreSweeps = S.nPreSweeps;
    a.nPostSweeps = S.nPostSweeps;
    a.shareAddressing(S);
    mfvCheck(mfv_amg_clone(ctx, S.amg, &a.amg), ctx, "cloneAMG");
    return a;
}

// amgGalerkin (device_amg.cuh:454): symmetric matrices only (amgRefuseAsymmetric)
inline void amgGalerkin(AMGData& A, const DeviceBuffer<scalar>& fineDiag, const DeviceBuffer<scalar>& fineUpper,
                        const DeviceBuffer<scalar>& fineLower)
{
    mfv_ctx ctx = threadContext();
    mfvCheck(mfv_amg_galerkin(ctx, A.amg, fineDiag.data(), fineUpper.data(), fineLower.data(), A.precision), ctx,
             "amgGalerkin");
}

// deviceAMGPCG (device_amg.cuh:550-553) with the normFactor on the device (NULL: computed here);
// captureVcycle replays captured groups of iterations (a HIP graph per hierarchy)
inline DeviceSolverPerf deviceAMGPCG(const DeviceLduView& Afine, AMGData& amg, const DeviceBuffer<scalar>& b,
                                     DeviceBuffer<scalar>& psi, const scalar* dNormFactor, scalar tol, scalar relTol,
                                     int maxIter, bool captureVcycle = false, int checkEvery = 1,
                                     bool corrScaling = false, int minIter = 0)
{
    mfv_ctx ctx = threadContext();
    mfv_pcg_controls c;
    mfv_pcg_default_controls(&c);
    c.tolerance = tol;
    c.relTol = relTol;
    c.maxIter = maxIter;
    c.minIter = minIter;
    c.precision = amg.precision;
    c.nPreSweeps = amg.nPreSweeps;
    c.nPostSweeps = amg.nPostSweeps;
    c.checkEvery = checkEvery;
    c.captureVcycle = captureVcycle ? 1 : 0;
    c.scaleCorrection = corrScaling ? 1 : 0;
    c.flexible = corrScaling ? 1 : 0;   // a scaled cycle is nonlinear: flexible CG, as the reference switches
    if (Afine.nCells != amg.nFine) throw std::runtime_error("brae hip: deviceAMGPCG: matrix and hierarchy sizes differ");
    mfv_perf perf{};
    const mfv_ldu l = Afine.ldu();
    // the hierarchy's own copy of the fine addressing (the same internal faces as Afine's mesh)
    mfvCheck(mfv_amg_pcg_solve(ctx, amg.addressing, amg.amg, &l, b.data(), psi.data(), dNormFactor, &c, &perf), ctx,
             "deviceAMGPCG");
    return DeviceSolverPerf{perf.initialResidual, perf.finalResidual, perf.nIterations};
}

// the same solve with the normFactor given on the host
inline DeviceSolverPerf deviceAMGPCG(const DeviceLduView& Afine, AMGData& amg, const DeviceBuffer<scalar>& b,
                                     DeviceBuffer<scalar>& psi, scalar normFactor, scalar tol, scalar relTol,
                                     int maxIter, bool captureVcycle = false, int checkEvery = 1,
                                     bool corrScaling = false, int minIter = 0)
{
    DeviceBuffer<scalar> nf(std::vector<scalar>{normFactor});
    return deviceAMGPCG(Afine, amg, b, psi, nf.data(), tol, relTol, maxIter, captureVcycle, checkEvery, corrScaling,
                        minIter);
}

// ---------------- OpenFOAM's GAMG (src/matrices/lduMatrix/solvers/GAMG/gamg_solver_cpp.cuh) ----------------

// SolverPerformance (src/OpenFOAM/matrices/pcg.cuh:16-21)
struct SolverPerformance
{
    scalar initialResidual = 0, finalResidual = 0;
    int nIterations = 0;
    bool converged = false;
};

// GamgControls (gamg_solver_cpp.cuh:67-85) with GAMGSolver.C's defaults
struct GamgControls
{
    std::string smoother = "GaussSeidel";
    scalar tolerance = 1e-6;
    scalar relTol = 0;
    int maxIter = 1000;
    int minIter = 0;
    int nPreSweeps = 0;
    int preSweepsLevelMultiplier = 1;
    int maxPreSweeps = 4;
    int nPostSweeps = 2;
    int postSweepsLevelMultiplier = 1;
    int maxPostSweeps = 4;
    int nFinestSweeps = 2;
    bool scaleCorrection = true;
    label nCellsInCoarsestLevel = 10;
};

// GamgPreconditionerControls (gamg_solver_cpp.cuh:127-131)
struct GamgPreconditionerControls
{
    GamgControls gamg;
    int nVcycles = 2;
};

// gamgSmootherPorted (gamg_solver_cpp.cuh:134): the smoothers the GAMG runs
inline bool gamgSmootherPorted(const std::string& s)
{
    return s == "GaussSeidel" || s == "symGaussSeidel" || s == "DIC" || s == "DICGaussSeidel";
}

inline mfv_gamg_controls toMfv(const GamgControls& g, int nVcycles, const char* who)
{
    mfv_gamg_controls c;
    mfv_gamg_default_controls(&c);
    if (g.smoother == "GaussSeidel")
        c.smoother = MFV_SMOOTHER_GAUSS_SEIDEL;
    else if (g.smoother == "symGaussSeidel")
        c.smoother = MFV_SMOOTHER_SYM_GAUSS_SEIDEL;
    else if (g.smoother == "DIC")
        c.smoother = MFV_SMOOTHER_DIC;
    else if (g.smoother == "DICGaussSeidel")
        c.smoother = MFV_SMOOTHER_DIC_GAUSS_SEIDEL;
    else
        throw std::runtime_error(std::string("brae hip: ") + who + ": smoother " + g.smoother + " is not carried");
    c.tolerance = g.tolerance;
    c.relTol = g.relTol;
    c.maxIter = g.maxIter;
    c.minIter = g.minIter;
    c.nPreSweeps = g.nPreSweeps;
    c.preSweepsLevelMultiplier = g.preSweepsLevelMultiplier;
    c.maxPreSweeps = g.maxPreSweeps;
    c.nPostSweeps = g.nPostSweeps;
    c.postSweepsLevelMultiplier = g.postSweepsLevelMultiplier;
    c.maxPostSweeps = g.maxPostSweeps;
    c.nFinestSweeps = g.nFinestSweeps;
    c.scaleCorrection = g.scaleCorrection ? 1 : 0;
    c.nVcycles = nVcycles;
    return c;
}

// faceAreaPairGamgAgglomeration (pair_gamg_agglomeration_cpp.cu:286-345): OpenFOAM's pairwise levels
// from the face weights |(Sf_x, 1.01 Sf_y, 1.02 Sf_z)|/sqrt|Sf| down to nCellsInCoarsestLevel. Sf is
// xyz-interleaved per internal face.
inline AMGData faceAreaPairGamgAgglomeration(const std::vector<label>& fineOwner, const std::vector<label>& fineNei,
                                             const std::vector<scalar>& Sf, const std::vector<scalar>& magSf,
                                             int nFine, label nCellsInCoarsestLevel = 10)
{
    mfv_ctx ctx = threadContext();
    const int nF = static_cast<int>(fineOwner.size());
    if (Sf.size() != 3 * (size_t)nF || magSf.size() != (size_t)nF)
        throw std::runtime_error("brae hip: faceAreaPairGamgAgglomeration: Sf/magSf sizes differ from the faces");
    std::vector<scalar> w(nF);
    mfvCheck(mfv_faceareapair_weights(nF, Sf.data(), magSf.data(), w.data()), ctx, "faceAreaPair weights");
    AMGData a;
    a.nFine = nFine;
    a.precision = MFV_PREC_FP64;
    mfv_mesh m = nullptr;
    mfvCheck(mfv_mesh_create(ctx, nFine, nF, fineOwner.data(), fineNei.data(), 0, nullptr, &m), ctx,
             "faceAreaPairGamgAgglomeration addressing");
    a.adoptAddressing(m);
    mfv_amg_controls c{MFV_AGG_FACEAREAPAIR, nCellsInCoarsestLevel, 1};
    mfvCheck(mfv_amg_build(ctx, a.addressing, w.data(), &c, &a.amg), ctx, "faceAreaPairGamgAgglomeration");
    return a;
}

// gamgSolveFolded (gamg_solver_cpp.cuh:185-195) on the device: the folded symmetric system (diag,
// upper), psi in place; the coarse matrices are agglomerated from it on every call
inline SolverPerformance gamgSolveFolded(AMGData& agglomeration, const DeviceBuffer<scalar>& diag,
                                         const DeviceBuffer<scalar>& upper, const DeviceBuffer<scalar>& source,
                                         DeviceBuffer<scalar>& psi, const GamgControls& controls)
{
    mfv_ctx ctx = threadContext();
    const mfv_gamg_controls c = toMfv(controls, 2, "gamgSolveFolded");
    const mfv_ldu l{diag.data(), upper.data(), upper.data()};
    mfv_perf perf{};
    mfvCheck(mfv_gamg_solve(ctx, agglomeration.addressing, agglomeration.amg, &l, source.data(), psi.data(), &c,
                            &perf),
             ctx, "gamgSolveFolded");
    return SolverPerformance{perf.initialResidual, perf.finalResidual, perf.nIterations, perf.converged != 0};
}

// pcgGamgSolve (gamg_solver_cpp.cuh:197-216) on the device, on a folded symmetric system: PCG's
// tolerance..minIter, the preconditioner's GAMG controls and nVcycles
inline SolverPerformance pcgGamgSolveFolded(AMGData& agglomeration, const DeviceBuffer<scalar>& diag,
                                            const DeviceBuffer<scalar>& upper, const DeviceBuffer<scalar>& source,
                                            DeviceBuffer<scalar>& psi, scalar tolerance, scalar relTol, int maxIter,
                                            int minIter, const GamgPreconditionerControls& precond)
{
    mfv_ctx ctx = threadContext();
    const mfv_gamg_controls c = toMfv(precond.gamg, precond.nVcycles, "pcgGamgSolve");
    const mfv_ldu l{diag.data(), upper.data(), upper.data()};
    mfv_perf perf{};
    mfvCheck(mfv_pcg_gamg_solve(ctx, agglomeration.addressing, agglomeration.amg, &l, source.data(), psi.data(),
                                tolerance, relTol, maxIter, minIter, &c, &perf),
             ctx, "pcgGamgSolve");
    return SolverPerformance{perf.initialResidual, perf.finalResidual, perf.nIterations, perf.converged != 0};
}

} // namespace brae
