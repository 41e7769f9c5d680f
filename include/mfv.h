/*
 * mfv.h -- C ABI of the MI355X-native finite-volume assembly and pressure-solve library (libmfv.so).
 *
 * It replaces the reference's device lduMatrix / assembly / AMG-PCG entry points for the pressure
 * path. Every entry point says which reference interface it stands for (file:line, relative to the
 * reference root). Conventions:
 *   - int32 labels, double scalars, structure-of-arrays; internal faces in OpenFOAM's
 *     upper-triangular order (sorted by owner, owner < neighbour);
 *   - pointers named d* are DEVICE pointers (HBM), h* are host pointers;
 *   - every call returns an int status (MFV_OK or an MFV_ERR_*), never throws, and records a
 *     message mfv_last_error(ctx) returns; work is enqueued on the context's HIP stream and is
 *     asynchronous unless the entry point says it synchronises;
 *   - a context is used by one host thread at a time; different contexts are independent.
 *   - a symmetric matrix passes dLower == dUpper (one off-diagonal, device_amg.cu:1703-1710).
 * Non-convergence of a solve is not an error: it returns MFV_OK with nIterations == maxIter.
 */
#ifndef MFV_H
#define MFV_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MFV_OK 0
#define MFV_ERR_INVALID 1     /* bad argument or mesh (names it) */
#define MFV_ERR_HIP 2         /* a HIP runtime call failed */
#define MFV_ERR_UNSUPPORTED 3 /* an input the path does not carry (refused by name) */
#define MFV_ERR_ALLOC 4       /* device allocation failed */
#define MFV_ERR_STATE 5       /* call order (e.g. a solve before amg_galerkin) */

typedef struct mfv_ctx_s* mfv_ctx;
typedef struct mfv_mesh_s* mfv_mesh;
typedef struct mfv_amg_s* mfv_amg;

/* Values of an LDU matrix on a mesh's (or an AMG level's) addressing; device pointers.
 * Stands for DeviceLduView's value half (src/matrices/lduMatrix/lduMatrix/device_ldu.cuh:20-66). */
typedef struct
{
    const double* dDiag;
    const double* dUpper;
    const double* dLower;
} mfv_ldu;

/* ---------------------------------------------------------------------------------------------
 * Context: one HIP stream, the reduction scratch and a pinned read-back mailbox. Replaces the
 * process-wide singletons of reductions.cu:89-111 so concurrent contexts do not share state.
 * hStream may be NULL (the context creates its own default-priority stream).
 * Contexts and meshes are reference counted: whatever is built on them (meshes, hierarchies, SIMPLE and
 * kEpsilon workspaces) keeps them alive, so destroy calls may come in any order; the last one frees.
 * ------------------------------------------------------------------------------------------- */
int mfv_ctx_create(int device, void* hStream, mfv_ctx* out);
int mfv_ctx_destroy(mfv_ctx ctx);
const char* mfv_last_error(mfv_ctx ctx);
/* hipStream_t of the context */
void* mfv_ctx_stream(mfv_ctx ctx);
/* waits for the context's stream */
int mfv_ctx_sync(mfv_ctx ctx);
/* 1: every dot and sum|x| runs in OpenFOAM's sequential order (one thread; the reference's
 * BRAE_DEVICE_OF_REDUCE, reductions.cu:36-56); 0 (default): the deterministic wave64 tree. */
int mfv_ctx_set_sequential_reduce(mfv_ctx ctx, int sequential);
/* Kernel timing with HIP events recorded around the launches of each kernel class on the context's
 * stream (no counterpart in the reference, whose timing is BRAE_PHASE_TIME/BRAE_AMG_PCG_SPLIT by
 * draining the device). Each tag below 16 times one kernel: 0 PCG p-update + FP64 product
 * (pcgDirAmulK), 1 PCG update (pcgUpdateK), 2 PCG z.r (pcgZrK), 3 one whole V-cycle, 4 grid-0
 * residual (residualK), 5 grid-0 correction update + first post-sweep (corrSmoothK / prolongSmoothK),
 * 6 grid 0 -> 1 restriction (restrictK), 7 grid-0 scaled correction product (corrAmulK);
 * 16 + g: grid g's way down; 80 + g: grid g's way up. */
int mfv_ctx_profile(mfv_ctx ctx, int enable);
/* launches recorded for `tag` since the last read and their summed duration; synchronises, resets */
int mfv_ctx_profile_read(mfv_ctx ctx, int tag, int* count, double* totalMs);

/* ---------------------------------------------------------------------------------------------
 * Mesh addressing: lduAddressing of the internal faces (owner, neighbour, ownerStart, losort,
 * losortStart) and the boundary gather (bFaceCell). Host arrays in; uploaded and kept on the
 * device. Refuses faces not in upper-triangular order.
 * Replaces buildDeviceLdu (src/matrices/lduMatrix/lduMatrix/device_ldu.cuh:109-146) and the
 * DeviceMesh addressing upload (src/cuda/device_mesh.cuh:31-60).
 * ------------------------------------------------------------------------------------------- */
int mfv_mesh_create(mfv_ctx ctx, int nCells, int nFaces, const int* hOwner, const int* hNeighbour, int nBFaces,
                    const int* hBFaceCell, mfv_mesh* out);
int mfv_mesh_destroy(mfv_mesh mesh);
/* device pointers owned by the mesh: owner, neighbour [nFaces]; ownerStart, losortStart [nCells+1];
 * losort [nFaces] (any may be NULL) */
int mfv_mesh_addressing(mfv_mesh mesh, const int** dOwner, const int** dNeighbour, const int** dOwnerStart,
                        const int** dLosortStart, const int** dLosort);

/* ---------------------------------------------------------------------------------------------
 * Assembly. All device pointers. Boundary arrays are per boundary face (nBFaces), bcType
 * 0 zeroGradient, 1 fixedValue, 2 empty.
 * ------------------------------------------------------------------------------------------- */
/* deviceLaplacianCoeffs / deviceLaplacianCoeffsSymmetric (src/cuda/device_fvm.cu:304-367):
 * upper = dc*gammaf*magSf per internal face (dLower may equal dUpper or be NULL), diag = -sum. */
int mfv_laplacian_coeffs(mfv_ctx ctx, mfv_mesh mesh, const double* dDeltaCoeffs, const double* dGammaf,
                         const double* dMagSf, double* dDiag, double* dUpper, double* dLower);
/* laplacianBoundaryCoeffs (src/finiteVolume/finiteVolume/fvm.cuh:43-103), fixedValue/zeroGradient */
int mfv_laplacian_bcoeffs(mfv_ctx ctx, mfv_mesh mesh, const int* dBcType, const double* dBcValue,
                          const double* dGammab, const double* dBMagSf, const double* dBDeltaCoeffs, double* dIC,
                          double* dBC);
/* deviceDivUpwindCoeffs (src/cuda/device_fvm.cu:646-662): lower = -pos0(phi)*phi, upper = lower + phi */
int mfv_div_upwind_coeffs(mfv_ctx ctx, mfv_mesh mesh, const double* dPhi, double* dDiag, double* dUpper,
                          double* dLower);
/* deviceDivCentralCoeffs (src/cuda/device_fvm.cu:628-643; Gauss linear): lower = -w*phi, upper = lower + phi
 * with the geometric weights dWeights [nFaces] (mfv_mesh_geometry's) */
int mfv_div_central_coeffs(mfv_ctx ctx, mfv_mesh mesh, const double* dPhi, const double* dWeights, double* dDiag,
                           double* dUpper, double* dLower);
/* fvm::div boundary coefficients (fvm.cuh:545-552) */
int mfv_div_bcoeffs(mfv_ctx ctx, mfv_mesh mesh, const int* dBcType, const double* dBcValue, const double* dPhib,
                    double* dIC, double* dBC);
/* deviceFold (src/cuda/device_fvm.cu:895-912): diagC = diag + sum iC, b = source + sum bC */
int mfv_fold(mfv_ctx ctx, mfv_mesh mesh, const double* dDiag, const double* dSource, const double* dIC,
             const double* dBC, double* dDiagC, double* dB);
/* deviceSetReference (src/applications/solvers/simpleFoam/device_simple.cuh:104-106):
 * b[refCell] += diag[refCell]*refValue, then diag[refCell] += diag[refCell] (the source takes the OLD diagonal).
 * nCells is the length of both arrays; refCell < 0 is a no-op, refCell >= nCells is refused. */
int mfv_set_reference(mfv_ctx ctx, int nCells, double* dDiag, double* dB, int refCell, double refValue);
/* deviceMatrixFluxInternal (device_simple.cuh:51-52): flux = upper*psi[nei] - lower*psi[own] */
int mfv_face_flux(mfv_ctx ctx, mfv_mesh mesh, const mfv_ldu* A, const double* dPsi, double* dFlux);

/* fvc::interpolate, linear (src/finiteVolume/finiteVolume/fvc.cu:626-656): face = w*vol[own] +
 * (1-w)*vol[nei]; dBFace (may be NULL) takes the owner cell's value (an extrapolated field such as rAU,
 * src/applications/solvers/simpleFoam/pEqn_cpp.cu:217-222) */
int mfv_interpolate(mfv_ctx ctx, mfv_mesh mesh, const double* dWeights, const double* dVol, double* dFace,
                    double* dBFace);
/* the pressure equation's right-hand side (pEqn_cpp.cu:238-242): source += sign * div(phi)*V, with
 * div(phi) = (sum_own phi - sum_nei phi + sum_boundary phiB)/V (fvc.cu:706-730); dPhiB may be NULL */
int mfv_div_source(mfv_ctx ctx, mfv_mesh mesh, const double* dPhi, const double* dPhiB, const double* dV,
                   double sign, double* dSource);
/* correctFlux (pEqn_cpp.cuh:134-139) with fvMatrix::flux's stored correction (fv_matrix_ops.cuh:33-50):
 * phi = phiHbyA - sign*(upper*psi[nei] - lower*psi[own] + ffc); dFfc may be NULL (no non-orthogonal
 * correction); sign -1 when A is the negated (positive-definite) pressure matrix */
int mfv_flux_correct(mfv_ctx ctx, mfv_mesh mesh, const mfv_ldu* A, const double* dPsi, const double* dFfc,
                     const double* dPhiHbyA, double sign, double* dPhi);
/* ---------------------------------------------------------------------------------------------
 * The SIMPLE pressure predictor's momentum glue (pressurePredictor,
 * src/applications/solvers/simpleFoam/pEqn.cu:94-335; host twin pEqn_cpp.cu:40-133): rAU, H() per
 * component and HbyA, constrainHbyA, phiHbyA and adjustPhi, for uncoupled patches (no SIMPLEC, MRF or
 * cyclic pair). The momentum matrix comes as an mfv_ldu (upper/lower) plus, per boundary face and
 * component, its internalCoeffs (dIcX/Y/Z) and boundaryCoeffs. Boundary arrays in boundary-face order.
 * ------------------------------------------------------------------------------------------- */
/* A()'s numerator (pEqn.cu:117-138, cmptAvKernel + foldBoundaryDiagKernel; fv_matrix_ops.cuh:211-226):
 * D = diag + cmptAv(internalCoeffs) face by face, cmptAv = (x + y + z)/3; dDiag the (relaxed) diagonal */
int mfv_fold_boundary_diag(mfv_ctx ctx, mfv_mesh mesh, const double* dDiag, const double* dIcX,
                           const double* dIcY, const double* dIcZ, double* dD);
/* deviceReciprocalV (src/cuda/device_simple.cuh:26): rAU = 1/(D/V) */
int mfv_reciprocal_v(mfv_ctx ctx, mfv_mesh mesh, const double* dD, const double* dV, double* dRAU);
/* H()'s boundary-diagonal factor of component k (pEqn.cu:146-149): cmptAv(ic) - ic_k per boundary face */
int mfv_boundary_diag_term(mfv_ctx ctx, mfv_mesh mesh, const double* dIcX, const double* dIcY, const double* dIcZ,
                           int k, double* dBdDiag);
/* deviceMatrixH (src/cuda/device_simple.cuh:14-23) of one component: H = (bdDiag*psi + lduMatrix::H(psi) +
 * source + bdSrc)/V in the host loop's order (fv_matrix_ops.cuh:277-340). A's upper and lower are read.
 * dRAU (may be NULL): the output is HbyA = rAU*H (pEqn.cu:185, deviceHadamard). valid = 0: a component
 * the mesh's empty patches knock out (fvMatrix::H's validComponents): zeros. */
int mfv_matrix_h(mfv_ctx ctx, mfv_mesh mesh, const mfv_ldu* A, const double* dPsi, const double* dSource,
                 const double* dBdDiag, const double* dBdSrc, const double* dV, const double* dRAU, int valid,
                 double* dOut);
/* No counterpart of its own: the three deviceMatrixH calls of pressurePredictor (pEqn.cu:143-185) in one pass
 * over the matrix and the addressing. Host arrays of three device pointers per argument (component x, y, z);
 * valid (host, may be NULL: all valid) knocks components out. The values are three mfv_matrix_h calls'. */
int mfv_matrix_h3(mfv_ctx ctx, mfv_mesh mesh, const mfv_ldu* A, const double* const* dPsi,
                  const double* const* dSource, const double* const* dBdDiag, const double* const* dBdSrc,
                  const double* dV, const double* dRAU, const int* valid, double* const* dOut);
/* deviceRelaxDiag (src/cuda/device_simple.cuh:118-120; fvMatrix::relax, fv_matrix_ops.cuh:158-207) for
 * uncoupled patches: sumOff = sum|off-diagonal| of the row (+ dCycSumOff, may be NULL); D = diag +
 * cmptMagMax(ic) per boundary face; D = max(|D|, sumOff)/alpha; D -= cmptMin(ic) per boundary face.
 * dICmaxMag/dICmin (may be NULL: |dICbnd| and dICbnd) are per boundary face. Writes dRelaxedDiag = D and
 * dDelta = D - diag (either may be NULL). nComp (0..3) components also get their source update in the same
 * pass, dSource[k] += dDelta*dPsi[k] (UEqn.cu:753-760), host arrays of device pointers. Sums in the host
 * loops' order. MFV_ERR_INVALID for alpha <= 0 (the host skips the relaxation: leave the matrix as is). */
int mfv_relax_diag(mfv_ctx ctx, mfv_mesh mesh, const mfv_ldu* A, const double* dICbnd, double alpha,
                   const double* dCycSumOff, const double* dICmaxMag, const double* dICmin, double* dRelaxedDiag,
                   double* dDelta, int nComp, const double* const* dPsi, double* const* dSource);
/* deviceLinearUpwindCorr (src/cuda/device_simple.cuh:34-36; fvm.cuh:395-433): linearUpwind's deferred
 * correction of div(phi, U_k), corr = sum over internal faces of +-phi*((Cf - C_up) . grad_k(up)) (+ on the
 * owner, - on the neighbour; up the owner where phi > 0), for nComp (1..3) components in one pass. dGrad[k]
 * xyz per cell (mfv_grad_gauss's layout), dCf and dC mfv_mesh_geometry's. Uncoupled boundary faces add
 * nothing. The caller subtracts dCorr[k] from component k's source (fvMatrix += surfaceIntegrate). Face
 * order per cell, uncontracted: the host loop's bits. */
int mfv_linear_upwind_corr(mfv_ctx ctx, mfv_mesh mesh, const double* dPhi, const double* dCf, const double* dC,
                           int nComp, const double* const* dGrad, double* const* dCorr);
/* deviceScaledDev2T (src/cuda/device_divdevreff.cuh:102-103; linearViscousStress_cpp.cu:75-79): sigma = nu*dev2(T(gradU))
 * per cell, tensors component-major [(i*3 + j)*n + c] with gradU_ij = d(U_j)/d(x_i); dSigma must not
 * alias dGradU */
int mfv_scaled_dev2t(mfv_ctx ctx, int n, const double* dGradU, const double* dNu, double* dSigma);
/* deviceTensorDivSource (src/cuda/device_divdevreff.cuh:104-107; fvc.cu:867-905 before its /V): V*fvc::div(T) per
 * cell, linear interpolation on internal faces, dTbnd [9*nBFaces] component-major on the boundary faces not
 * marked in dIsEmpty (may be NULL); dSf xyz per face, internal then boundary (mfv_mesh_geometry's), dWeights
 * its weights. Outputs the raw sums per component. Face order per cell, uncontracted: the host loop's bits.
 * The boundary tensors (the corrected boundary gradient's stress) are the caller's. */
int mfv_tensor_div_source(mfv_ctx ctx, mfv_mesh mesh, const double* dWeights, const double* dSf,
                          const int* dIsEmpty, const double* dTcell, const double* dTbnd, double* dSrcX,
                          double* dSrcY, double* dSrcZ);
/* The sigma and tensor-divergence stages of deviceDivDevReff (src/cuda/device_divdevreff.cu:408-620) in one
 * pass: V*div(nu*dev2(T(gradU))) with each cell's stress formed from dGradU [9*nCells] and dNu where it is
 * read, never stored; dSigmaB the boundary faces' stress [9*nBFaces]. The values are mfv_scaled_dev2t's
 * followed by mfv_tensor_div_source's, to the bit. */
int mfv_div_dev_reff_explicit(mfv_ctx ctx, mfv_mesh mesh, const double* dWeights, const double* dSf,
                              const int* dIsEmpty, const double* dGradU, const double* dNu, const double* dSigmaB,
                              double* dSrcX, double* dSrcY, double* dSrcZ);
/* deviceCmptMaxMag3 / deviceCmptMin3 (src/cuda/device_simple.cuh:123-128): per element max(|a|,|b|,|c|) and
 * min(a,b,c) */
int mfv_cmpt_max_mag3(mfv_ctx ctx, int n, const double* dA, const double* dB, const double* dC, double* dOut);
int mfv_cmpt_min3(mfv_ctx ctx, int n, const double* dA, const double* dB, const double* dC, double* dOut);
/* relaxField (src/applications/solvers/simpleFoam/pEqn.cuh:154): p = pPrev + alpha*(p - pPrev) in place;
 * nothing unless 0 < alpha < 1 */
int mfv_relax_field(mfv_ctx ctx, int n, double* dP, const double* dPPrev, double alpha);
/* constrainHbyA (pEqn.cu:187-203, constrainHbyAKernel): HbyAb = Ub where dTakeU is set (U not assignable),
 * the owner cell's HbyA elsewhere; one component */
int mfv_constrain_hbya(mfv_ctx ctx, mfv_mesh mesh, const int* dTakeU, const double* dUb, const double* dHbyA,
                       double* dHbyAb);
/* deviceVectorFlux (src/cuda/device_simple.cuh:48): phi = (w U_own + (1-w) U_nei) . Sf per internal face;
 * dSf xyz per face, internal faces first (as mfv_mesh_geometry writes it) */
int mfv_vector_flux(mfv_ctx ctx, mfv_mesh mesh, const double* dWeights, const double* dSf, const double* dUx,
                    const double* dUy, const double* dUz, double* dPhi);
/* deviceBoundaryFlux (src/cuda/device_simple.cuh:59): phiB = Ub . Sf_b per boundary face, zero where dIsEmpty
 * (may be NULL) is set; dSf the same whole-mesh array as mfv_vector_flux's */
int mfv_boundary_flux(mfv_ctx ctx, mfv_mesh mesh, const double* dSf, const double* dUxb, const double* dUyb,
                      const double* dUzb, const int* dIsEmpty, double* dPhiB);
/* deviceAdjustPhi (src/cuda/device_simple.cuh:99; adjustPhi, pEqn_cpp.cu:80-133): scales the adjustable
 * (dAdjustable) outflow of dPhiB [nBFaces] in place so the boundary flux balances; dPhiInt [nInternalFaces]
 * (NULL with 0) is the internal flux of the normaliser 1e-300 + sum|phi|. massCorr/closedVolume may be NULL.
 * Synchronises. MFV_ERR_INVALID when the continuity error cannot be removed (OpenFOAM's FatalError; dPhiB
 * untouched). The sums are the context's reductions (tree, or sequential in the host's order). */
int mfv_adjust_phi(mfv_ctx ctx, int nBFaces, const int* dAdjustable, double* dPhiB, int nInternalFaces,
                   const double* dPhiInt, double* massCorr, int* closedVolume);

/* fvc::gaussGrad of a cell field with its patch values (src/finiteVolume/finiteVolume/fvc.cu:67-121):
 * dSf xyz per face (internal then boundary, as mfv_mesh_geometry writes it), dWeights per internal face,
 * patch values from dBcType/dBcValue (fixedValue: the value; zeroGradient: the cell's; empty: none);
 * dGrad xyz per cell */
int mfv_grad_gauss(mfv_ctx ctx, mfv_mesh mesh, const int* dBcType, const double* dBcValue, const double* dSf,
                   const double* dWeights, const double* dV, const double* dPsi, double* dGrad);
/* The explicit non-orthogonal correction of fvm::laplacian (fvm::laplacianCorrFlux and
 * laplacianNonOrthSource, src/finiteVolume/finiteVolume/fvm.cuh:170-222,276-310; pEqn_cpp.cu:223-235):
 * dFfc = (gammaf*magSf) * (corrVec . interpolate(grad)) per internal face, and when dSource is not NULL
 * dSource -= div(ffc) (face order per cell). Pass dFfc to mfv_flux_correct afterwards. */
int mfv_laplacian_corr(mfv_ctx ctx, mfv_mesh mesh, const double* dGrad, const double* dWeights,
                       const double* dCorrVec, const double* dGammaf, const double* dMagSf, double* dFfc,
                       double* dSource);
/* `limited k corrected` (deviceLaplacianCorrFluxLimited, src/cuda/device_fvm.cu:497-517; fvm.cuh:170-211):
 * as mfv_laplacian_corr with each face's correction scaled by min(k|orth| / ((1-k)|corr| + 1e-15), 1),
 * orth = dNonOrthDeltaCoeffs*(dVf[nei] - dVf[own]) the orthogonal snGrad of the corrected field dVf;
 * k outside (0, 1) is the unlimited correction */
int mfv_laplacian_corr_limited(mfv_ctx ctx, mfv_mesh mesh, const double* dGrad, const double* dWeights,
                               const double* dCorrVec, const double* dGammaf, const double* dMagSf,
                               const double* dNonOrthDeltaCoeffs, const double* dVf, double limitCoeff,
                               double* dFfc, double* dSource);

/* FvGeometry::build (src/finiteVolume/fvMesh/fv_geometry.cu:13-158) and the patches' deltaCoeffs
 * (fv_patch.cu:160-165) on the device. Faces are polygons given by dFaceOffsets [nFaces+1] into
 * dFaceVerts; all nInternal + nBFaces faces, boundary faces after the internal ones. Outputs: Cf, Sf
 * [3*nFaces], magSf [nFaces], C [3*nCells], V [nCells], weights, deltaCoeffs, nonOrthDeltaCoeffs
 * [nInternal], corrVec [3*nInternal], bDeltaCoeffs [nBFaces]. */
int mfv_mesh_geometry(mfv_ctx ctx, mfv_mesh mesh, const double* dPoints, const int* dFaceOffsets,
                      const int* dFaceVerts, double* dCf, double* dSf, double* dMagSf, double* dC, double* dV,
                      double* dWeights, double* dDeltaCoeffs, double* dNonOrthDeltaCoeffs, double* dCorrVec,
                      double* dBDeltaCoeffs);

/* ---------------------------------------------------------------------------------------------
 * Products and reductions.
 * ------------------------------------------------------------------------------------------- */
/* deviceAmul (src/matrices/lduMatrix/lduMatrix/device_ldu.cuh:241, device_spmv.cu:302) */
int mfv_amul(mfv_ctx ctx, mfv_mesh mesh, const mfv_ldu* A, const double* dPsi, double* dApsi);
/* residualKernel (device_spmv.cu:116): r = b - A psi */
int mfv_residual(mfv_ctx ctx, mfv_mesh mesh, const mfv_ldu* A, const double* dPsi, const double* dB, double* dR);
/* deviceDotInto / deviceSumMagInto (src/matrices/lduMatrix/lduMatrix/reductions.cu:315,387):
 * result to a device scalar */
int mfv_dot(mfv_ctx ctx, int n, const double* dA, const double* dB, double* dResult);
int mfv_sum_mag(mfv_ctx ctx, int n, const double* dA, double* dResult);
/* deviceNormFactorInto (src/matrices/lduMatrix/solvers/PCG/device_pcg.cu:133-170) */
int mfv_norm_factor(mfv_ctx ctx, mfv_mesh mesh, const mfv_ldu* A, const double* dPsi, const double* dB,
                    double* dNormFactor);

/* ---------------------------------------------------------------------------------------------
 * AMG hierarchy (pairwise agglomeration; the addressing is built on the host once per mesh).
 * ------------------------------------------------------------------------------------------- */
#define MFV_AGG_BRAE 0          /* brae's device-path pairing: forward greedy, sorted coarse faces */
#define MFV_AGG_FACEAREAPAIR 1  /* OpenFOAM faceAreaPair: alternating direction, first-appearance faces */
#define MFV_AGG_PAIR 2          /* the host CPU path's agglomeratePair (gamg.cu:41-172): alternating
                                   direction, sorted coarse faces; weights |Sf|; nCellsInCoarsestLevel 10 */

#define MFV_PREC_FP64 64
#define MFV_PREC_FP32 32
#define MFV_PREC_FP16 16        /* FP16 coefficients and vectors scaled per grid by powers of two, FP32
                                   arithmetic */

/* How the fine grid's matrix is read (mfv_amg_controls.fineLayout). Automatic: as diagonals when its faces
 * form a fixed stencil (mfv_amg_fine_stencil), else as sliced rows; rows: always sliced rows. Both sum
 * every row in face order; the choice changes speed, not results. */
#define MFV_LAYOUT_AUTO 0
#define MFV_LAYOUT_ROWS 1

typedef struct
{
    int agglomeration;          /* MFV_AGG_* */
    int nCellsInCoarsestLevel;  /* brae: coarsen while n > this (its TARGET, default 64);
                                   faceAreaPair: nCellsInCoarsestLevel (default 10) */
    int mergeLevels;            /* pairwise passes composed into one level (GAMG's mergeLevels, 1..4;
                                   0 or 1: every pass is a level, the reference's only setting) */
    int fineLayout;             /* MFV_LAYOUT_*; not part of the hierarchy's signature */
} mfv_amg_controls;

/* buildAMG (src/matrices/lduMatrix/preconditioners/GAMGPreconditioner/device_amg.cuh:393,
 * device_amg.cu:1346-1431), and for MFV_AGG_FACEAREAPAIR faceAreaPairGamgAgglomeration
 * (src/matrices/lduMatrix/solvers/GAMG/GAMGAgglomerations/pair_gamg_agglomeration_cpp.cu:286-345).
 * hFaceWeights: one per internal face (brae passes |Sf|; faceAreaPair its perturbed weights). */
/* faceAreaPair's face weights |(Sf_x, 1.01 Sf_y, 1.02 Sf_z)| / sqrt(|Sf|) for MFV_AGG_FACEAREAPAIR
 * (src/matrices/lduMatrix/solvers/GAMG/GAMGAgglomerations/pair_gamg_agglomeration_cpp.cu:301-311); host
 * arrays: hSf [3*nFaces] xyz-interleaved, hMagSf, hWeights [nFaces] */
int mfv_faceareapair_weights(int nFaces, const double* hSf, const double* hMagSf, double* hWeights);
int mfv_amg_build(mfv_ctx ctx, mfv_mesh mesh, const double* hFaceWeights, const mfv_amg_controls* controls,
                  mfv_amg* out);
int mfv_amg_destroy(mfv_amg amg);

/* Hierarchy persistence (amgHierarchySignature, device_amg.cu:1262-1285; writeAMGCache/readAMGCache,
 * device_amg_cache.cu:176-326; buildOrLoadAMG, device_amg.cuh:448; cloneAMG, device_amg_cache.cu:70).
 * The structure only (levels, coarse addressing, face maps), never matrix values. */
#define MFV_CACHE_LOADED 0
#define MFV_CACHE_ABSENT 1
#define MFV_CACHE_OTHER_MESH_OR_BUILD 2
#define MFV_CACHE_UNREADABLE 3
/* 64-bit fingerprint of what a build depends on: the internal addressing, the face weights, the
 * controls and this library's build version. Host only, no context. */
int mfv_amg_signature(int nCells, int nFaces, const int* hOwner, const int* hNeighbour, const double* hFaceWeights,
                      const mfv_amg_controls* controls, unsigned long long* signature);
/* writes the hierarchy's structure to path (a temporary file renamed over it) */
int mfv_amg_save(mfv_ctx ctx, mfv_amg amg, const char* path);
/* loads <cacheDir>/.mfv_amgcache when it holds this mesh's signature, else builds (and with writeCache
 * writes the file); *how (may be NULL) says what the file held: MFV_CACHE_* */
int mfv_amg_build_or_load(mfv_ctx ctx, mfv_mesh mesh, const double* hFaceWeights, const mfv_amg_controls* controls,
                          const char* cacheDir, int writeCache, mfv_amg* out, int* how);
/* a second hierarchy of the same structure on the same mesh; its values start unset (amg_galerkin) */
int mfv_amg_clone(mfv_ctx ctx, mfv_amg src, mfv_amg* out);
/* number of grids (the mesh plus its coarse levels) */
int mfv_amg_num_grids(mfv_amg amg, int* nGrids);
int mfv_amg_grid_size(mfv_amg amg, int grid, int* nCells, int* nFaces);
/* Inspection, no counterpart in the reference: the fine grid's fixed stencil, if it has one. When every
 * face's neighbour - owner is one of at most four positive offsets and the faces are in (owner,
 * neighbour) order, the outer product reads the fine matrix as diagonals (each coefficient stored once):
 * *nOffsets is their number and hOffsets[0..*nOffsets) (may be NULL; room for four) the ascending
 * offsets. 0 offsets: the fine matrix is read as sliced rows. */
int mfv_amg_fine_stencil(mfv_amg amg, int* nOffsets, int* hOffsets);
/* How the coarsest grids' cycle runs, no counterpart in the reference (its small levels run one kernel
 * per step). Below 4096 cells the grids from the first such one down run in one workgroup (the tail):
 * mode 2 (default) stages the tail's coarsest grids -- as many as fit the device's LDS per workgroup --
 * in that workgroup's LDS and reads the others where they are; mode 1 reads them all where they are;
 * mode 0 runs every grid with its own kernels. All modes give the same bits (the scaled correction
 * keeps mode 1's kernel). mode -1 only queries. *tailStart: the tail's first grid (0: no tail);
 * *ldsGrid: the first grid staged (0: none); *ldsBytes: the LDS the staged grids take, in the precision
 * of the last mfv_amg_galerkin; any of the three may be NULL. */
int mfv_amg_coarse_mode(mfv_amg amg, int mode, int* tailStart, int* ldsGrid, int* ldsBytes);
/* host copies of the step from grid `level` to grid `level+1` (any pointer may be NULL):
 * restrict map [nCells(level)], coarse owner/neighbour [nFaces(level+1)], faceRestrict and faceFlip
 * [nFaces(level)] -- AMGLevel's map/cOwn/cNei/faceRestrict/faceFlip (device_amg.cuh:119-182) */
int mfv_amg_level_addressing(mfv_amg amg, int level, int* hMap, int* hCoarseOwner, int* hCoarseNeighbour,
                             int* hFaceRestrict, signed char* hFaceFlip);
/* amgGalerkin (device_amg.cuh:454, device_amg.cu:1693-1846): coarse matrices from the folded fine
 * matrix (symmetric only: dFineLower must be NULL or equal dFineUpper), the V-cycle copies in
 * `precision`, the coarsest dense inverse. */
int mfv_amg_galerkin(mfv_ctx ctx, mfv_amg amg, const double* dFineDiag, const double* dFineUpper,
                     const double* dFineLower, int precision);
/* host copies of a grid's FP64 Galerkin values (grid >= 1): diag [nCells], upper [nFaces]; synchronises */
int mfv_amg_grid_values(mfv_ctx ctx, mfv_amg amg, int grid, double* hDiag, double* hUpper);
/* amgVCycleApply (device_amg_vcycle.cu:1793-1875): dZ = one V(nPre,nPost) cycle applied to dR;
 * scaleCorrection as in mfv_pcg_controls */
int mfv_amg_vcycle(mfv_ctx ctx, mfv_amg amg, const double* dR, double* dZ, int nPreSweeps, int nPostSweeps,
                   int scaleCorrection);

/* ---------------------------------------------------------------------------------------------
 * AMG-preconditioned CG (deviceAMGPCG, device_amg.cuh:550-553; device_amg_pcg.cu:638-848).
 * ------------------------------------------------------------------------------------------- */
typedef struct
{
    double tolerance;      /* on OpenFOAM's normalised residual sum|r|/normFactor */
    double relTol;
    int maxIter, minIter;
    int precision;         /* MFV_PREC_*: the V-cycle's precision; must match mfv_amg_galerkin's */
    int nPreSweeps, nPostSweeps;
    int flexible;          /* 1: Polak-Ribiere+ beta (flexBetaK); 0: Fletcher-Reeves */
    int smoother;          /* 0: weighted Jacobi (omega 0.8), the only value carried */
    int checkEvery;        /* host polls the device stop flag every k iterations (result independent of k) */
    int nVcycles;          /* 1 (only value carried) */
    int scaleCorrection;   /* 1: each grid's coarse correction scaled by (r.e)/(e.Ae) (the reference's
                              corrScaling, device_amg_vcycle.cu:1113-1126), fused with the first
                              post-sweep; use with flexible = 1 */
    int captureVcycle;     /* 1: the iterations between two polls (checkEvery, rounded up to even) are
                              captured once into a HIP graph and replayed (the reference's captureVcycle,
                              deviceAMGPCGGraph, device_amg_pcg.cu:183-411); the result is the same as
                              without. Needs a context stream other than the legacy default stream. */
} mfv_pcg_controls;

/* SolverPerformance (src/OpenFOAM/matrices/pcg.cuh:16-21) */
typedef struct
{
    double initialResidual, finalResidual;
    int nIterations;
    int converged;
} mfv_perf;

int mfv_pcg_default_controls(mfv_pcg_controls* controls);
/* A: the folded symmetric fine matrix the hierarchy's last mfv_amg_galerkin saw; dNormFactor: a device
 * scalar (deviceNormFactorInto's) or NULL to compute OpenFOAM's normFactor; psi updated in place.
 * Synchronises. */
int mfv_amg_pcg_solve(mfv_ctx ctx, mfv_mesh mesh, mfv_amg amg, const mfv_ldu* A, const double* dB, double* dPsi,
                      const double* dNormFactor, const mfv_pcg_controls* controls, mfv_perf* perf);

/* Small solves batched into one launch, no counterpart in the reference (a solve of tens of thousands
 * of cells is latency-bound on the per-grid kernels): every solve runs whole in one workgroup -- the
 * AMG-PCG of mfv_amg_pcg_solve with OpenFOAM's normFactor, the V-cycle in controls.precision with or
 * without the scaled correction -- and all of them in the same launch, each to its own tolerance. A
 * solve's result is the same bits in any batch. Each hierarchy appears once per call, its Galerkin done
 * in that precision; checkEvery/captureVcycle are ignored. The meshes and
 * hierarchies may belong to other contexts on ctx's device (the launch runs on ctx's stream; their own
 * streams must have nothing in flight on them meanwhile). Synchronises. */
typedef struct
{
    mfv_mesh mesh;
    mfv_amg amg;
    const double* dDiag;   /* the folded FP64 diagonal; the upper is the one the hierarchy's Galerkin saw */
    const double* dB;
    double* dPsi;          /* initial guess in, solution out */
    mfv_pcg_controls controls;
} mfv_batch_solve;
int mfv_amg_pcg_solve_batch(mfv_ctx ctx, int n, const mfv_batch_solve* solves, mfv_perf* perfs);

/* ---------------------------------------------------------------------------------------------
 * OpenFOAM's sequential sweeps, level-scheduled on the device and equal to the host loops to the bit.
 * ------------------------------------------------------------------------------------------- */
/* DICPreconditioner: rD (calcReciprocalD) then w = forward/backward substitution of r
 * (src/matrices/lduMatrix/solvers/GAMG/gamg_solver_cpp.cu:194-231; pcg.cu:6-149); synchronises */
int mfv_dic_precondition(mfv_ctx ctx, mfv_mesh mesh, const mfv_ldu* A, const double* dR, double* dW);
/* nSweeps forward Gauss-Seidel sweeps on A psi = b in place (GaussSeidelSmoother,
 * src/OpenFOAM/matrices/smooth_solver_cpp.cu:87-155; gamg.cu:187-205) */
int mfv_gauss_seidel(mfv_ctx ctx, mfv_mesh mesh, const mfv_ldu* A, const double* dB, double* dPsi, int nSweeps);
/* OpenFOAM's PCG with the DIC preconditioner (the reference's `pcg`, src/OpenFOAM/matrices/pcg.cu:6-149):
 * controls' tolerance, relTol, maxIter, minIter and checkEvery are read; dNormFactor as in
 * mfv_amg_pcg_solve. Synchronises. */
int mfv_pcg_dic_solve(mfv_ctx ctx, mfv_mesh mesh, const mfv_ldu* A, const double* dB, double* dPsi,
                      const double* dNormFactor, const mfv_pcg_controls* controls, mfv_perf* perf);
/* symGaussSeidel: each of nSweeps is the forward sweep, then the reverse walk from the forward walk's
 * accumulated lower-neighbour terms (src/OpenFOAM/matrices/smooth_solver_cpp.cu:87-163) */
int mfv_sym_gauss_seidel(mfv_ctx ctx, mfv_mesh mesh, const mfv_ldu* A, const double* dB, double* dPsi, int nSweeps);

/* ---------------------------------------------------------------------------------------------
 * OpenFOAM's GAMG (src/matrices/lduMatrix/solvers/GAMG/gamg_solver_cpp.cuh): the V-cycle of
 * GAMGSolver -- residual restricted by summation, pre-sweeps (default none), the coarsest level solved
 * by PCG-DIC from zero to the GAMG's tolerance, injection, the correction scaling, post-sweeps
 * min(nPostSweeps + postSweepsLevelMultiplier*level, maxPostSweeps), the finest level's correction
 * scaled against the residual and psi smoothed nFinestSweeps times -- in FP64 on the hierarchy's grids
 * (build it with MFV_AGG_OPENFOAM for OpenFOAM's faceAreaPair levels). With the context's sequential
 * reductions a solve equals the host solve to the bit.
 * ------------------------------------------------------------------------------------------- */
#define MFV_SMOOTHER_GAUSS_SEIDEL 0
#define MFV_SMOOTHER_SYM_GAUSS_SEIDEL 1
#define MFV_SMOOTHER_DIC 2
#define MFV_SMOOTHER_DIC_GAUSS_SEIDEL 3

/* GamgControls (gamg_solver_cpp.cuh:67-85) and GamgPreconditionerControls::nVcycles (:127-131) */
typedef struct
{
    int smoother;                       /* MFV_SMOOTHER_* */
    double tolerance, relTol;           /* the solver's; as a preconditioner, the coarsest solve's */
    int maxIter, minIter;
    int nPreSweeps, preSweepsLevelMultiplier, maxPreSweeps;
    int nPostSweeps, postSweepsLevelMultiplier, maxPostSweeps;
    int nFinestSweeps;
    int scaleCorrection;
    int nVcycles;                       /* V-cycles per preconditioner application */
} mfv_gamg_controls;

/* GAMGSolver.C's defaults: GaussSeidel, 1e-6, 0, 1000, 0, 0/1/4, 2/1/4, 2, scaled, nVcycles 2 */
int mfv_gamg_default_controls(mfv_gamg_controls* controls);
/* gamgSolveFolded (gamg_solver_cpp.cuh:185-195): A the folded symmetric matrix; the FP64 coarse matrices
 * are agglomerated from it here; psi updated in place. Synchronises once per V-cycle. */
int mfv_gamg_solve(mfv_ctx ctx, mfv_mesh mesh, mfv_amg amg, const mfv_ldu* A, const double* dB, double* dPsi,
                   const mfv_gamg_controls* controls, mfv_perf* perf);
/* pcgGamgSolve (gamg_solver_cpp.cuh:197-216): PCG with the GAMG preconditioner; tolerance..minIter are
 * the PCG's. Synchronises once per iteration. */
int mfv_pcg_gamg_solve(mfv_ctx ctx, mfv_mesh mesh, mfv_amg amg, const mfv_ldu* A, const double* dB, double* dPsi,
                       double tolerance, double relTol, int maxIter, int minIter, const mfv_gamg_controls* controls,
                       mfv_perf* perf);
/* gamg (src/OpenFOAM/matrices/gamg.cu:267-400): the GAMG the host CPU path's SIMPLE step solves p with
 * (simpleFoam_cpp.cu:156), on a hierarchy built with MFV_AGG_PAIR on |Sf| (mergeLevels 1): the residual
 * restricted straight down, the coarsest level by dense Gaussian elimination with partial pivoting, then per
 * level injection, its correction scaling and nPostSweeps Gauss-Seidel sweeps; the finest correction scaled,
 * added and smoothed nFinestSweeps times. The folded symmetric system (A with lower == upper), psi in place.
 * With the context's sequential reductions the solve is the host loop's to the bit. Synchronises per cycle. */
int mfv_gamg_ref_solve(mfv_ctx ctx, mfv_mesh mesh, mfv_amg amg, const mfv_ldu* A, const double* dB, double* dPsi,
                       double tolerance, double relTol, int maxIter, int minIter, int nPostSweeps, int nFinestSweeps,
                       mfv_perf* perf);

/* ---------------------------------------------------------------------------------------------
 * PBiCGStab for an asymmetric matrix with its uncoupled boundary coefficients (the reference's
 * `pbicgstab`, src/OpenFOAM/matrices/pbicgstab.cu:6-150, the solver solveVector runs per momentum
 * component, src/OpenFOAM/matrices/solve_vector.cuh:27-116). The raw matrix comes in with internalCoeffs
 * and boundaryCoeffs per boundary face (dIC, dBC, in the mesh's boundary-face order); the solve folds them
 * (diag + iC, source + bC) as the reference does. MFV_PRECOND_DILU is the reference's DILU preconditioner,
 * walked level by level in the host loop's order (with the context's sequential reductions the solve is the
 * host solve to the bit); MFV_PRECOND_DIAGONAL divides by the folded diagonal. Synchronises every
 * checkEvery iterations.
 * ------------------------------------------------------------------------------------------- */
#define MFV_PRECOND_DILU 0
#define MFV_PRECOND_DIAGONAL 1
typedef struct
{
    double tolerance, relTol;
    int maxIter, minIter;
    int preconditioner;    /* MFV_PRECOND_* */
    int checkEvery;
} mfv_bicg_controls;
/* the reference's defaults for a momentum component (cpu::StepInput): 1e-10, 0, 1000, 0, DILU, poll every 4 */
int mfv_bicg_default_controls(mfv_bicg_controls* controls);
int mfv_pbicgstab_solve(mfv_ctx ctx, mfv_mesh mesh, const mfv_ldu* A, const double* dSource, const double* dIC,
                        const double* dBC, double* dPsi, const mfv_bicg_controls* controls, mfv_perf* perf);
/* No counterpart in the reference. For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc): mfv_pbicgstab_solve with its work vectors in
 * a static device array of the library (MFV_STATIC_BICG_MAX_CELLS cells, about 44 MiB per GPU, held until
 * the process ends) instead of the mesh's buffers; it calls no hipMalloc or hipFree. Same arguments, same
 * kernels, same bits. Larger meshes are refused (MFV_ERR_INVALID). DILU needs the mesh's level schedule,
 * which a first mfv_pbicgstab_solve, mfv_pcg_dic_solve or sweep on the mesh builds; the diagonal
 * preconditioner needs nothing. Every call on one GPU shares that GPU's array: calls that overlap on one
 * GPU (two contexts, two threads) race. */
#define MFV_STATIC_BICG_MAX_CELLS 524288
int mfv_pbicgstab_solve_noalloc(mfv_ctx ctx, mfv_mesh mesh, const mfv_ldu* A, const double* dSource,
                                const double* dIC, const double* dBC, double* dPsi, const mfv_bicg_controls* controls,
                                mfv_perf* perf);

/* ---------------------------------------------------------------------------------------------
 * One SIMPLE iteration on the device: cpu::simpleStep (src/applications/solvers/simpleFoam/
 * simpleFoam_cpp.cu:15-237; its device twin gpu::simpleStep, simpleFoam.cu:92-420) for uncoupled patches
 * without MRF, fvOptions or SIMPLEC. UEqn = div(phi,U) (upwind, optionally linearUpwind's deferred
 * correction) - laplacian(nuEff,U) (optionally corrected) - div(nuEff dev2(T(grad U))), relaxed; the
 * momentum predictor with PBiCGStab per component; rAU, HbyA, constrainHbyA, phiHbyA, adjustPhi (when
 * pRefCell >= 0); nNonOrthCorr + 1 pressure solves by AMG-PCG with phi corrected after the last; p relaxed;
 * U corrected. Boundary conditions per boundary face: MFV_BC_* (noSlip = fixedValue 0).
 * Geometry: the arrays mfv_mesh_geometry writes (device).
 * ------------------------------------------------------------------------------------------- */
#define MFV_BC_ZERO_GRADIENT 0
#define MFV_BC_FIXED_VALUE 1
#define MFV_BC_EMPTY 2
typedef struct mfv_simple_s* mfv_simple;
typedef struct
{
    const double *dSf, *dCf, *dMagSf;   /* internal then boundary faces (xyz for Sf, Cf) */
    const double *dC, *dV;              /* cells */
    const double *dWeights, *dDeltaCoeffs, *dNonOrthDeltaCoeffs, *dCorrVec;  /* internal faces */
    const double* dBDeltaCoeffs;        /* boundary faces */
} mfv_geometry;
typedef struct
{
    double nu;                          /* laminar viscosity (nuEff when the step is given none) */
    double relaxU, relaxP;              /* relaxationFactors: equations U, fields p */
    int momentumPredictor;
    int linearUpwind;                   /* div(phi,U) Gauss linearUpwind grad(U) (else upwind) */
    int correctedLaplacian;             /* laplacianSchemes Gauss linear corrected (nonOrthDeltaCoeffs) */
    int nNonOrthCorr;                   /* nNonOrthogonalCorrectors, 0..7 */
    int pRefCell;                       /* -1: p is fixed on a patch (no reference, no adjustPhi) */
    double pRefValue;
    mfv_bicg_controls U;                /* the momentum components' PBiCGStab */
    mfv_pcg_controls p;                 /* the pressure's AMG-PCG */
    int mergeLevels;                    /* pairwise passes per pressure AMG level */
    int pSolver;                        /* MFV_P_AMG_PCG, or MFV_P_GAMG_REF: the host CPU path's gamg.cu GAMG
                                           (nCellsInCoarsest 10, V(0,2,2)) to p's tolerance..minIter */
} mfv_simple_controls;
#define MFV_P_AMG_PCG 0
#define MFV_P_GAMG_REF 1
typedef struct
{
    double *dUx, *dUy, *dUz;            /* cells */
    double *dUbx, *dUby, *dUbz;         /* boundary faces: U's patch values */
    double *dP, *dPb;                   /* p in cells and on the boundary faces */
    double *dPhi, *dPhiB;               /* face fluxes: internal, boundary */
} mfv_simple_fields;
typedef struct
{
    double UInitial[3], UFinal[3];      /* per component; OpenFOAM reports the max over solved components */
    int UIters[3];
    double pInitial, pFinal;            /* the first solve's initial, the last solve's final residual */
    int pIters[8];                      /* per pressure solve */
    double massCorr;                    /* adjustPhi's factor (1 without a reference cell) */
} mfv_simple_result;
/* cpu::StepInput's defaults: nu 1e-5, U 0.7, p 0.3, momentum predictor, upwind, uncorrected, no
 * non-orthogonal correctors, no reference cell; PBiCGStab-DILU to 1e-10; AMG-PCG FP64 to 1e-10 */
int mfv_simple_default_controls(mfv_simple_controls* controls);
/* builds the step's workspace and the pressure hierarchy (synchronises); hUType/hPType: MFV_BC_* per
 * boundary face, hUValue 3 per face (xyz), hPValue 1 per face */
int mfv_simple_create(mfv_ctx ctx, mfv_mesh mesh, const mfv_geometry* geometry, const int* hUType,
                      const double* hUValue, const int* hPType, const double* hPValue,
                      const mfv_simple_controls* controls, mfv_simple* out);
int mfv_simple_destroy(mfv_simple simple);
/* one iteration in place on the fields; dNuEff/dNuEffB (nullable): nuEff per cell and boundary face in
 * place of nu (nu + nut of the turbulence model's last correction). Synchronises. */
int mfv_simple_step(mfv_simple simple, const mfv_simple_fields* fields, const double* dNuEff,
                    const double* dNuEffB, mfv_simple_result* result);

/* ---------------------------------------------------------------------------------------------
 * The standard k-epsilon model as the reference's host SIMPLE driver corrects it after each iteration:
 * kepsilon::correct (src/TurbulenceModels/turbulenceModels/RAS/kEpsilon/k_epsilon.cuh:221-337; the
 * device twin deviceKEpsilonCorrect, src/cuda/device_kepsilon.cu:1139-1378). Production G = nut*GbyNu,
 * the epsilonWallFunction's near-wall epsilon and production (corner weights 1/nWallFaces) and the
 * nutkWallFunction on the faces of wall patches, the epsilon equation (upwind convection, laplacian of
 * nut/sigmaEps + nu, C2 eps/k destruction, C1 Cmu k GbyNu production, the (2/3 C1 - C3) divU SuSp term),
 * relaxed and constrained at the near-wall cells (fvMatrix::setValues), solved by PBiCGStab and bounded;
 * then k with the new epsilon; then nut = Cmu k^2/epsilon. Incompressible, upwind convection, uncoupled
 * patches; k and epsilon carry zeroGradient values on the wall-function faces.
 * ------------------------------------------------------------------------------------------- */
typedef struct mfv_kepsilon_s* mfv_kepsilon;
typedef struct
{
    double Cmu, C1, C2, C3, sigmaK, sigmaEps, kappa, E;   /* kepsilon_coeffs.cuh:11-22 */
} mfv_kepsilon_coeffs;
typedef struct
{
    double nu;                          /* laminar viscosity */
    double relaxEps, relaxK;            /* relaxationFactors: equations epsilon, k */
    mfv_bicg_controls solver;           /* the two solves' PBiCGStab */
    mfv_kepsilon_coeffs coeffs;
} mfv_kepsilon_controls;
typedef struct
{
    const double *dUx, *dUy, *dUz, *dUbx, *dUby, *dUbz;  /* U in cells and on the boundary faces */
    const double *dPhi, *dPhiB;                          /* face fluxes: internal, boundary */
    double *dK, *dKb, *dEps, *dEpsb, *dNut, *dNutb;      /* cells, boundary faces (in place) */
} mfv_kepsilon_fields;
typedef struct
{
    double epsInitial, epsFinal, kInitial, kFinal;
    int epsIters, kIters, nWallCells;
} mfv_kepsilon_result;
/* OpenFOAM v2412's coefficients, nu 1e-5, no relaxation (factor 1), PBiCGStab-DILU to 1e-10 */
int mfv_kepsilon_default_controls(mfv_kepsilon_controls* controls);
/* nearWallDist (src/finiteVolume/fvMesh/near_wall_dist.cuh:120-177) on the host: y per boundary face of the
 * faces flagged in hIsWall; hFaceOffsets/hFaceVerts over every face (internal then boundary), hC the cell
 * centres (xyz). Needs no context. */
int mfv_near_wall_dist(int nInternalFaces, int nBFaces, const double* hPoints, const int* hFaceOffsets,
                       const int* hFaceVerts, const int* hBFaceCell, const double* hC, const int* hIsWall,
                       double* hY);
/* hIsWall, hYWall (mfv_near_wall_dist), hKType/hEpsType (MFV_BC_*) and values per boundary face */
int mfv_kepsilon_create(mfv_ctx ctx, mfv_mesh mesh, const mfv_geometry* geometry, const int* hIsWall,
                        const double* hYWall, const int* hKType, const double* hKValue, const int* hEpsType,
                        const double* hEpsValue, const mfv_kepsilon_controls* controls, mfv_kepsilon* out);
int mfv_kepsilon_destroy(mfv_kepsilon ke);
/* nuEff = nu + nut per cell and boundary face (the momentum equation's viscosity) */
int mfv_kepsilon_nu_eff(mfv_kepsilon ke, const double* dNut, const double* dNutB, double* dNuEff, double* dNuEffB);
/* one correct() in place on k, epsilon and nut. Synchronises. */
int mfv_kepsilon_correct(mfv_kepsilon ke, const mfv_kepsilon_fields* fields, mfv_kepsilon_result* result);
/* one turbulent SIMPLE iteration, as cpu::simpleStep orders it (simpleFoam_cpp.cu:27-46, 208-233): nuEff
 * from nut, the laminar step's UEqn, pressure and corrections with it, then the model's correct() from the
 * corrected U and phi. The U/phi members of keFields are ignored (the step's fields are used). */
int mfv_simple_step_kepsilon(mfv_simple simple, const mfv_simple_fields* fields, mfv_kepsilon ke,
                             const mfv_kepsilon_fields* keFields, mfv_simple_result* result,
                             mfv_kepsilon_result* keResult);

/* ---------------------------------------------------------------------------------------------
 * Serving: a stream of pressure-Poisson requests (no counterpart in the reference, which runs one
 * case per process). Each request is a box of nx*ny*nz hex cells (lengths lx, ly, lz; interior points
 * jittered by +-jitter of the spacing), solved nSteps times for right-hand sides -div(U) of seeded
 * random velocities, with the 7-point FP64 laplacian, zeroGradient walls and reference cell 0, by
 * AMG-PCG to `tolerance` in the server's cycle precision. Requests are admitted in arrival order when
 * they have arrived (seconds after mfv_server_run starts), a slot is free and their HBM estimate fits
 * the budget; each slot is a host thread with its own stream, so admitted requests overlap on the
 * device. A request's mesh, matrix and hierarchy stay resident across its steps, and -- with
 * mfv_server_set_resident -- between requests: a request for a box that a slot still holds (the same
 * nx..seed) is handed to that slot when it is idle and solves on the system it keeps, without building
 * anything; its result is the same bits as on a new build.
 * ------------------------------------------------------------------------------------------- */
typedef struct mfv_server_s* mfv_server;
typedef struct
{
    int nx, ny, nz;
    double lx, ly, lz;
    double jitter;
    unsigned long long seed;
    int nSteps;
    double tolerance;
    double arrival;
} mfv_serve_request;
typedef struct
{
    double arrival, admitted, finished;   /* seconds after the run started */
    int status;                           /* MFV_OK or the request's error */
    int nCells, nSteps, totalIterations;
    double finalResidual;                 /* of the last step */
    unsigned long long fingerprint;       /* of every step's solution bits */
    int resident;                         /* 1: solved on a system a slot kept from an earlier request */
    int batched;                          /* 1: its solves ran in batches (mfv_server_set_batch) */
    double built;                         /* seconds after the run started when its system was ready */
} mfv_serve_result;
int mfv_server_create(int device, int nSlots, double hbmBudgetBytes, int precision, mfv_server* out);
int mfv_server_destroy(mfv_server server);
/* bytes of the HBM budget that hold systems (mesh, geometry, matrix, hierarchy, work vectors) between
 * requests, an equal share per slot, least recently used dropped first; held back from admission for
 * the whole run. 0 (the default) keeps none. */
int mfv_server_set_resident(mfv_server server, double residentBytes);
/* requests of at most maxCells cells solve in batches: each slot hands its solve (right-hand side ready)
 * to the server's batcher, which launches every solve waiting as one mfv_amg_pcg_solve_batch (one
 * workgroup each, the same cycle) and the next arrivals in the next launch. 0 (the default): every
 * request on the per-grid kernels. */
int mfv_server_set_batch(mfv_server server, int maxCells);
/* the HBM estimate admission charges a request */
double mfv_server_request_bytes(mfv_server server, const mfv_serve_request* request);
/* runs the stream to the end; returns MFV_OK or the first request error (each result has its status) */
int mfv_server_run(mfv_server server, const mfv_serve_request* requests, int nRequests, mfv_serve_result* results);

#ifdef __cplusplus
}
#endif
#endif
