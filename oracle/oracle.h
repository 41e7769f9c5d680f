/*
 * oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C FP64 restatement of the reference's CPU path for the finite-volume pressure
 * assembly and solve. It is the checker the parity tests, smoke() and bench.py's CPU-baseline
 * leg compare the HIP library against; nothing in the product links, loads or calls it.
 * Every function names the reference file:line it restates (paths relative to the
 * reference root).
 *
 * Conventions: int32 labels, double scalars, structure-of-arrays, vectors as xyz triples
 * (3*n doubles). Internal faces are in OpenFOAM's upper-triangular order (sorted by owner,
 * owner < neighbour). A "folded" system has the boundary's internalCoeffs in diag and its
 * boundaryCoeffs in the source.
 *
 * Parity pins (what checks each function; see tests/test_oracle_*.py):
 *   geometry + laplacian ........ validation/matrixDump/matrix.dat (OpenFOAM dump, 1e-12)
 *   amul ........................ validation/matrixDump/amul.dat (1e-12) + dense brute force
 *   pcg_dic ..................... validation/matrixDump/solution.dat (224 iterations exactly)
 *   of agglomeration ............ demo/amgpcg/case20/trace/of_hierarchy.csv (bit-exact maps)
 *   gamg_solve (OpenFOAM GAMG) .. validation/matrixDumpGAMG/solution.dat (101 iterations)
 *   brae agglomeration .......... demo/amgpcg/case20/trace/brae_hierarchy.csv (bit-exact maps)
 *   brae galerkin ............... demo/amgpcg/case20/trace/brae_levels.csv
 *   brae amg-pcg ................ demo/amgpcg/case20/trace/brae_cycles.csv (per-iteration psi)
 *   div upwind .................. validation/matrixDump/matrix_div.dat
 *   fold / setReference / flux .. closed forms and definitions (tests state them)
 */
#ifndef MFV_ORACLE_H
#define MFV_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- geometry: src/finiteVolume/fvMesh/fv_geometry.cu:13-134 (OpenFOAM primitiveMeshTools) ---- */
int orc_geometry(int nPoints, const double* points, int nFaces, const int* faceOffsets, const int* faceVerts,
                 int nCells, int nInternal, const int* owner, const int* neighbour,
                 double* Cf, double* Sf, double* magSf, double* C, double* V,
                 double* weights, double* deltaCoeffs, double* nonOrthDeltaCoeffs, double* corrVec);

/* boundary-face deltaCoeffs 1/|nHat.(Cf - C)|: src/finiteVolume/fvMesh/fv_patch.cu:144-166 */
void orc_patch_delta(int nBFaces, const int* bFaceCell, const double* bCf, const double* bSf,
                     const double* bMagSf, const double* C, double* bDeltaCoeffs);

/* ---- lduAddressing: src/matrices/lduMatrix/lduMatrix/device_ldu.cuh:109-146 ---- */
int orc_ldu_addressing(int nCells, int nFaces, const int* owner, const int* neighbour,
                       int* ownerStart, int* losortStart, int* losort);

/* ---- fvm::laplacian interior: src/finiteVolume/finiteVolume/fvm.cuh:113-151 ---- */
void orc_laplacian(int nCells, int nFaces, const int* owner, const int* neighbour,
                   const double* dc, const double* gammaf, const double* magSf,
                   double* diag, double* upper, double* lower);

/* Boundary coefficients of a patch field, per boundary face (bcType: 0 zeroGradient, 1 fixedValue,
 * 2 empty -> no contribution). fvm.cuh:43-103 (laplacian) and fvm.cuh:556-619 (div). */
enum { ORC_BC_ZERO_GRADIENT = 0, ORC_BC_FIXED_VALUE = 1, ORC_BC_EMPTY = 2 };
void orc_laplacian_bcoeffs(int nBFaces, const int* bcType, const double* bcValue, const double* gammab,
                           const double* bMagSf, const double* bDeltaCoeffs, double* iC, double* bC);

/* ---- fvm::div(phi, psi) upwind interior + boundary: fvm.cuh:476-554, 556-619 ---- */
void orc_div_upwind(int nCells, int nFaces, const int* owner, const int* neighbour, const double* phi,
                    double* diag, double* upper, double* lower);
void orc_div_linear(int nCells, int nFaces, const int* owner, const int* neighbour, const double* phi,
                    const double* weights, double* diag, double* upper, double* lower);
void orc_div_bcoeffs(int nBFaces, const int* bcType, const double* bcValue, const double* phib,
                     double* iC, double* bC);

/* ---- fold + setReference: gamg_solver_cpp.cu:769-797; device_simple.cuh:104-106 ---- */
void orc_fold(int nCells, const double* diag, const double* source, int nBFaces, const int* bFaceCell,
              const int* bcType, const double* iC, const double* bC, double* diagC, double* b);
void orc_set_reference(double* diag, double* b, int refCell, double refValue);

/* ---- LDU products and reductions: ldu_spmv.cu:5-24; gamg_solver_cpp.cu:62-159 ---- */
void orc_amul(int nCells, int nFaces, const int* owner, const int* neighbour, const double* diag,
              const double* upper, const double* lower, const double* psi, double* Apsi);
void orc_residual(int nCells, int nFaces, const int* owner, const int* neighbour, const double* diag,
                  const double* upper, const double* lower, const double* psi, const double* source, double* r);
double orc_norm_factor(int nCells, int nFaces, const int* owner, const int* neighbour, const double* diag,
                       const double* upper, const double* lower, const double* psi, const double* source);
double orc_sum_mag(int n, const double* x);
double orc_sum_prod(int n, const double* a, const double* b);

/* flux = upper*psi[nei] - lower*psi[own]: pEqn_cpp.cuh:134-139, device_simple.cuh:51-52 */
void orc_face_flux(int nFaces, const int* owner, const int* neighbour, const double* upper,
                   const double* lower, const double* psi, double* flux);

/* ---- non-orthogonal correction: fvc.cu:67-121 (Gauss grad), fvm.cuh:170-222 + 276-299 (face flux and
 *      its divergence off the source), fv_matrix_ops.cuh:33-50 + pEqn_cpp.cu:255-258 (phi) ---- */
void orc_grad_gauss(int nCells, int nFaces, const int* owner, const int* neighbour, int nBFaces,
                    const int* bFaceCell, const int* bcType, const double* bcValue, const double* Sf,
                    const double* weights, const double* V, const double* psi, double* grad);
void orc_laplacian_corr(int nCells, int nFaces, const int* owner, const int* neighbour, const double* grad,
                        const double* weights, const double* corrVec, const double* gammaf, const double* magSf,
                        double* ffc, double* source);
void orc_laplacian_corr_limited(int nCells, int nFaces, const int* owner, const int* neighbour, const double* grad,
                                const double* weights, const double* corrVec, const double* gammaf,
                                const double* magSf, const double* nonOrthDc, const double* vf, double limitCoeff,
                                double* ffc, double* source);
void orc_flux_correct(int nFaces, const int* owner, const int* neighbour, const double* upper, const double* lower,
                      const double* psi, const double* ffc, const double* phiHbyA, double* phi);

typedef struct { double initialResidual, finalResidual; int nIterations; } orc_perf;

/* ---- PCG with DIC on a folded symmetric system: src/OpenFOAM/matrices/pcg.cu:6-149 ---- */
orc_perf orc_pcg_dic(int nCells, int nFaces, const int* owner, const int* neighbour, const double* diag,
                     const double* upper, const double* b, double* psi, double tol, double relTol,
                     int maxIter, int minIter);

/* ---- agglomeration hierarchies ---- */
typedef struct orc_aggl orc_aggl;
/* OpenFOAM faceAreaPair weights |Sf/sqrt(magSf) * (1,1.01,1.02)|: pair_gamg_agglomeration_cpp.cu:301-311 */
void orc_faceareapair_weights(int nFaces, const double* Sf, const double* magSf, double* w);
/* OpenFOAM faceAreaPair hierarchy: pair_gamg_agglomeration_cpp.cu:28-345. *forward is the static flag. */
orc_aggl* orc_of_agglomeration(int nCells, int nFaces, const int* owner, const int* neighbour,
                               const double* weights, int nCellsInCoarsestLevel, int* forward);
/* brae's device-path hierarchy: device_amg.cu:444-685 (agglomerate, coarseFacesByBucket), :1346-1431
 * (buildAMG, injection branch, MERGE 1). */
orc_aggl* orc_brae_agglomeration(int nCells, int nFaces, const int* owner, const int* neighbour,
                                 const double* weights, int target);
int orc_aggl_nlevels(const orc_aggl* a);            /* number of coarse levels */
/* level i describes the step from grid i to grid i+1 (grid 0 = the mesh) */
void orc_aggl_sizes(const orc_aggl* a, int i, int* nFineCells, int* nFineFaces, int* nCoarseCells, int* nCoarseFaces);
void orc_aggl_level(const orc_aggl* a, int i, int* restrictMap, int* coarseOwner, int* coarseNeighbour,
                    int* faceRestrict, signed char* faceFlip);
void orc_aggl_free(orc_aggl* a);

/* ---- Galerkin coarse matrices (symmetric): gamg_solver_cpp.cu:474-526 (OpenFOAM order);
 *      device_amg.cu:156-223 (brae gather order). Levels are written back to back. ---- */
void orc_galerkin_of(const orc_aggl* a, const double* fineDiag, const double* fineUpper,
                     double* coarseDiag, double* coarseUpper);
void orc_galerkin_brae(const orc_aggl* a, const double* fineDiag, const double* fineUpper,
                       double* coarseDiag, double* coarseUpper);

/* ---- OpenFOAM GAMG: gamg_solver_cpp.cu:627-735 (vCycle), :979-1023 (solve), :1027-1143 (PCG+GAMG) ---- */
typedef struct {
    int smoother;             /* 0 GaussSeidel, 1 symGaussSeidel, 2 DIC, 3 DICGaussSeidel */
    double tolerance, relTol; /* the GAMG's own (coarsest PCG uses them) */
    int maxIter, minIter;
    int nPreSweeps, preSweepsLevelMultiplier, maxPreSweeps;
    int nPostSweeps, postSweepsLevelMultiplier, maxPostSweeps;
    int nFinestSweeps;
    int scaleCorrection;
    int nVcycles;             /* GAMG as a preconditioner */
} orc_gamg_controls;
void orc_gamg_default_controls(orc_gamg_controls* c);
orc_perf orc_gamg_solve(const orc_aggl* a, const double* diag, const double* upper, const double* b,
                        double* psi, const orc_gamg_controls* c);
orc_perf orc_pcg_gamg_solve(const orc_aggl* a, const double* diag, const double* upper, const double* b,
                            double* psi, double tol, double relTol, int maxIter, int minIter,
                            const orc_gamg_controls* c);

/* ---- brae AMG-PCG (device path semantics, FP64 arithmetic here):
 *      V(nPre,nPost) weighted Jacobi (omega 0.8) from zero, injection, coarsest by the explicit
 *      inverse (Gauss-Jordan, partial pivoting; the reference's default coarse kind):
 *      device_amg_vcycle.cu:982-1143 / 1658-1766, amg_kernels.cuh:27-53,257-282;
 *      PCG recurrence device_amg_pcg.cu:60-170 (Fletcher-Reeves) or flexBetaK :37-49 (flexible).
 *      scale: each grid's coarse correction scaled by alpha = (r.e)/(e.Ae) (corrScaling,
 *      device_amg_vcycle.cu:1113-1126); 1: fused with the first post-sweep as the HIP library
 *      computes it (x + alpha*e + omega*(r - alpha*Ae)/D); 2 (FP64 cycle only): the reference's own
 *      order, x += alpha*e and then full post-sweeps.
 *      normFactorIn > 0 replaces OpenFOAM's normFactor (the demo trace passes sum|b| + 1e-20).
 *      psiTrace (nullable, (maxIter+1)*nCells) receives psi after each iteration. ---- */
orc_perf orc_brae_amgpcg(const orc_aggl* a, const double* diag, const double* upper, const double* b,
                         double* psi, double tol, double relTol, int maxIter, int minIter,
                         int nPre, int nPost, int flexible, int single, int scale, double normFactorIn,
                         double* psiTrace);

/* OpenFOAM's GaussSeidel sweeps (smooth_solver_cpp.cu:87-155) and DIC preconditioner (calcReciprocalD +
 * precondition, gamg_solver_cpp.cu:194-231) on a folded symmetric system, as the GAMG restatement runs them */
void orc_gauss_seidel(int nCells, int nFaces, const int* owner, const int* neighbour, const double* diag,
                      const double* upper, const double* b, double* psi, int nSweeps);
/* symGaussSeidel: each sweep is the forward sweep, then a reverse walk (cells descending) from the
 * forward walk's accumulated lower-neighbour terms (smooth_solver_cpp.cu:141-163) */
void orc_sym_gauss_seidel(int nCells, int nFaces, const int* owner, const int* neighbour, const double* diag,
                          const double* upper, const double* b, double* psi, int nSweeps);
void orc_dic_precondition(int nCells, int nFaces, const int* owner, const int* neighbour, const double* diag,
                          const double* upper, const double* r, double* w);

/* one brae V-cycle z = M^-1 r on the hierarchy (same cycle as orc_brae_amgpcg applies) */
void orc_brae_vcycle(const orc_aggl* a, const double* diag, const double* upper, const double* r, double* z,
                     int nPre, int nPost, int single, int scale);

/* ---- The SIMPLE pressure predictor's momentum glue (src/applications/solvers/simpleFoam/pEqn_cpp.cu:40-133):
 * rAU = 1/A, H() per component and HbyA = rAU*H, constrainHbyA, phiHbyA = fvc::flux(HbyA), adjustPhi.
 * Boundary arrays per boundary face in face (patch) order; Sf xyz-interleaved. ---- */
/* A()'s numerator (fv_matrix_ops.cuh:211-226, matrixA): D = diag + cmptAv(internalCoeffs) face by face,
 * cmptAv = (x + y + z)/3 */
void orc_fold_boundary_diag(int nCells, int nBFaces, const int* bFaceCell, const double* icX, const double* icY,
                            const double* icZ, const double* diag, double* D);
/* rAU = 1/A, A = D/V (fv_matrix_ops.cuh:224-225; pEqn_cpp.cu:44-47) */
void orc_reciprocal_v(int nCells, const double* D, const double* V, double* rAU);
/* H()'s boundary-diagonal factor of component k: cmptAv(ic) - ic_k (fv_matrix_ops.cuh:290-300) */
void orc_boundary_diag_term(int nBFaces, const double* icX, const double* icY, const double* icZ, int k,
                            double* out);
/* H() of one component (fv_matrix_ops.cuh:277-340, matrixH; uncoupled patches): from zero, += bdDiag*psi per
 * boundary face; per internal face in order H[nei] -= lower*psi[own], H[own] -= upper*psi[nei]; += source;
 * += boundaryCoeffs per boundary face; /= V */
void orc_relax_matrix(int nCells, int nFaces, const int* owner, const int* neighbour, const double* upper,
                      const double* lower, int nBFaces, const int* bFaceCell, const double* icMag,
                      const double* icMin, const double* cycSumOff, double alpha, double* diag, int nComp,
                      const double* const* psi, double* const* source, double* delta);
void orc_linear_upwind_corr(int nCells, int nFaces, const int* owner, const int* neighbour, const double* phi,
                            const double* Cf, const double* C, const double* grad, double* corr);
void orc_scaled_dev2t(int n, const double* gradU, const double* nu, double* sigma);
void orc_tensor_div(int nCells, int nFaces, const int* owner, const int* neighbour, const double* w,
                    const double* Sf, int nBFaces, const int* bFaceCell, const int* isEmpty, const double* Tcell,
                    const double* Tbnd, double* dX, double* dY, double* dZ);
void orc_cmpt_mag_max_min(int n, const double* x, const double* y, const double* z, double* mag, double* mn);
void orc_relax_field(int n, double* p, const double* pPrev, double alpha);
void orc_matrix_h(int nCells, int nFaces, const int* owner, const int* neighbour, const double* upper,
                  const double* lower, const double* psi, const double* source, int nBFaces, const int* bFaceCell,
                  const double* bdDiag, const double* bdSrc, const double* V, double* H);
/* fvc::flux of a cell vector field (fvc.cu:545-580): internal (w*U_own + (1-w)*U_nei) . Sf */
void orc_vector_flux(int nFaces, const int* owner, const int* neighbour, const double* weights, const double* Sf,
                     const double* Ux, const double* Uy, const double* Uz, double* phi);
/* ...and its boundary faces: Ub . Sf_b, zero on an empty face (SfB: the boundary faces' Sf) */
void orc_boundary_flux(int nBFaces, const double* SfB, const double* uxb, const double* uyb, const double* uzb,
                       const int* isEmpty, double* phiB);
/* adjustPhi (pEqn_cpp.cu:80-133): totalFlux = 1e-300 + sum|phi| over internal faces; massIn, fixed and adjustable
 * outflow over the boundary; massCorr = (massIn - fixedOut)/adjOut when adjOut is not negligible, and the
 * adjustable outflow scaled by it. Returns 0, or 1 when the continuity error cannot be removed (phiB
 * untouched). closedVolume (nullable): all three negligible against totalFlux (device_simple.cu:282-346). */
int orc_adjust_phi(int nFaces, const double* phiInt, int nBFaces, const int* adjustable, double* phiB,
                   double* massCorr, int* closedVolume);

/* ---- one SIMPLE iteration (oracle/simple.c): simpleFoam_cpp.cu:15-237 ---- */
/* PBiCGStab with DILU on an asymmetric LDU matrix plus uncoupled boundary coefficients: pbicgstab.cu:6-150 */
orc_perf orc_pbicgstab(int nCells, int nFaces, const int* owner, const int* neighbour, const double* diag,
                       const double* upper, const double* lower, const double* source, int nBFaces,
                       const int* bFaceCell, const double* iC, const double* bC, double* psi, double tol,
                       double relTol, int maxIter, int minIter);
/* the GAMG of gamg.cu:267-400 (magSf face weights, agglomeratePair, coarsen, GaussSeidel, dense coarsest);
 * nLevels (nullable) receives the number of levels */
orc_perf orc_gamg_ref(int nCells, int nFaces, const int* owner, const int* neighbour, const double* diag,
                      const double* upper, const double* lower, const double* source, int nBFaces,
                      const int* bFaceCell, const double* iC, const double* bC, const double* magSf,
                      double* psi, double tol, double relTol, int maxIter, int minIter, int nCellsInCoarsest,
                      int nPostSweeps, int nFinestSweeps, int* nLevels);
/* one agglomeratePair pass of gamg.cu:41-107; returns the number of coarse cells */
int orc_gamg_ref_pair(int nCells, int nFaces, const int* lowerAddr, const int* upperAddr, const double* weights,
                      int forward, int* map);
/* the GAMG's parts: one coarsening (returns the coarse face count; arrays may be NULL to size them),
 * GaussSeidel sweeps on an asymmetric level, the dense coarsest solve and the correction scaling */
int orc_gamg_ref_coarsen(int nCells, int nFaces, const int* lowerAddr, const int* upperAddr, const double* diag,
                         const double* upper, const double* lower, const double* weights, const int* map,
                         int nCoarse, int* cLower, int* cUpper, double* cDiag, double* cUpperV, double* cLowerV,
                         double* cWeights);
void orc_gamg_ref_gauss_seidel(int nCells, int nFaces, const int* lowerAddr, const int* upperAddr,
                               const double* diag, const double* upper, const double* lower, const double* source,
                               double* psi, int nSweeps);
void orc_gamg_ref_dense(int nCells, int nFaces, const int* lowerAddr, const int* upperAddr, const double* diag,
                        const double* upper, const double* lower, const double* b, double* x);
void orc_gamg_ref_scale(int nCells, int nFaces, const int* lowerAddr, const int* upperAddr, const double* diag,
                        const double* upper, const double* lower, const double* source, double* field);
/* fvc::gradUBoundary for fixedValue/zeroGradient/empty patches (fvc.cu:781-865); gradU [9*nCells] and
 * gb [9*nBFaces] component-major, (i*3 + j): dU_j/dx_i */
void orc_grad_u_boundary(int nCells, int nFaces, int nBFaces, const int* bFaceCell, const int* uType,
                         const double* Sf, const double* magSf, const double* bDeltaCoeffs, const double* Ux,
                         const double* Uy, const double* Uz, const double* Ubx, const double* Uby,
                         const double* Ubz, const double* gradU, double* gb);

enum { ORC_P_GAMG_REF = 0, ORC_P_BRAE_AMGPCG = 1, ORC_P_PCG_GAMG = 2 };
typedef struct {
    int nCells, nFaces, nBFaces;
    const int *owner, *neighbour, *bFaceCell;        /* internal owner/neighbour, boundary faces' cells */
    /* geometry as orc_geometry writes it: Sf, Cf, magSf over internal then boundary faces */
    const double *Sf, *Cf, *magSf, *C, *V, *weights, *deltaCoeffs, *nonOrthDeltaCoeffs, *corrVec, *bDeltaCoeffs;
    const int* uType;          /* per boundary face: ORC_BC_FIXED_VALUE (noSlip: value 0), ZERO_GRADIENT, EMPTY */
    const double* uValue;      /* 3*nBFaces, xyz: the fixed values */
    const int* pType;
    const double* pValue;      /* nBFaces */
    double nu;                 /* laminar: nuEff = nu in every cell and on every boundary face */
    const double *nuEff, *nuEffB; /* nullable: a frozen nuEff per cell / boundary face instead */
    double relaxU, relaxP;
    int momentumPredictor, linearUpwind, correctedLaplacian, nNonOrthCorr;
    double tolU, relTolU; int maxIterU, minIterU;
    double tolP, relTolP; int maxIterP, minIterP;
    int pSolver;               /* ORC_P_*: gamg.cu's GAMG (the reference CPU path), brae AMG-PCG, PCG + GAMG */
    const orc_aggl* pAggl;     /* the hierarchy of ORC_P_BRAE_AMGPCG / ORC_P_PCG_GAMG */
    int amgPre, amgPost, amgFlexible, amgSingle, amgScale;
    int pRefCell; double pRefValue;
} orc_simple_case;
typedef struct {
    double *Ux, *Uy, *Uz;      /* cells */
    double *Ubx, *Uby, *Ubz;   /* boundary faces */
    double *p, *pb, *phi, *phiB;
} orc_simple_fields;
typedef struct {
    double UInitial[3], UFinal[3]; int UIters[3];
    double pInitial, pFinal; int pIters[8];
    double massCorr; int adjustStatus;
} orc_simple_result;
/* one iteration in place on the fields; returns 0, or 1 when adjustPhi refuses */
int orc_simple_step(const orc_simple_case* cs, orc_simple_fields* fields, orc_simple_result* res);

/* ---- the standard k-epsilon model as the host SIMPLE driver calls it (kepsilon.c) ---- */
typedef struct {
    double Cmu, C1, C2, C3, sigmaK, sigmaEps, kappa, E;
} orc_kepsilon_coeffs;
void orc_kepsilon_default_coeffs(orc_kepsilon_coeffs* co);
typedef struct {
    int nCells, nFaces, nBFaces;
    const int *owner, *neighbour, *bFaceCell;
    /* geometry as orc_geometry writes it (Sf, magSf over internal then boundary faces) */
    const double *Sf, *magSf, *V, *weights, *deltaCoeffs, *bDeltaCoeffs;
    const int* isWall;          /* per boundary face: the face of a `wall` patch (the wall functions) */
    const double* yWall;        /* per boundary face: the near-wall distance (wall faces) */
    const int *kType, *epsType; /* ORC_BC_* per boundary face (the wall functions' patches: zeroGradient) */
    const double *kValue, *epsValue;
    double nu, relaxEps, relaxK, tol, relTol;
    int maxIter, minIter;
    orc_kepsilon_coeffs co;
} orc_kepsilon_case;
typedef struct {
    double epsInitial, epsFinal, kInitial, kFinal;
    int epsIters, kIters, nWallCells;
} orc_kepsilon_result;
/* nearWallDist (near_wall_dist.cuh:120-177): y per boundary face of a wall patch; faceOffsets/faceVerts over
 * every face (internal then boundary), C the cell centres [3*nCells] */
void orc_near_wall_dist(int nFaces, int nBFaces, const double* points, const int* faceOffsets,
                        const int* faceVerts, const int* bFaceCell, const double* C, const int* isWall, double* y);
double orc_y_plus_lam(double kappa, double E);
double orc_nutk_wall(double Cmu, double kappa, double E, double y, double kCell, double nu);
void orc_gbynu(int n, const double* gradU, double* out);
/* kEqn and epsilonEqn as assembled matrices (k_epsilon.cuh:137-186): nutb NULL takes the owner cell's
 * diffusivity on the boundary */
void orc_kepsilon_eqns(const orc_kepsilon_case* cs, const double* k, const double* kb, const double* eps,
                       const double* epsb, const double* nut, const double* nutb, const double* G,
                       const double* GbyNu, const double* phi, const double* phiB, double* kD, double* kU,
                       double* kL, double* kS, double* kIC, double* kBC, double* eD, double* eU, double* eL,
                       double* eS, double* eIC, double* eBC);
/* one kepsilon::correct (k_epsilon.cuh:221-337) in place on k, epsilon and nut (cells and boundary faces) */
int orc_kepsilon_correct(const orc_kepsilon_case* cs, const double* const* U, const double* const* Ub,
                         const double* phi, const double* phiB, double* k, double* kb, double* eps, double* epsb,
                         double* nut, double* nutb, orc_kepsilon_result* res);

#ifdef __cplusplus
}
#endif
#endif
