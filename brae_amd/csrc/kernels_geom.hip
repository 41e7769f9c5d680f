// Finite-volume mesh geometry on the device, in OpenFOAM's algorithm and operation order
// (primitiveMeshTools via the reference's fv_geometry.cu:13-158 and fv_patch.cu:160-165):
// face centres/areas from the triangle fan about the point average, cell centres/volumes from the
// pyramid decomposition about the face-centre average, then weights, deltaCoeffs,
// nonOrthDeltaCoeffs and the non-orthogonal correction vectors; boundary deltaCoeffs.
//
// The reference accumulates a cell's sums in two passes (all faces it owns in ascending face index,
// internal faces then boundary faces; then the internal faces it neighbours, ascending). The gathers
// below walk the same lists in the same order, so the results equal the host loops to the bit.
#include "mfv_internal.hpp"

namespace mfv {

namespace {

constexpr int kTpb = 256;
inline int blocks(long long n) { return (int)((n + kTpb - 1) / kTpb); }

struct V3
{
    double x, y, z;
};
__device__ __forceinline__ V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 scl(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ V3 dvd(V3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ double mag(V3 a) { return sqrt(dot(a, a)); }
__device__ __forceinline__ V3 cross(V3 a, V3 b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ V3 ld(const double* p, long i) { return {p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }
__device__ __forceinline__ void st(double* p, long i, V3 v)
{
    p[3 * i] = v.x;
    p[3 * i + 1] = v.y;
    p[3 * i + 2] = v.z;
}

// face centres and area vectors (fv_geometry.cu:13-77)
__global__ __launch_bounds__(kTpb) void faceGeomK(int nF, const double* __restrict__ P, const int* __restrict__ off,
                                                  const int* __restrict__ fv, double* __restrict__ Cf,
                                                  double* __restrict__ Sf, double* __restrict__ magSf)
{
    const int f = blockIdx.x * kTpb + threadIdx.x;
    if (f >= nF) return;
    const int k0 = off[f], k = off[f + 1] - k0;
    V3 cf, sf;
    if (k == 3)
    {
        const V3 a = ld(P, fv[k0]), b = ld(P, fv[k0 + 1]), c = ld(P, fv[k0 + 2]);
        cf = scl(1.0 / 3.0, add(add(a, b), c));
        sf = scl(0.5, cross(sub(b, a), sub(c, a)));
    }
    else
    {
        V3 fc = ld(P, fv[k0]);
        for (int i = 1; i < k; ++i) fc = add(fc, ld(P, fv[k0 + i]));
        fc = dvd(fc, (double)k);
        V3 sumN{0, 0, 0}, sumAc{0, 0, 0};
        double sumA = 0.0;
        for (int i = 0; i < k; ++i)
        {
            const V3 t = ld(P, fv[k0 + i]);
            const V3 nx = ld(P, fv[k0 + (i == k - 1 ? 0 : i + 1)]);
            const V3 c = add(add(t, nx), fc);
            const V3 n = cross(sub(nx, t), sub(fc, t));
            const double a = mag(n);
            sumN = add(sumN, n);
            sumA += a;
            sumAc = add(sumAc, scl(a, c));
        }
        if (sumA < 1.0e-150)
        {
            cf = fc;
            sf = {0, 0, 0};
        }
        else
        {
            cf = dvd(scl(1.0 / 3.0, sumAc), sumA);
            sf = scl(0.5, sumN);
        }
    }
    st(Cf, f, cf);
    st(Sf, f, sf);
    magSf[f] = mag(sf);
}

// cell centres and volumes (fv_geometry.cu:79-127) by gathering the cell's faces in the host loops'
// order: owned internal, owned boundary, then neighboured internal, each ascending
__global__ __launch_bounds__(kTpb) void cellGeomK(int nC, int nI, const int* __restrict__ ownerStart,
                                                  const int* __restrict__ losortStart,
                                                  const int* __restrict__ losort, const int* __restrict__ bndStart,
                                                  const int* __restrict__ bndFaces, const double* __restrict__ Cf,
                                                  const double* __restrict__ Sf, double* __restrict__ C,
                                                  double* __restrict__ V)
{
    const int c = blockIdx.x * kTpb + threadIdx.x;
    if (c >= nC) return;
    const int o0 = ownerStart[c], o1 = ownerStart[c + 1];
    const int b0 = bndStart[c], b1 = bndStart[c + 1];
    const int l0 = losortStart[c], l1 = losortStart[c + 1];
    V3 ce{0, 0, 0};
    for (int f = o0; f < o1; ++f) ce = add(ce, ld(Cf, f));
    for (int k = b0; k < b1; ++k) ce = add(ce, ld(Cf, (long)nI + bndFaces[k]));
    for (int k = l0; k < l1; ++k) ce = add(ce, ld(Cf, losort[k]));
    ce = dvd(ce, (double)((o1 - o0) + (b1 - b0) + (l1 - l0)));
    V3 cc{0, 0, 0};
    double vol = 0.0;
    auto ownFace = [&](long f) {
        const V3 cf = ld(Cf, f);
        const double pv = dot(ld(Sf, f), sub(cf, ce));
        const V3 pc = add(scl(0.75, cf), scl(0.25, ce));
        cc = add(cc, scl(pv, pc));
        vol += pv;
    };
    for (int f = o0; f < o1; ++f) ownFace(f);
    for (int k = b0; k < b1; ++k) ownFace((long)nI + bndFaces[k]);
    for (int k = l0; k < l1; ++k)
    {
        const int f = losort[k];
        const V3 cf = ld(Cf, f);
        const double pv = dot(ld(Sf, f), sub(ce, cf));
        const V3 pc = add(scl(0.75, cf), scl(0.25, ce));
        cc = add(cc, scl(pv, pc));
        vol += pv;
    }
    st(C, c, fabs(vol) > 1.0e-300 ? dvd(cc, vol) : ce);
    V[c] = vol * (1.0 / 3.0);
}

// weights, deltaCoeffs, nonOrthDeltaCoeffs, correction vectors (fv_geometry.cu:129-158)
__global__ __launch_bounds__(kTpb) void interpGeomK(int nI, const int* __restrict__ own, const int* __restrict__ nei,
                                                    const double* __restrict__ Cf, const double* __restrict__ Sf,
                                                    const double* __restrict__ magSf, const double* __restrict__ C,
                                                    double* __restrict__ w, double* __restrict__ dc,
                                                    double* __restrict__ ndc, double* __restrict__ corr)
{
    const int f = blockIdx.x * kTpb + threadIdx.x;
    if (f >= nI) return;
    const V3 sf = ld(Sf, f), cf = ld(Cf, f), co = ld(C, own[f]), cn = ld(C, nei[f]);
    const double sOwn = fabs(dot(sf, sub(cf, co)));
    const double sNei = fabs(dot(sf, sub(cn, cf)));
    w[f] = fabs(sOwn + sNei) > 1.0e-150 ? sNei / (sOwn + sNei) : 0.5;
    const V3 delta = sub(cn, co);
    dc[f] = 1.0 / mag(delta);
    const V3 unit = dvd(sf, magSf[f]);
    const double nd = 1.0 / fmax(dot(unit, delta), 0.05 * mag(delta));
    ndc[f] = nd;
    st(corr, f, sub(unit, scl(nd, delta)));
}

// boundary deltaCoeffs 1/|nHat.(Cf - C)| (fv_patch.cu:160-165)
__global__ __launch_bounds__(kTpb) void bndDeltaK(int nB, int nI, const int* __restrict__ bFaceCell,
                                                  const double* __restrict__ Cf, const double* __restrict__ Sf,
                                                  const double* __restrict__ magSf, const double* __restrict__ C,
                                                  double* __restrict__ bd)
{
    const int i = blockIdx.x * kTpb + threadIdx.x;
    if (i >= nB) return;
    const long f = (long)nI + i;
    const V3 nh = dvd(ld(Sf, f), magSf[f]);
    bd[i] = 1.0 / fabs(dot(sub(ld(Cf, f), ld(C, bFaceCell[i])), nh));
}

} // namespace

void launchGeometry(Ctx& c, const Mesh& m, const double* points, const int* faceOffsets, const int* faceVerts,
                    double* Cf, double* Sf, double* magSf, double* C, double* V, double* weights, double* dc,
                    double* ndc, double* corr, double* bDelta)
{
    const int nF = m.nFaces + m.nBFaces;
    faceGeomK<<<blocks(nF), kTpb, 0, c.stream>>>(nF, points, faceOffsets, faceVerts, Cf, Sf, magSf);
    cellGeomK<<<blocks(m.nCells), kTpb, 0, c.stream>>>(m.nCells, m.nFaces, m.ownerStart.p, m.losortStart.p,
                                                       m.losort.p, m.bndStart.p, m.bndFaces.p, Cf, Sf, C, V);
    if (m.nFaces)
        interpGeomK<<<blocks(m.nFaces), kTpb, 0, c.stream>>>(m.nFaces, m.owner.p, m.neighbour.p, Cf, Sf, magSf, C,
                                                             weights, dc, ndc, corr);
    if (m.nBFaces)
        bndDeltaK<<<blocks(m.nBFaces), kTpb, 0, c.stream>>>(m.nBFaces, m.nFaces, m.bFaceCell.p, Cf, Sf, magSf, C,
                                                            bDelta);
    hipCheck(hipGetLastError(), "geometry");
}

} // namespace mfv
