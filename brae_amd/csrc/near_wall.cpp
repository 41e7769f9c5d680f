// The near-wall distance the wall functions read (OpenFOAM's turbulence.y(); the reference's nearWallDist,
// src/finiteVolume/fvMesh/near_wall_dist.cuh:31-177), on the host, once per mesh: for each face of a wall
// patch, the smallest distance from its cell's centre to that face or to any wall face sharing a vertex
// with it, on any wall patch. The distance to a face is the smallest over the triangles of its fan about the
// area-weighted centre, each by the closest point on a triangle (Ericson, Real-Time Collision Detection,
// 5.1.5).
#include "simple.hpp"

#include <array>
#include <cmath>

namespace mfv {

namespace {

using V3 = std::array<double, 3>;

inline V3 sub(const V3& a, const V3& b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
inline double dot(const V3& a, const V3& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline V3 axpy(const V3& a, double s, const V3& d) { return {a[0] + s * d[0], a[1] + s * d[1], a[2] + s * d[2]}; }

V3 closestOnTriangle(const V3& p, const V3& a, const V3& b, const V3& c)
{
    const V3 ab = sub(b, a), ac = sub(c, a), ap = sub(p, a);
    const double d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0.0 && d2 <= 0.0) return a;
    const V3 bp = sub(p, b);
    const double d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0.0 && d4 <= d3) return b;
    const double vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) return axpy(a, d1 / (d1 - d3), ab);
    const V3 cp = sub(p, c);
    const double d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0.0 && d5 <= d6) return c;
    const double vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) return axpy(a, d2 / (d2 - d6), ac);
    const double va = d3 * d6 - d5 * d4;
    if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0)
        return axpy(b, (d4 - d3) / ((d4 - d3) + (d5 - d6)), sub(c, b));
    const double denom = 1.0 / (va + vb + vc);
    const double v = vb * denom, w = vc * denom;
    return {a[0] + v * ab[0] + w * ac[0], a[1] + v * ab[1] + w * ac[1], a[2] + v * ab[2] + w * ac[2]};
}

V3 point(const double* pts, int i) { return {pts[3 * (size_t)i], pts[3 * (size_t)i + 1], pts[3 * (size_t)i + 2]}; }

double pointToFace(const V3& p, const double* pts, const int* fv, int v0, int v1)
{
    const int n = v1 - v0;
    V3 cp{0.0, 0.0, 0.0};
    for (int j = v0; j < v1; ++j)
    {
        const V3 q = point(pts, fv[j]);
        for (int i = 0; i < 3; ++i) cp[i] += q[i];
    }
    for (int i = 0; i < 3; ++i) cp[i] = (1.0 / n) * cp[i];
    double sumA = 0.0;
    V3 sumAc{0.0, 0.0, 0.0};
    for (int j = 0; j < n; ++j)
    {
        const V3 a = point(pts, fv[v0 + j]), b = point(pts, fv[v0 + (j + 1) % n]);
        const V3 u = sub(a, cp), v = sub(b, cp);
        const V3 cr{u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
        const double ta = std::sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
        sumA += ta;
        for (int i = 0; i < 3; ++i) sumAc[i] += ta * (a[i] + b[i] + cp[i]);
    }
    V3 ctr = cp;
    if (sumA > 1.0e-30)
        for (int i = 0; i < 3; ++i) ctr[i] = sumAc[i] / (3.0 * sumA);
    double best = 1.0e15;
    for (int j = 0; j < n; ++j)
    {
        const V3 a = point(pts, fv[v0 + j]), b = point(pts, fv[v0 + (j + 1) % n]);
        const V3 d = sub(p, closestOnTriangle(p, ctr, a, b));
        const double dist = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        if (dist < best) best = dist;
    }
    return best;
}

} // namespace

void nearWallDistHost(int nFaces, int nBFaces, const double* points, const int* faceOffsets, const int* faceVerts,
                      const int* bFaceCell, const double* C, const int* isWall, double* y)
{
    std::vector<int> wf;
    for (int i = 0; i < nBFaces; ++i)
        if (isWall[i]) wf.push_back(i);
    const int nW = (int)wf.size();
    int nP = 0;
    for (int w = 0; w < nW; ++w)
    {
        const int f = nFaces + wf[w];
        for (int j = faceOffsets[f]; j < faceOffsets[f + 1]; ++j) nP = std::max(nP, faceVerts[j] + 1);
    }
    // point -> the wall faces on it (CSR, the wall faces in order)
    std::vector<int> start(nP + 1, 0);
    for (int w = 0; w < nW; ++w)
    {
        const int f = nFaces + wf[w];
        for (int j = faceOffsets[f]; j < faceOffsets[f + 1]; ++j) ++start[faceVerts[j] + 1];
    }
    for (int p = 0; p < nP; ++p) start[p + 1] += start[p];
    std::vector<int> at(start.begin(), start.end() - 1), list(start[nP]);
    for (int w = 0; w < nW; ++w)
    {
        const int f = nFaces + wf[w];
        for (int j = faceOffsets[f]; j < faceOffsets[f + 1]; ++j) list[at[faceVerts[j]]++] = w;
    }
    for (int i = 0; i < nBFaces; ++i) y[i] = 0.0;
    for (int w = 0; w < nW; ++w)
    {
        const int i = wf[w];
        const V3 cc = point(C, bFaceCell[i]);
        const int f = nFaces + i;
        double best = pointToFace(cc, points, faceVerts, faceOffsets[f], faceOffsets[f + 1]);
        for (int j = faceOffsets[f]; j < faceOffsets[f + 1]; ++j)
        {
            const int p = faceVerts[j];
            for (int q = start[p]; q < start[p + 1]; ++q)
            {
                const int o = list[q];
                if (o == w) continue;
                const int g = nFaces + wf[o];
                const double d = pointToFace(cc, points, faceVerts, faceOffsets[g], faceOffsets[g + 1]);
                if (d < best) best = d;
            }
        }
        y[i] = best;
    }
}

} // namespace mfv
