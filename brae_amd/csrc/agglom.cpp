// Pairwise agglomeration on the host. The pairing is a sequential greedy walk whose result depends on
// the visiting order and on strict-'>' tie-breaks, so it stays on one CPU thread; coarse-face
// construction is a counting sort (linear time). Runs once per mesh.
#include "agglom.hpp"

#include <algorithm>
#include <stdexcept>
#include <utility>

namespace mfv {

namespace {

// A cell's faces in CSR form. Brae interleaves owner/neighbour occurrences in face order
// (device_amg.cu:547-558); OpenFOAM lists the faces a cell is the UPPER cell of first, then those it
// owns (pair_gamg_agglomeration_cpp.cu:176-205).
void cellFaces(int n, const std::vector<int>& own, const std::vector<int>& nei, bool upperFirst, std::vector<int>& off,
               std::vector<int>& cf)
{
    const int nF = (int)own.size();
    off.assign(n + 1, 0);
    for (int f = 0; f < nF; ++f)
    {
        ++off[own[f] + 1];
        ++off[nei[f] + 1];
    }
    for (int c = 0; c < n; ++c) off[c + 1] += off[c];
    cf.assign(2 * (size_t)nF, 0);
    std::vector<int> at(off.begin(), off.end() - 1);
    if (upperFirst)
    {
        for (int f = 0; f < nF; ++f) cf[at[nei[f]]++] = f;
        for (int f = 0; f < nF; ++f) cf[at[own[f]]++] = f;
    }
    else
    {
        for (int f = 0; f < nF; ++f)
        {
            cf[at[own[f]]++] = f;
            cf[at[nei[f]]++] = f;
        }
    }
}

// brae agglomerate(): forward greedy pairing over the heaviest unmatched face (strict >, -1e300 seed),
// else join the heaviest neighbour's cluster, else a singleton.
std::vector<int> braePairing(int n, const std::vector<int>& own, const std::vector<int>& nei,
                             const std::vector<double>& fw, int& nCoarse)
{
    std::vector<int> off, cf;
    cellFaces(n, own, nei, false, off, cf);
    std::vector<int> map(n, -1);
    nCoarse = 0;
    for (int c = 0; c < n; ++c)
    {
        if (map[c] >= 0) continue;
        int mf = -1;
        double mw = -1e300;
        for (int o = off[c]; o < off[c + 1]; ++o)
        {
            const int f = cf[o];
            if (map[own[f]] < 0 && map[nei[f]] < 0 && fw[f] > mw)
            {
                mf = f;
                mw = fw[f];
            }
        }
        if (mf >= 0)
        {
            map[own[mf]] = nCoarse;
            map[nei[mf]] = nCoarse;
            ++nCoarse;
            continue;
        }
        int cm = -1;
        double cmw = -1e300;
        for (int o = off[c]; o < off[c + 1]; ++o)
        {
            const int f = cf[o];
            if (fw[f] > cmw)
            {
                cm = f;
                cmw = fw[f];
            }
        }
        if (cm >= 0) map[c] = std::max(map[own[cm]], map[nei[cm]]);
    }
    for (int c = 0; c < n; ++c)
        if (map[c] < 0) map[c] = nCoarse++;
    return map;
}

// coarseFacesByBucket: crossing faces bucketed by lower coarse cell, each bucket sorted by (higher
// cell, fine face); a run of equal higher cells is one coarse face. Weights summed in ascending fine face.
void braeCoarseFaces(AggLevel& L, const std::vector<int>& own, const std::vector<int>& nei,
                     const std::vector<double>& fw, std::vector<double>& cfw)
{
    const int nF = (int)own.size(), nC = L.nCoarse;
    const std::vector<int>& map = L.map;
    std::vector<int> start(nC + 1, 0);
    for (int f = 0; f < nF; ++f)
    {
        const int co = map[own[f]], cn = map[nei[f]];
        if (co != cn) ++start[std::min(co, cn) + 1];
    }
    for (int c = 0; c < nC; ++c) start[c + 1] += start[c];
    std::vector<std::pair<int, int>> entry(start[nC]);
    {
        std::vector<int> at(start.begin(), start.end() - 1);
        for (int f = 0; f < nF; ++f)
        {
            const int co = map[own[f]], cn = map[nei[f]];
            if (co != cn) entry[at[std::min(co, cn)]++] = {std::max(co, cn), f};
        }
    }
    L.faceRestrict.assign(nF, 0);
    L.faceFlip.assign(nF, 0);
    L.cOwn.clear();
    L.cNei.clear();
    for (int c = 0; c < nC; ++c)
    {
        auto b = entry.begin() + start[c], e = entry.begin() + start[c + 1];
        std::sort(b, e);
        int last = -1;
        for (auto it = b; it != e; ++it)
        {
            if (it->first != last)
            {
                last = it->first;
                L.cOwn.push_back(c);
                L.cNei.push_back(last);
            }
            L.faceRestrict[it->second] = (int)L.cOwn.size() - 1;
        }
    }
    L.nCoarseFaces = (int)L.cOwn.size();
    cfw.assign(L.nCoarseFaces, 0.0);
    for (int f = 0; f < nF; ++f)
    {
        const int co = map[own[f]], cn = map[nei[f]];
        if (co == cn)
        {
            L.faceRestrict[f] = -1 - co;
            L.faceFlip[f] = 0;
        }
        else
        {
            L.faceFlip[f] = co > cn ? 1 : 0;
            cfw[L.faceRestrict[f]] += fw[f];
        }
    }
}

// gamg.cu agglomeratePair (gamg.cu:41-107), one pass: cells visited forward or in reverse; a cell's faces
// are its upper-side ones, then its lower-side ones, each in face order; pair with the heaviest face whose
// cells are both unmatched (strict >, -1e300 seed), else join the cluster across the heaviest face (the
// larger of its two cells' numbers; nothing when both are unmatched), singletons numbered after the walk in
// its direction; a reverse pass renumbers c -> nCoarse - 1 - c.
std::vector<int> pairRefPairing(int n, const std::vector<int>& lower, const std::vector<int>& upper,
                                const std::vector<double>& fw, bool fwd, int& nCoarse)
{
    std::vector<int> off, cf;
    cellFaces(n, lower, upper, true, off, cf);
    std::vector<int> map(n, -1);
    nCoarse = 0;
    for (int k = 0; k < n; ++k)
    {
        const int c = fwd ? k : n - k - 1;
        if (map[c] >= 0) continue;
        int mf = -1;
        double mw = -1e300;
        for (int o = off[c]; o < off[c + 1]; ++o)
        {
            const int f = cf[o];
            if (map[upper[f]] < 0 && map[lower[f]] < 0 && fw[f] > mw)
            {
                mf = f;
                mw = fw[f];
            }
        }
        if (mf >= 0)
        {
            map[upper[mf]] = nCoarse;
            map[lower[mf]] = nCoarse;
            ++nCoarse;
            continue;
        }
        int cm = -1;
        double cmw = -1e300;
        for (int o = off[c]; o < off[c + 1]; ++o)
        {
            const int f = cf[o];
            if (fw[f] > cmw)
            {
                cm = f;
                cmw = fw[f];
            }
        }
        if (cm >= 0) map[c] = std::max(map[upper[cm]], map[lower[cm]]);
    }
    for (int k = 0; k < n; ++k)
    {
        const int c = fwd ? k : n - k - 1;
        if (map[c] < 0) map[c] = nCoarse++;
    }
    if (!fwd)
        for (int& c : map) c = nCoarse - 1 - c;
    return map;
}

// pairGAMGAgglomeration::agglomerate, one pass (pair_gamg_agglomeration_cpp.cu:158-284)
std::vector<int> ofPairing(int n, const std::vector<int>& lower, const std::vector<int>& upper,
                           const std::vector<double>& fw, bool& forward, int& nCoarse)
{
    std::vector<int> off, cf;
    cellFaces(n, lower, upper, true, off, cf);
    std::vector<int> map(n, -1);
    nCoarse = 0;
    const double great = 1.0e+15;
    const bool fwd = forward;
    for (int cellfi = 0; cellfi < n; ++cellfi)
    {
        const int celli = fwd ? cellfi : n - cellfi - 1;
        if (map[celli] >= 0) continue;
        int matchFaceNo = -1;
        double maxFaceWeight = -great;
        for (int o = off[celli]; o < off[celli + 1]; ++o)
        {
            const int f = cf[o];
            if (map[upper[f]] < 0 && map[lower[f]] < 0 && fw[f] > maxFaceWeight)
            {
                matchFaceNo = f;
                maxFaceWeight = fw[f];
            }
        }
        if (matchFaceNo >= 0)
        {
            map[upper[matchFaceNo]] = nCoarse;
            map[lower[matchFaceNo]] = nCoarse;
            ++nCoarse;
            continue;
        }
        int clusterMatchFaceNo = -1;
        double clusterMaxFaceCoeff = -great;
        for (int o = off[celli]; o < off[celli + 1]; ++o)
        {
            const int f = cf[o];
            if (fw[f] > clusterMaxFaceCoeff)
            {
                clusterMatchFaceNo = f;
                clusterMaxFaceCoeff = fw[f];
            }
        }
        if (clusterMatchFaceNo >= 0)
            map[celli] = std::max(map[upper[clusterMatchFaceNo]], map[lower[clusterMatchFaceNo]]);
    }
    for (int cellfi = 0; cellfi < n; ++cellfi)
    {
        const int celli = fwd ? cellfi : n - cellfi - 1;
        if (map[celli] < 0) map[celli] = nCoarse++;
    }
    if (!fwd)
    {
        --nCoarse;
        for (int& c : map) c = nCoarse - c;
        ++nCoarse;
    }
    forward = !fwd;
    return map;
}

// agglomerateLduAddressing (pair_gamg_agglomeration_cpp.cu:28-154): coarse faces owner-major in the
// order the fine-face walk first meets them; flip where the fine UPPER cell lands in the coarse owner.
void ofCoarseFaces(AggLevel& L, const std::vector<int>& lower, const std::vector<int>& upper)
{
    const int nF = (int)lower.size(), nC = L.nCoarse;
    const std::vector<int>& rm = L.map;
    std::vector<int> head(nC, -1), tail(nC, -1), next(nF, -1), initNei(nF, 0);
    L.faceRestrict.assign(nF, 0);
    int nCoarseFaces = 0;
    for (int f = 0; f < nF; ++f)
    {
        const int rU = rm[upper[f]], rL = rm[lower[f]];
        if (rU == rL)
        {
            L.faceRestrict[f] = -(rU + 1);
            continue;
        }
        const int cOwn = std::min(rU, rL), cNei = std::max(rU, rL);
        int found = -1;
        for (int k = head[cOwn]; k >= 0; k = next[k])
            if (initNei[k] == cNei)
            {
                found = k;
                break;
            }
        if (found >= 0)
        {
            L.faceRestrict[f] = found;
            continue;
        }
        initNei[nCoarseFaces] = cNei;
        if (tail[cOwn] >= 0)
            next[tail[cOwn]] = nCoarseFaces;
        else
            head[cOwn] = nCoarseFaces;
        tail[cOwn] = nCoarseFaces;
        L.faceRestrict[f] = nCoarseFaces++;
    }
    L.nCoarseFaces = nCoarseFaces;
    L.cOwn.assign(nCoarseFaces, 0);
    L.cNei.assign(nCoarseFaces, 0);
    std::vector<int> renum(nCoarseFaces);
    int k = 0;
    for (int c = 0; c < nC; ++c)
        for (int cf = head[c]; cf >= 0; cf = next[cf])
        {
            L.cOwn[k] = c;
            L.cNei[k] = initNei[cf];
            renum[cf] = k++;
        }
    L.faceFlip.assign(nF, 0);
    for (int f = 0; f < nF; ++f)
    {
        if (L.faceRestrict[f] < 0) continue;
        L.faceRestrict[f] = renum[L.faceRestrict[f]];
        if (L.cOwn[L.faceRestrict[f]] == rm[upper[f]]) L.faceFlip[f] = 1;
    }
}

} // namespace

std::vector<AggLevel> agglomerate(AggKind kind, int nCells, const std::vector<int>& fineOwner,
                                  const std::vector<int>& fineNeighbour, const std::vector<double>& weights,
                                  int target, bool& forward)
{
    if (weights.size() != fineOwner.size() || fineNeighbour.size() != fineOwner.size())
        throw std::invalid_argument("agglomerate: face weights must be given per internal face");
    std::vector<AggLevel> levels;
    const std::vector<int>* own = &fineOwner;
    const std::vector<int>* nei = &fineNeighbour;
    std::vector<double> fw = weights;
    int n = nCells;
    if (kind == AggKind::Brae)
    {
        while (n > target)
        {
            AggLevel L;
            L.nFine = n;
            L.nFineFaces = (int)own->size();
            L.map = braePairing(n, *own, *nei, fw, L.nCoarse);
            std::vector<double> cfw;
            braeCoarseFaces(L, *own, *nei, fw, cfw);
            if (L.nCoarse >= n || L.nCoarseFaces == 0) break;
            levels.push_back(std::move(L));
            fw = std::move(cfw);
            own = &levels.back().cOwn;
            nei = &levels.back().cNei;
            n = levels.back().nCoarse;
        }
        return levels;
    }
    if (kind == AggKind::Pair)
    {
        // gamg.cu:290-312: at most 50 grids; a pass that reaches target cells or fewer, does not coarsen, or
        // leaves fewer than 2 cells ends the hierarchy
        bool fwd = true;
        while ((int)levels.size() < 49)
        {
            AggLevel L;
            L.nFine = n;
            L.nFineFaces = (int)own->size();
            L.map = pairRefPairing(n, *own, *nei, fw, fwd, L.nCoarse);
            fwd = !fwd;
            if (L.nCoarse <= target || L.nCoarse >= n || L.nCoarse < 2) break;
            std::vector<double> cfw;
            braeCoarseFaces(L, *own, *nei, fw, cfw);
            levels.push_back(std::move(L));
            fw = std::move(cfw);
            own = &levels.back().cOwn;
            nei = &levels.back().cNei;
            n = levels.back().nCoarse;
        }
        return levels;
    }
    // OpenFOAM faceAreaPair: nCellsInCoarsestLevel clamped to [1, nCells/2] (GAMGAgglomeration.C:285)
    const int nCoarsest = std::max(1, std::min(nCells / 2, target));
    const int maxLevels = 50;
    while ((int)levels.size() < maxLevels - 1)
    {
        AggLevel L;
        L.nFine = n;
        L.nFineFaces = (int)own->size();
        L.map = ofPairing(n, *own, *nei, fw, forward, L.nCoarse);
        if (!(L.nCoarse >= nCoarsest && L.nCoarse < n)) break;
        ofCoarseFaces(L, *own, *nei);
        std::vector<double> cfw(L.nCoarseFaces, 0.0);
        for (int f = 0; f < L.nFineFaces; ++f)
            if (L.faceRestrict[f] >= 0) cfw[L.faceRestrict[f]] += fw[f];
        levels.push_back(std::move(L));
        fw = std::move(cfw);
        own = &levels.back().cOwn;
        nei = &levels.back().cNei;
        n = levels.back().nCoarse;
    }
    return levels;
}

// OpenFOAM's mergeLevels (pairGAMGAgglomeration: combineLevels while nPairLevels % mergeLevels != 0):
// each run of `merge` consecutive passes becomes one level. The composed level maps a fine cell through
// both passes; a fine face lands on the second pass's coarse face, or inside a coarse cell when either
// pass swallowed it; its flip is the parity of the two flips.
static AggLevel compose(const AggLevel& A, const AggLevel& B)
{
    AggLevel C;
    C.nFine = A.nFine;
    C.nFineFaces = A.nFineFaces;
    C.nCoarse = B.nCoarse;
    C.nCoarseFaces = B.nCoarseFaces;
    C.map.resize(A.nFine);
    for (int i = 0; i < A.nFine; ++i) C.map[i] = B.map[A.map[i]];
    C.cOwn = B.cOwn;
    C.cNei = B.cNei;
    C.faceRestrict.resize(A.nFineFaces);
    C.faceFlip.assign(A.nFineFaces, 0);
    for (int f = 0; f < A.nFineFaces; ++f)
    {
        const int a = A.faceRestrict[f];
        if (a >= 0)
        {
            const int b = B.faceRestrict[a];
            C.faceRestrict[f] = b;
            if (b >= 0) C.faceFlip[f] = (int8_t)((A.faceFlip[f] != 0) != (B.faceFlip[a] != 0));
        }
        else
        {
            C.faceRestrict[f] = -(B.map[-1 - a] + 1);
        }
    }
    return C;
}

std::vector<AggLevel> mergeLevels(std::vector<AggLevel> passes, int merge)
{
    if (merge <= 1) return passes;
    std::vector<AggLevel> out;
    for (size_t i = 0; i < passes.size(); i += (size_t)merge)
    {
        AggLevel L = std::move(passes[i]);
        for (size_t k = i + 1; k < passes.size() && k < i + (size_t)merge; ++k) L = compose(L, passes[k]);
        out.push_back(std::move(L));
    }
    return out;
}

} // namespace mfv
