// OpenFOAM's on-disk formats through include/brae_foam.hpp, host only (no device calls): reads a case's
// polyMesh and a field, parses fvSolution dictionaries and reads their GAMG entries (with every refusal by
// name), writes fields and reads them back. Prints `key value...` lines that the Python test compares with
// the same files parsed independently; exit status 0 when every check it makes itself holds.
//
// usage: foam_io <case dir: constant/polyMesh, 0/T, system/fvSolution> <an fvSolution with a usable GAMG
//        entry `p`> <scratch dir>
#include "brae_foam.hpp"

#include <cmath>
#include <cstdio>
#include <functional>
#include <string>

using namespace brae;

namespace {

int failures = 0;

void check(bool ok, const std::string& what)
{
    if (!ok)
    {
        ++failures;
        std::printf("FAIL %s\n", what.c_str());
    }
}

// the message of the refusal f throws, or "" when it accepts
std::string refusal(const std::function<void()>& f)
{
    try
    {
        f();
    }
    catch (const std::runtime_error& e)
    {
        return e.what();
    }
    return "";
}

std::string gamgEntry(const std::string& extra)
{
    return "p { solver GAMG; smoother GaussSeidel; tolerance 1e-7; relTol 0.01; " + extra + " }";
}

} // namespace

int main(int argc, char** argv)
{
    if (argc < 4)
    {
        std::printf("usage: foam_io <case> <fvSolution> <scratch>\n");
        return 2;
    }
    const std::string caseDir = argv[1], fvs = argv[2], scratch = argv[3];

    // ---- polyMesh ----
    const PolyMesh m = readPolyMesh(caseDir + "/constant/polyMesh");
    std::printf("mesh %d %d %d %zu\n", m.nCells(), m.nFaces(), m.nInternalFaces(), m.points.size() / 3);
    std::printf("owner");
    for (label o : m.owner) std::printf(" %d", o);
    std::printf("\nneighbour");
    for (label o : m.neighbour) std::printf(" %d", o);
    std::printf("\nfaceOffsets");
    for (label o : m.faceOffsets) std::printf(" %d", o);
    std::printf("\nfaceVerts");
    for (label o : m.faceVerts) std::printf(" %d", o);
    std::printf("\npoints");
    for (scalar v : m.points) std::printf(" %.17g", v);
    std::printf("\n");
    for (const FoamPatch& p : m.patches) std::printf("patch %s %s %d %d\n", p.name.c_str(), p.type.c_str(), p.startFace, p.nFaces);

    // ---- a field ----
    const FoamField T = readField(caseDir + "/0/T", m);
    std::printf("T internal %zu %.17g\n", T.internal.size(), T.internal.empty() ? 0.0 : T.internal[0]);
    for (const FoamPatchField& p : T.boundary)
        std::printf("T patch %s %s %d %.17g\n", p.name.c_str(), p.type.c_str(), p.hasValue ? 1 : 0,
                    p.value.empty() ? 0.0 : p.value[0]);

    // ---- write and read back: a nonuniform vector field and a scalar one ----
    FoamField U;
    U.nComp = 3;
    for (int c = 0; c < m.nCells(); ++c)
        U.internal.insert(U.internal.end(), {std::sin(0.1 * c) / 3.0, -1e-300 * c, 1.0 / (c + 7.0)});
    for (const FoamPatch& p : m.patches)
    {
        FoamPatchField pf;
        pf.name = p.name;
        pf.type = p.nFaces % 2 ? "zeroGradient" : "fixedValue";
        pf.hasValue = pf.type == "fixedValue";
        if (pf.hasValue)
            for (int i = 0; i < p.nFaces; ++i) pf.value.insert(pf.value.end(), {1.0 + i, 2.5, -0.125 * i});
        U.boundary.push_back(pf);
    }
    writeField(scratch + "/U", "U", "[0 1 -1 0 0 0 0]", U);
    const FoamField U2 = readField(scratch + "/U", m);
    bool same = U2.nComp == 3 && U2.internal == U.internal && U2.boundary.size() == U.boundary.size();
    for (size_t p = 0; same && p < U.boundary.size(); ++p)
        same = U2.boundary[p].type == U.boundary[p].type && U2.boundary[p].hasValue == U.boundary[p].hasValue &&
               U2.boundary[p].value == U.boundary[p].value;
    check(same, "vector field round trip");
    FoamField S;
    S.nComp = 1;
    for (int c = 0; c < m.nCells(); ++c) S.internal.push_back(std::exp(-0.3 * c) * 1e-7);
    for (const FoamPatch& p : m.patches)
    {
        FoamPatchField pf;
        pf.name = p.name;
        pf.type = "calculated";
        pf.hasValue = true;
        pf.value.assign(p.nFaces, 0.000851986);
        S.boundary.push_back(pf);
    }
    writeField(scratch + "/nut", "nut", "[0 2 -1 0 0 0 0]", S);
    const FoamField S2 = readField(scratch + "/nut", m);
    same = S2.nComp == 1 && S2.internal == S.internal;
    for (size_t p = 0; same && p < S.boundary.size(); ++p) same = S2.boundary[p].value == S.boundary[p].value;
    check(same, "scalar field round trip");
    std::printf("roundtrip %d\n", failures == 0 ? 1 : 0);

    // ---- the case's own fvSolution: its GAMG entry asks for directSolveCoarsest, refused by name ----
    {
        const FoamDict d = readFoamDict(caseDir + "/system/fvSolution");
        const FoamDict* T = d.subDict("solvers") ? d.subDict("solvers")->subDict("T") : nullptr;
        check(T != nullptr, "case fvSolution solvers/T");
        const std::string msg = T ? refusal([&] { readGamgControls(*T, 1e-10, 0.0, 1000, "T: "); }) : "";
        std::printf("case_refusal %s\n", msg.c_str());
        check(msg.find("directSolveCoarsest") != std::string::npos, "directSolveCoarsest refused");
    }

    // ---- a usable entry, with the regex keys of the same file ----
    {
        const FoamDict d = readFoamDict(fvs);
        const FoamDict* solvers = d.subDict("solvers");
        const FoamDict* p = solvers ? solvers->subDict("p") : nullptr;
        check(p != nullptr, "fvSolution solvers/p");
        if (p)
        {
            const FoamGamgControls c =
                readGamgControls(*p, p->scalarOr("tolerance", 1e-6), p->scalarOr("relTol", 0.0), p->intOr("maxIter", 1000), "p: ");
            std::printf("gamg %s %.17g %.17g %d %d %d %d %d %d %d %d %d %d %d\n", c.smoother.c_str(), c.tolerance, c.relTol,
                        c.maxIter, c.minIter, c.nPreSweeps, c.preSweepsLevelMultiplier, c.maxPreSweeps, c.nPostSweeps,
                        c.postSweepsLevelMultiplier, c.maxPostSweeps, c.nFinestSweeps, c.scaleCorrection ? 1 : 0,
                        c.nCellsInCoarsestLevel);
        }
        // "(U|k|epsilon|omega|f|v2)" answers for k, a literal U entry wins over it for U
        check(solvers && solvers->subDict("k") != nullptr, "regex key (U|k|...) matches k");
        check(solvers && solvers->subDict("U") && solvers->subDict("U")->wordOr("solver", "") == "PBiCGStab",
              "literal U wins over the regex");
        const FoamDict* eq = d.subDict("relaxationFactors") ? d.subDict("relaxationFactors")->subDict("equations") : nullptr;
        if (eq) std::printf("relax %.17g %.17g\n", eq->scalarOr("U", -1.0), eq->scalarOr("k", -1.0));
    }

    // ---- every refusal, by name ----
    const std::pair<std::string, std::string> cases[] = {
        {"", ""},
        {"smoother DILU;", "asymmetric"},
        {"smoother FDIC;", "not carried"},
        {"agglomerator algebraicPair;", "agglomerator"},
        {"mergeLevels 2;", "mergeLevels"},
        {"updateInterval 5;", "updateInterval"},
        {"cacheAgglomeration no;", "cacheAgglomeration"},
        {"interpolateCorrection yes;", "interpolateCorrection"},
        {"directSolveCoarsest on;", "directSolveCoarsest"},
        {"coarsestLevelCorr { solver PCG; }", "coarsestLevelCorr"},
        {"processorAgglomerator procFaces;", "processorAgglomerator"},
        {"scaleCorrection maybe;", "Switch"},
        {"nPreSweeps 1.5;", "integer"},
    };
    for (const auto& cs : cases)
    {
        std::string text = gamgEntry(cs.first);
        const FoamDict d = parseFoamDict(text);
        const std::string msg = refusal([&] { readGamgControls(*d.subDict("p"), 1e-7, 0.01, 1000, "p: "); });
        const bool ok = cs.second.empty() ? msg.empty() : msg.find(cs.second) != std::string::npos;
        check(ok, "entry `" + cs.first + "` -> " + (msg.empty() ? "accepted" : msg));
        std::printf("refusal %d %s\n", ok ? 1 : 0, cs.first.empty() ? "(none)" : cs.first.c_str());
    }
    {
        const FoamDict d = parseFoamDict("p { solver GAMG; tolerance 1e-7; }");
        const std::string msg = refusal([&] { readGamgControls(*d.subDict("p"), 1e-7, 0, 1000, "p: "); });
        check(msg.find("smoother") != std::string::npos, "a missing smoother is refused");
    }
    {
        // the non-default controls are read, and a sub-dictionary later in the file replaces an earlier one
        const FoamDict d = parseFoamDict(
            "p { smoother DIC; } p { smoother symGaussSeidel; nPreSweeps 1; preSweepsLevelMultiplier 2; "
            "maxPreSweeps 5; nPostSweeps 3; postSweepsLevelMultiplier 1; maxPostSweeps 6; nFinestSweeps 4; "
            "scaleCorrection off; nCellsInCoarsestLevel 25; minIter 2; }");
        const FoamGamgControls c = readGamgControls(*d.subDict("p"), 1e-6, 0.0, 50, "p: ");
        check(c.smoother == "symGaussSeidel" && c.nPreSweeps == 1 && c.preSweepsLevelMultiplier == 2 &&
                  c.maxPreSweeps == 5 && c.nPostSweeps == 3 && c.maxPostSweeps == 6 && c.nFinestSweeps == 4 &&
                  !c.scaleCorrection && c.nCellsInCoarsestLevel == 25 && c.minIter == 2 && c.maxIter == 50,
              "non-default controls");
    }
    // FvMatrix's addEqual is term by term
    {
        FvScalarMatrix a, b;
        a.diag = {1.0, 2.0};
        b.diag = {0.5, -1.0};
        a.upper = a.lower = {3.0};
        b.upper = {1.0};
        b.lower = {2.0};
        a.source = {0.0, 1.0};
        b.source = {4.0, 4.0};
        a.internalCoeffs = {{1.0}};
        a.boundaryCoeffs = {{2.0}};
        b.internalCoeffs = {{0.25}};
        b.boundaryCoeffs = {{-2.0}};
        addEqual(a, b, -1.0);
        check(a.diag[0] == 0.5 && a.diag[1] == 3.0 && a.upper[0] == 2.0 && a.lower[0] == 1.0 && a.source[1] == -3.0 &&
                  a.internalCoeffs[0][0] == 0.75 && a.boundaryCoeffs[0][0] == 4.0,
              "addEqual");
    }
    std::printf("failures %d\n", failures);
    return failures == 0 ? 0 : 1;
}
