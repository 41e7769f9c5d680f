// OpenFOAM's on-disk formats and host-side types for brae's C++ callers (header-only, host code, no HIP):
//
//   FoamDict          a parsed OpenFOAM dictionary (the reference's FoamDict, src/OpenFOAM/db/foam_dict.cuh:62-160):
//                     leaves (key -> tokens) and sub-dictionaries, literal keys first, else the last matching
//                     regex key ("(U|k|epsilon)", "p.*"), as OpenFOAM looks keys up
//   readFoamDict      the tokenizer/parser of an ASCII dictionary file (src/OpenFOAM/db/foam_token_reader.cu)
//   readGamgControls  an fvSolution GAMG entry (src/matrices/lduMatrix/solvers/GAMG/gamg_solver_cpp.cu:801-928):
//                     the controls GAMGSolver reads, with its defaults, and a refusal by name of each entry
//                     that selects a branch the GAMG here does not have
//   PolyMesh          constant/polyMesh: points, faces, owner, neighbour, boundary
//                     (src/OpenFOAM/meshes/primitive_mesh.cu, PrimitiveMesh::read)
//   readField /       an ASCII volScalarField or volVectorField: internalField and boundaryField, uniform or
//   writeField        nonuniform (src/finiteVolume/fields/foam_field_reader.cuh, foam_field_writer.cuh)
//   FvMatrix<T>       an assembled fvMatrix on the host (src/OpenFOAM/matrices/ldu_matrix.cuh:10-55) and addEqual
//
// Errors throw std::runtime_error("brae foam: ...").
#pragma once

#include <cctype>
#include <cstdio>
#include <fstream>
#include <memory>
#include <regex>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace brae {

#ifndef BRAE_SCALAR_DEFINED
#define BRAE_SCALAR_DEFINED
using scalar = double;
using label = int;
#endif

namespace foam_detail {

[[noreturn]] inline void fail(const std::string& what) { throw std::runtime_error("brae foam: " + what); }

// tokens of an OpenFOAM text: words (with dots, signs, quotes kept whole), and the punctuation ; { } ( ) [ ]
inline std::vector<std::string> tokenize(const std::string& text)
{
    std::vector<std::string> out;
    const size_t n = text.size();
    size_t i = 0;
    while (i < n)
    {
        const char c = text[i];
        if (std::isspace(static_cast<unsigned char>(c)))
        {
            ++i;
            continue;
        }
        if (c == '/' && i + 1 < n && text[i + 1] == '/')
        {
            while (i < n && text[i] != '\n') ++i;
            continue;
        }
        if (c == '/' && i + 1 < n && text[i + 1] == '*')
        {
            const size_t e = text.find("*/", i + 2);
            if (e == std::string::npos) fail("unterminated comment");
            i = e + 2;
            continue;
        }
        if (c == '"')
        {
            const size_t e = text.find('"', i + 1);
            if (e == std::string::npos) fail("unterminated string");
            out.push_back(text.substr(i, e - i + 1));
            i = e + 1;
            continue;
        }
        if (c == ';' || c == '{' || c == '}' || c == '(' || c == ')' || c == '[' || c == ']')
        {
            out.emplace_back(1, c);
            ++i;
            continue;
        }
        size_t j = i;
        while (j < n)
        {
            const char d = text[j];
            if (std::isspace(static_cast<unsigned char>(d)) || d == ';' || d == '{' || d == '}' || d == '(' ||
                d == ')' || d == '[' || d == ']' || d == '"')
                break;
            if (d == '/' && j + 1 < n && (text[j + 1] == '/' || text[j + 1] == '*')) break;
            ++j;
        }
        out.push_back(text.substr(i, j - i));
        i = j;
    }
    return out;
}

inline std::string unquote(const std::string& s)
{
    return s.size() >= 2 && s.front() == '"' && s.back() == '"' ? s.substr(1, s.size() - 2) : s;
}

// an OpenFOAM key pattern ("(U|k).*") as a std::regex over the whole key
inline bool keyMatches(const std::string& pattern, const std::string& key)
{
    try
    {
        return std::regex_match(key, std::regex(pattern));
    }
    catch (const std::regex_error&)
    {
        return false;
    }
}

inline std::string readText(const std::string& path)
{
    std::ifstream in(path, std::ios::binary);
    if (!in) fail("cannot open " + path);
    std::stringstream ss;
    ss << in.rdbuf();
    return ss.str();
}

} // namespace foam_detail

struct FoamDict
{
    std::vector<std::pair<std::string, std::vector<std::string>>> leaves;   // key -> value tokens
    std::vector<std::pair<std::string, FoamDict>> subs;                     // key -> sub-dictionary

    // literal key first, else the last regex key that matches (OpenFOAM's dictionary lookup)
    const FoamDict* subDict(const std::string& key) const
    {
        for (const auto& s : subs)
            if (foam_detail::unquote(s.first) == key) return &s.second;
        const FoamDict* hit = nullptr;
        for (const auto& s : subs)
            if (s.first.size() && s.first.front() == '"' && foam_detail::keyMatches(foam_detail::unquote(s.first), key))
                hit = &s.second;
        return hit;
    }
    const FoamDict* optionalSubDict(const std::string& key) const
    {
        const FoamDict* d = subDict(key);
        return d ? d : this;
    }
    const std::vector<std::string>* find(const std::string& key) const
    {
        for (const auto& l : leaves)
            if (foam_detail::unquote(l.first) == key) return &l.second;
        const std::vector<std::string>* hit = nullptr;
        for (const auto& l : leaves)
            if (l.first.size() && l.first.front() == '"' && foam_detail::keyMatches(foam_detail::unquote(l.first), key))
                hit = &l.second;
        return hit;
    }
    bool found(const std::string& key) const { return find(key) != nullptr || subDict(key) != nullptr; }
    std::string wordOr(const std::string& key, const std::string& def) const
    {
        const auto* v = find(key);
        return v && !v->empty() ? foam_detail::unquote(v->front()) : def;
    }
    int intOr(const std::string& key, int def) const
    {
        const auto* v = find(key);
        if (!v || v->empty()) return def;
        try
        {
            size_t pos = 0;
            const int r = std::stoi(v->front(), &pos);
            if (pos != v->front().size()) throw std::invalid_argument("trailing");
            return r;
        }
        catch (const std::exception&)
        {
            foam_detail::fail("`" + key + " " + v->front() + "` is not an integer");
        }
    }
    scalar scalarOr(const std::string& key, scalar def) const
    {
        const auto* v = find(key);
        if (!v || v->empty()) return def;
        try
        {
            return std::stod(v->back());
        }
        catch (const std::exception&)
        {
            foam_detail::fail("`" + key + "` is not a number");
        }
    }
};

namespace foam_detail {

// a dictionary body from tokens[i], up to the closing brace (or the end when top-level)
inline size_t parseBody(const std::vector<std::string>& t, size_t i, FoamDict& d, bool topLevel)
{
    while (i < t.size())
    {
        if (t[i] == "}")
        {
            if (topLevel) fail("unbalanced }");
            return i + 1;
        }
        if (t[i] == ";")
        {
            ++i;
            continue;
        }
        const std::string key = t[i++];
        if (i < t.size() && t[i] == "{")
        {
            FoamDict sub;
            i = parseBody(t, i + 1, sub, false);
            // a repeated key replaces the earlier entry, as OpenFOAM's dictionary merges by default
            bool replaced = false;
            for (auto& s : d.subs)
                if (s.first == key)
                {
                    s.second = std::move(sub);
                    replaced = true;
                }
            if (!replaced) d.subs.emplace_back(key, std::move(sub));
            continue;
        }
        std::vector<std::string> vals;
        int depth = 0;
        while (i < t.size())
        {
            const std::string& tok = t[i];
            if (depth == 0 && tok == ";") break;
            if (depth == 0 && tok == "}") fail("entry `" + key + "` is not terminated by ;");
            if (tok == "(" || tok == "[" || tok == "{") ++depth;
            if (tok == ")" || tok == "]" || tok == "}") --depth;
            vals.push_back(tok);
            ++i;
        }
        if (i >= t.size()) fail("entry `" + key + "` is not terminated by ;");
        ++i;
        bool replaced = false;
        for (auto& l : d.leaves)
            if (l.first == key)
            {
                l.second = vals;
                replaced = true;
            }
        if (!replaced) d.leaves.emplace_back(key, std::move(vals));
    }
    if (!topLevel) fail("unbalanced {");
    return i;
}

} // namespace foam_detail

// an ASCII OpenFOAM dictionary (the FoamFile header is kept as a sub-dictionary like any other)
inline FoamDict parseFoamDict(const std::string& text)
{
    const std::vector<std::string> t = foam_detail::tokenize(text);
    FoamDict d;
    foam_detail::parseBody(t, 0, d, true);
    return d;
}

inline FoamDict readFoamDict(const std::string& path) { return parseFoamDict(foam_detail::readText(path)); }

// ---------------- fvSolution GAMG entries ----------------

// GAMGSolver's controls with GAMGSolver.C's defaults (gamg_solver_cpp.cuh:67-85); the tolerances are
// lduMatrix::solver's
struct FoamGamgControls
{
    std::string smoother;
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

// the smoothers the GAMG carries (gamg_solver_cpp.cuh:134)
inline bool foamGamgSmootherCarried(const std::string& s)
{
    return s == "GaussSeidel" || s == "symGaussSeidel" || s == "DIC" || s == "DICGaussSeidel";
}

// readGamgControls (gamg_solver_cpp.cu:801-928): `who` opens every message (the solver and field). Refuses,
// by name: no smoother; a smoother the GAMG does not carry (DILU/DILUGaussSeidel are the asymmetric table's,
// which OpenFOAM itself refuses for the symmetric pressure matrix); an agglomerator other than faceAreaPair;
// mergeLevels != 1; updateInterval != 1; cacheAgglomeration off; interpolateCorrection; directSolveCoarsest;
// a coarsestLevelCorr sub-dictionary; a processorAgglomerator.
inline FoamGamgControls readGamgControls(const FoamDict& d, scalar tol, scalar relTol, int maxIter,
                                         const std::string& who)
{
    auto switchOr = [&](const std::string& key, bool def) {
        if (!d.found(key)) return def;
        const std::string w = d.wordOr(key, "");
        if (w == "yes" || w == "true" || w == "on" || w == "y" || w == "t" || w == "1") return true;
        if (w == "no" || w == "false" || w == "off" || w == "n" || w == "f" || w == "0") return false;
        foam_detail::fail(who + "`" + key + " " + w + "` is not a Switch");
    };
    FoamGamgControls c;
    c.smoother = d.wordOr("smoother", "");
    c.tolerance = tol;
    c.relTol = relTol;
    c.maxIter = maxIter;
    c.minIter = d.intOr("minIter", 0);
    c.nPreSweeps = d.intOr("nPreSweeps", c.nPreSweeps);
    c.preSweepsLevelMultiplier = d.intOr("preSweepsLevelMultiplier", c.preSweepsLevelMultiplier);
    c.maxPreSweeps = d.intOr("maxPreSweeps", c.maxPreSweeps);
    c.nPostSweeps = d.intOr("nPostSweeps", c.nPostSweeps);
    c.postSweepsLevelMultiplier = d.intOr("postSweepsLevelMultiplier", c.postSweepsLevelMultiplier);
    c.maxPostSweeps = d.intOr("maxPostSweeps", c.maxPostSweeps);
    c.nFinestSweeps = d.intOr("nFinestSweeps", c.nFinestSweeps);
    c.scaleCorrection = switchOr("scaleCorrection", true);
    c.nCellsInCoarsestLevel = static_cast<label>(d.intOr("nCellsInCoarsestLevel", 10));
    if (c.smoother.empty()) foam_detail::fail(who + "names no `smoother` (OpenFOAM's lookup of it is mandatory)");
    if (!foamGamgSmootherCarried(c.smoother))
    {
        if (c.smoother == "DILU" || c.smoother == "DILUGaussSeidel")
            foam_detail::fail(who + "`smoother " + c.smoother + "` is registered for asymmetric matrices only; the "
                              "pressure matrix is symmetric and OpenFOAM refuses it as well");
        foam_detail::fail(who + "`smoother " + c.smoother + "` is not carried (GaussSeidel, symGaussSeidel, DIC, "
                          "DICGaussSeidel)");
    }
    const std::string agglomerator = d.wordOr("agglomerator", "faceAreaPair");
    if (agglomerator != "faceAreaPair")
        foam_detail::fail(who + "`agglomerator " + agglomerator + "`: only faceAreaPair is carried");
    if (d.intOr("mergeLevels", 1) != 1)
        foam_detail::fail(who + "`mergeLevels " + std::to_string(d.intOr("mergeLevels", 1)) +
                          "`: combined pairwise levels are not carried by OpenFOAM's GAMG here");
    if (d.intOr("updateInterval", 1) != 1)
        foam_detail::fail(who + "`updateInterval`: only the default weighting by the perturbed area vector is carried");
    if (!switchOr("cacheAgglomeration", true))
        foam_detail::fail(who + "`cacheAgglomeration no`: the hierarchy is built once per mesh here");
    if (switchOr("interpolateCorrection", false))
        foam_detail::fail(who + "`interpolateCorrection yes` is not carried");
    if (switchOr("directSolveCoarsest", false))
        foam_detail::fail(who + "`directSolveCoarsest yes`: the coarsest level is solved by PCG-DIC here");
    if (d.subDict("coarsestLevelCorr"))
        foam_detail::fail(who + "a `coarsestLevelCorr` sub-dictionary: the coarsest level's own solver is not carried");
    if (d.found("processorAgglomerator"))
        foam_detail::fail(who + "a `processorAgglomerator`: this is a one-processor solver");
    return c;
}

// ---------------- polyMesh ----------------

struct FoamPatch
{
    std::string name, type;
    label startFace = 0, nFaces = 0;
};

// constant/polyMesh as PrimitiveMesh::read takes it: faces in CSR (faceOffsets [nFaces+1], faceVerts)
struct PolyMesh
{
    std::vector<scalar> points;                 // xyz per point
    std::vector<label> faceOffsets, faceVerts;
    std::vector<label> owner, neighbour;        // owner over every face, neighbour over the internal ones
    std::vector<FoamPatch> patches;
    label nCells() const
    {
        label n = -1;
        for (label o : owner) n = std::max(n, o);
        for (label o : neighbour) n = std::max(n, o);
        return n + 1;
    }
    label nInternalFaces() const { return static_cast<label>(neighbour.size()); }
    label nFaces() const { return static_cast<label>(owner.size()); }
};

namespace foam_detail {

// the body after the FoamFile header: the tokens of the single list `N ( ... )`
inline std::vector<std::string> listTokens(const std::string& path, size_t& count)
{
    const std::vector<std::string> t = tokenize(readText(path));
    size_t i = 0;
    if (i < t.size() && t[i] == "FoamFile")
    {
        while (i < t.size() && t[i] != "}") ++i;
        ++i;
    }
    if (i >= t.size()) fail(path + ": no list");
    count = static_cast<size_t>(std::stoull(t[i]));
    if (i + 1 >= t.size() || t[i + 1] != "(") fail(path + ": a list must open with N (");
    std::vector<std::string> body(t.begin() + static_cast<long>(i) + 2, t.end());
    while (!body.empty() && body.back() != ")") body.pop_back();
    if (body.empty()) fail(path + ": the list is not closed");
    body.pop_back();
    return body;
}

inline std::vector<label> readLabels(const std::string& path)
{
    size_t n = 0;
    const std::vector<std::string> b = listTokens(path, n);
    if (b.size() != n) fail(path + ": " + std::to_string(b.size()) + " labels where the list says " + std::to_string(n));
    std::vector<label> out(n);
    for (size_t k = 0; k < n; ++k) out[k] = std::stoi(b[k]);
    return out;
}

} // namespace foam_detail

inline PolyMesh readPolyMesh(const std::string& dir)
{
    using namespace foam_detail;
    PolyMesh m;
    {
        size_t n = 0;
        const std::vector<std::string> b = listTokens(dir + "/points", n);
        m.points.reserve(3 * n);
        for (size_t k = 0; k < b.size(); ++k)
        {
            if (b[k] == "(" || b[k] == ")") continue;
            m.points.push_back(std::stod(b[k]));
        }
        if (m.points.size() != 3 * n) fail(dir + "/points: not " + std::to_string(n) + " (x y z) triples");
    }
    {
        size_t n = 0;
        const std::vector<std::string> b = listTokens(dir + "/faces", n);
        m.faceOffsets.assign(1, 0);
        size_t k = 0;
        while (k < b.size())
        {
            const int nv = std::stoi(b[k]);
            if (k + 1 >= b.size() || b[k + 1] != "(") fail(dir + "/faces: a face must read N(v0 v1 ...)");
            k += 2;
            for (int j = 0; j < nv; ++j) m.faceVerts.push_back(std::stoi(b[k++]));
            if (k >= b.size() || b[k] != ")") fail(dir + "/faces: a face's vertex count does not match its list");
            ++k;
            m.faceOffsets.push_back(static_cast<label>(m.faceVerts.size()));
        }
        if (m.faceOffsets.size() != n + 1) fail(dir + "/faces: face count differs from the list's");
    }
    m.owner = readLabels(dir + "/owner");
    m.neighbour = readLabels(dir + "/neighbour");
    if (m.owner.size() + 1 != m.faceOffsets.size()) fail(dir + ": owner and faces differ in length");
    {
        const std::vector<std::string> t = tokenize(readText(dir + "/boundary"));
        size_t i = 0;
        if (i < t.size() && t[i] == "FoamFile")
        {
            while (i < t.size() && t[i] != "}") ++i;
            ++i;
        }
        const size_t n = static_cast<size_t>(std::stoull(t.at(i)));
        FoamDict d;
        std::vector<std::string> inner(t.begin() + static_cast<long>(i) + 2, t.end());
        while (!inner.empty() && inner.back() != ")") inner.pop_back();
        inner.pop_back();
        parseBody(inner, 0, d, true);
        if (d.subs.size() != n) fail(dir + "/boundary: " + std::to_string(d.subs.size()) + " patches, the list says " +
                                     std::to_string(n));
        for (const auto& s : d.subs)
            m.patches.push_back(FoamPatch{s.first, s.second.wordOr("type", "patch"), s.second.intOr("startFace", 0),
                                          s.second.intOr("nFaces", 0)});
    }
    label at = m.nInternalFaces();
    for (const FoamPatch& p : m.patches)
    {
        if (p.startFace != at) fail(dir + "/boundary: patch " + p.name + " does not start where the last one ended");
        at += p.nFaces;
    }
    if (at != m.nFaces()) fail(dir + "/boundary: the patches do not cover the boundary faces");
    return m;
}

// ---------------- fields ----------------

// one patch of a field's boundaryField: its type, and its value per face when the file gives one
struct FoamPatchField
{
    std::string name, type;
    bool hasValue = false;
    std::vector<scalar> value;   // nComp per face
};

// a volScalarField (nComp 1) or volVectorField (nComp 3) as read: internal values nComp per cell
struct FoamField
{
    int nComp = 1;
    std::vector<scalar> internal;
    std::vector<FoamPatchField> boundary;
};

namespace foam_detail {

// `uniform v`, `uniform (x y z)` or `nonuniform List<...> N ( ... )`, expanded to `count` entries
inline std::vector<scalar> values(const std::vector<std::string>& v, int nComp, size_t count, const std::string& what)
{
    std::vector<scalar> out;
    if (v.empty()) fail(what + ": no value");
    if (v[0] == "uniform")
    {
        std::vector<scalar> one;
        for (size_t k = 1; k < v.size(); ++k)
            if (v[k] != "(" && v[k] != ")") one.push_back(std::stod(v[k]));
        if ((int)one.size() != nComp) fail(what + ": a uniform value of the wrong rank");
        for (size_t c = 0; c < count; ++c) out.insert(out.end(), one.begin(), one.end());
        return out;
    }
    if (v[0] != "nonuniform" || v.size() < 4) fail(what + ": neither uniform nor nonuniform");
    const size_t n = static_cast<size_t>(std::stoull(v[2]));
    for (size_t k = 4; k + 1 < v.size(); ++k)
        if (v[k] != "(" && v[k] != ")") out.push_back(std::stod(v[k]));
    if (out.size() != n * nComp) fail(what + ": " + std::to_string(out.size() / nComp) + " values, the list says " +
                                      std::to_string(n));
    if (count != n && n != 0) fail(what + ": " + std::to_string(n) + " values for " + std::to_string(count) + " entries");
    return out;
}

} // namespace foam_detail

// patch sizes from the mesh fill `uniform` values; a patch without `value` keeps hasValue false
inline FoamField readField(const std::string& path, const PolyMesh& mesh)
{
    using namespace foam_detail;
    const FoamDict d = readFoamDict(path);
    const FoamDict* hdr = d.subDict("FoamFile");
    const std::string cls = hdr ? hdr->wordOr("class", "") : "";
    FoamField f;
    if (cls == "volScalarField")
        f.nComp = 1;
    else if (cls == "volVectorField")
        f.nComp = 3;
    else
        fail(path + ": class `" + cls + "` is not carried (volScalarField, volVectorField)");
    const auto* internal = d.find("internalField");
    if (!internal) fail(path + ": no internalField");
    f.internal = values(*internal, f.nComp, static_cast<size_t>(mesh.nCells()), path + " internalField");
    const FoamDict* bf = d.subDict("boundaryField");
    if (!bf) fail(path + ": no boundaryField");
    for (const FoamPatch& p : mesh.patches)
    {
        const FoamDict* pd = bf->subDict(p.name);
        if (!pd) fail(path + ": no boundaryField entry for patch " + p.name);
        FoamPatchField pf;
        pf.name = p.name;
        pf.type = pd->wordOr("type", "");
        if (const auto* v = pd->find("value"))
        {
            pf.hasValue = true;
            pf.value = values(*v, f.nComp, static_cast<size_t>(p.nFaces), path + " " + p.name + " value");
        }
        f.boundary.push_back(std::move(pf));
    }
    return f;
}

// an ASCII field file OpenFOAM reads back: the internal values nonuniform, each patch's type and, when it
// has one, its value list
inline void writeField(const std::string& path, const std::string& name, const std::string& dimensions,
                       const FoamField& f)
{
    std::ofstream out(path);
    if (!out) foam_detail::fail("cannot write " + path);
    out.precision(17);
    const bool vec = f.nComp == 3;
    const std::string cls = vec ? "volVectorField" : "volScalarField";
    const std::string list = vec ? "List<vector>" : "List<scalar>";
    auto writeList = [&](const std::vector<scalar>& v) {
        const size_t n = v.size() / f.nComp;
        out << "nonuniform " << list << "\n" << n << "\n(\n";
        for (size_t i = 0; i < n; ++i)
        {
            if (vec)
                out << "(" << v[3 * i] << " " << v[3 * i + 1] << " " << v[3 * i + 2] << ")\n";
            else
                out << v[i] << "\n";
        }
        out << ")";
    };
    out << "FoamFile\n{\n    version     2.0;\n    format      ascii;\n    class       " << cls
        << ";\n    object      " << name << ";\n}\n\n";
    out << "dimensions      " << dimensions << ";\n\ninternalField   ";
    writeList(f.internal);
    out << ";\n\nboundaryField\n{\n";
    for (const FoamPatchField& p : f.boundary)
    {
        out << "    " << p.name << "\n    {\n        type            " << p.type << ";\n";
        if (p.hasValue)
        {
            out << "        value           ";
            writeList(p.value);
            out << ";\n";
        }
        out << "    }\n";
    }
    out << "}\n";
    if (!out) foam_detail::fail("write failed: " + path);
}

// ---------------- FvMatrix (ldu_matrix.cuh:10-55) ----------------

template <typename T>
struct FvMatrix
{
    std::vector<scalar> diag;                    // nCells, before the boundary's diagonal
    std::vector<scalar> upper, lower;            // nInternalFaces
    std::vector<T> source;                       // nCells, before the boundary's source
    std::vector<std::vector<T>> internalCoeffs;  // [patch][face] -> diag of the face's cell
    std::vector<std::vector<T>> boundaryCoeffs;  // [patch][face] -> source of the face's cell
    std::vector<T> faceFluxCorrection;           // nInternalFaces, or empty
};
using FvScalarMatrix = FvMatrix<scalar>;

// a += s*b on the same addressing (ldu_matrix.cuh:40-55), term by term
template <typename T>
inline void addEqual(FvMatrix<T>& a, const FvMatrix<T>& b, scalar s)
{
    for (size_t i = 0; i < a.diag.size(); ++i) a.diag[i] += s * b.diag[i];
    for (size_t i = 0; i < a.upper.size(); ++i) a.upper[i] += s * b.upper[i];
    for (size_t i = 0; i < a.lower.size(); ++i) a.lower[i] += s * b.lower[i];
    for (size_t i = 0; i < a.source.size(); ++i) a.source[i] += s * b.source[i];
    for (size_t p = 0; p < a.internalCoeffs.size(); ++p)
        for (size_t i = 0; i < a.internalCoeffs[p].size(); ++i)
        {
            a.internalCoeffs[p][i] += s * b.internalCoeffs[p][i];
            a.boundaryCoeffs[p][i] += s * b.boundaryCoeffs[p][i];
        }
}

} // namespace brae
