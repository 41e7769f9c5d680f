"""Write the golden fixtures the parity tests read, from the reference's own data files.

Run in the build container only (it reads the reference checkout; the GPU box never does):

    python tests/golden/make_fixtures.py /root/reference

Everything written is DATA the reference's own tests hold -- OpenFOAM v2412 dumps of the
pitzDaily matrixDump cases and the 20-cell AMG trace CSVs -- converted to compressed .npz so
the tests load them without an OpenFOAM parser:

  pitzdaily.npz  validation/matrixDump/constant/polyMesh (mesh), 282/T boundary types and
                 values, 282/phi, matrix.dat (fvm::laplacian(1,T)), matrix_div.dat
                 (fvm::div(phi,T) upwind), amul.dat (psi, A.psi), solution.dat (PCG-DIC),
                 validation/matrixDumpGAMG/solution.dat (OpenFOAM GAMG)
  simple_step.npz  validation/matrixDumpSimple/282/{U,p,phi}, constant/transportProperties (nu) and
                 step.dat (OpenFOAM's p, U, phi after one SIMPLE iteration from those fields)
  kepsilon.npz   validation/kEpsCase and validation/kEpsCorrect: 282/{U,k,epsilon,nut,phi}, keqn.dat and
                 epseqn.dat (OpenFOAM's k and epsilon equations), correct.dat (k, epsilon, nut after one
                 kEpsilon::correct())
  foam/          OpenFOAM ASCII files as they are: case20's polyMesh, 0/T and system/fvSolution, and
                 pitzDaily's system/fvSolution (validation/matrixDumpSimple), for the readers' tests
  case20.npz     demo/amgpcg/case20/constant/polyMesh, trace/{matrix,rhs,of_hierarchy,
                 brae_hierarchy,brae_levels,brae_cycles,of_gamg_cycles}.csv
"""
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def _body(path):
    """Text of an OpenFOAM file after its FoamFile header."""
    txt = open(path).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    txt = re.sub(r"//[^\n]*", "", txt)
    i = txt.find("FoamFile")
    if i >= 0:
        j = txt.find("}", i)
        txt = txt[j + 1:]
    return txt


def _list_block(txt, start=0):
    """(count, inner text, end index) of the first `N ( ... )` list at or after `start`."""
    m = re.compile(r"(\d+)\s*\(").search(txt, start)
    n = int(m.group(1))
    depth, i = 1, m.end()
    while depth:
        c = txt[i]
        if c == "(":
            depth += 1
        elif c == ")":
            depth -= 1
        i += 1
    return n, txt[m.end():i - 1], i


def read_points(path):
    n, inner, _ = _list_block(_body(path))
    v = np.array(re.findall(r"\(([^()]*)\)", inner)[0:n], dtype=object)
    return np.array([[float(t) for t in s.split()] for s in v], dtype=np.float64)


def read_faces(path):
    n, inner, _ = _list_block(_body(path))
    faces = re.findall(r"(\d+)\(([^()]*)\)", inner)
    assert len(faces) == n
    off = [0]
    verts = []
    for k, s in faces:
        vv = [int(t) for t in s.split()]
        assert len(vv) == int(k)
        verts.extend(vv)
        off.append(len(verts))
    return np.array(off, dtype=np.int32), np.array(verts, dtype=np.int32)


def read_labels(path):
    n, inner, _ = _list_block(_body(path))
    a = np.array(inner.split(), dtype=np.int32)
    assert a.size == n
    return a


def read_boundary(path):
    txt = _body(path)
    _, inner, _ = _list_block(txt)
    out = []
    for m in re.finditer(r"(\w+)\s*\{([^}]*)\}", inner):
        d = dict(re.findall(r"(\w+)\s+([^;]+);", m.group(2)))
        out.append((m.group(1), d["type"].strip(), int(d["startFace"]), int(d["nFaces"])))
    return out


def read_mesh(d):
    pts = read_points(os.path.join(d, "points"))
    off, verts = read_faces(os.path.join(d, "faces"))
    own = read_labels(os.path.join(d, "owner"))
    nei = read_labels(os.path.join(d, "neighbour"))
    bnd = read_boundary(os.path.join(d, "boundary"))
    return dict(points=pts, faceOffsets=off, faceVerts=verts, owner=own, neighbour=nei,
                patchNames=np.array([b[0] for b in bnd]), patchTypes=np.array([b[1] for b in bnd]),
                patchStart=np.array([b[2] for b in bnd], np.int32), patchSize=np.array([b[3] for b in bnd], np.int32))


def read_scalar_field(path, patchNames, patchSize):
    """internal values and per-boundary-face values of a volScalarField/surfaceScalarField."""
    txt = _body(path)
    m = re.search(r"internalField\s+(uniform|nonuniform)", txt)
    if m.group(1) == "uniform":
        internal = np.array([float(re.search(r"internalField\s+uniform\s+([^;]+);", txt).group(1))])
    else:
        _, inner, _ = _list_block(txt, m.end())
        internal = np.array(inner.split(), dtype=np.float64)
    bi = txt.find("boundaryField")
    types, vals = [], []
    for name, size in zip(patchNames, patchSize):
        pm = re.search(r"\b%s\s*\{" % re.escape(str(name)), txt[bi:])
        body_start = bi + pm.end()
        body_end = txt.find("}", body_start)
        body = txt[body_start:body_end]
        t = re.search(r"type\s+(\w+);", body).group(1)
        v = np.zeros(int(size))
        um = re.search(r"value\s+uniform\s+([^;]+);", body)
        nm = re.search(r"value\s+nonuniform", body)
        if um:
            v[:] = float(um.group(1))
        elif nm:
            _, inner, _ = _list_block(txt, body_start + nm.end())
            v = np.array(inner.split(), dtype=np.float64)
            if v.size != size:
                v = np.zeros(int(size))
        types.append(t)
        vals.append(v)
    return internal, types, vals


def read_vector_field(path, patchNames, patchSize):
    """internal (n, 3) values, boundary types and per-boundary-face (size, 3) values of a volVectorField;
    noSlip is a fixedValue of zero, a patch without a value reads as zeros"""
    txt = _body(path)
    m = re.search(r"internalField\s+(uniform|nonuniform)", txt)
    if m.group(1) == "uniform":
        v = re.search(r"internalField\s+uniform\s+\(([^)]*)\)", txt).group(1)
        internal = np.array([[float(t) for t in v.split()]])
    else:
        n, inner, _ = _list_block(txt, m.end())
        internal = np.array([[float(t) for t in s.split()] for s in re.findall(r"\(([^()]*)\)", inner)])
        assert internal.shape == (n, 3)
    bi = txt.find("boundaryField")
    types, vals = [], []
    for name, size in zip(patchNames, patchSize):
        pm = re.search(r"\b%s\s*\{" % re.escape(str(name)), txt[bi:])
        body_start = bi + pm.end()
        body_end = txt.find("}", body_start)
        body = txt[body_start:body_end]
        t = re.search(r"type\s+(\w+);", body).group(1)
        v = np.zeros((int(size), 3))
        um = re.search(r"value\s+uniform\s+\(([^)]*)\)", body)
        if um:
            v[:] = [float(x) for x in um.group(1).split()]
        types.append(t)
        vals.append(v)
    return internal, types, vals


def read_matrix_dat(path):
    tok = open(path).read().split()
    nC, nIf, npch = int(tok[0]), int(tok[1]), int(tok[2])
    k = 3
    arr = lambda n: np.array(tok[k:k + n], dtype=np.float64)
    diag = arr(nC); k += nC
    upper = arr(nIf); k += nIf
    lower = arr(nIf); k += nIf
    source = arr(nC); k += nC
    patches = {}
    for _ in range(npch):
        name, sz = tok[k], int(tok[k + 1]); k += 2
        v = np.array(tok[k:k + 2 * sz], dtype=np.float64).reshape(sz, 2); k += 2 * sz
        patches[name] = v
    return diag, upper, lower, source, patches


def read_solution(path):
    tok = open(path).read().split()
    nC = int(tok[0])
    return float(tok[1]), float(tok[2]), int(tok[3]), np.array(tok[4:4 + nC], dtype=np.float64)


def read_csv(path):
    with open(path) as f:
        head = f.readline().strip().split(",")
        rows = [ln.strip().split(",") for ln in f if ln.strip()]
    return head, rows


def pitzdaily(ref):
    case = os.path.join(ref, "validation/matrixDump")
    mesh = read_mesh(os.path.join(case, "constant/polyMesh"))
    out = dict(mesh)
    Ti, Ttypes, Tvals = read_scalar_field(os.path.join(case, "282/T"), mesh["patchNames"], mesh["patchSize"])
    out["T_types"] = np.array(Ttypes)
    out["T_values"] = np.concatenate(Tvals)
    phiI, _, phiB = read_scalar_field(os.path.join(case, "282/phi"), mesh["patchNames"], mesh["patchSize"])
    out["phi_internal"] = phiI
    out["phi_boundary"] = np.concatenate(phiB)
    for tag, fn in (("lap", "matrix.dat"), ("div", "matrix_div.dat")):
        d, u, l, s, p = read_matrix_dat(os.path.join(case, fn))
        out[tag + "_diag"], out[tag + "_upper"], out[tag + "_lower"], out[tag + "_source"] = d, u, l, s
        for name in mesh["patchNames"]:
            v = p.get(str(name), np.zeros((0, 2)))
            out[tag + "_iC_" + str(name)] = v[:, 0].copy()
            out[tag + "_bC_" + str(name)] = v[:, 1].copy()
    tok = open(os.path.join(case, "amul.dat")).read().split()
    n = int(tok[0])
    a = np.array(tok[1:1 + 2 * n], dtype=np.float64).reshape(n, 2)
    out["amul_psi"], out["amul_Apsi"] = a[:, 0].copy(), a[:, 1].copy()
    for tag, sub in (("pcg", "validation/matrixDump"), ("gamg", "validation/matrixDumpGAMG")):
        i0, f0, it, T = read_solution(os.path.join(ref, sub, "solution.dat"))
        out[tag + "_init"], out[tag + "_final"], out[tag + "_iters"], out[tag + "_T"] = i0, f0, it, T
    np.savez_compressed(os.path.join(HERE, "pitzdaily.npz"), **out)


def simple_step(ref):
    """validation/matrixDumpSimple (pitzDaily, the same mesh as matrixDump): OpenFOAM's step-282 fields
    (U, p, phi with their patch types and values) and its dumpSimpleStep output step.dat (p, U and phi
    after one SIMPLE iteration, phi per patch), which tests/test_simple_step_cpp.cu compares at 1e-9"""
    case = os.path.join(ref, "validation/matrixDumpSimple")
    mesh = read_mesh(os.path.join(case, "constant/polyMesh"))
    names, sizes = mesh["patchNames"], mesh["patchSize"]
    out = {}
    Ui, Ut, Uv = read_vector_field(os.path.join(case, "282/U"), names, sizes)
    out["U0"], out["U_types"], out["U_values"] = Ui, np.array(Ut), np.concatenate(Uv)
    pi, pt, pv = read_scalar_field(os.path.join(case, "282/p"), names, sizes)
    out["p0"], out["p_types"], out["p_values"] = pi, np.array(pt), np.concatenate(pv)
    phiI, _, phiB = read_scalar_field(os.path.join(case, "282/phi"), names, sizes)
    out["phi0_internal"], out["phi0_boundary"] = phiI, np.concatenate(phiB)
    nu = re.search(r"\bnu\s+([^;]+);", _body(os.path.join(case, "constant/transportProperties"))).group(1)
    out["nu"] = float(nu.split()[-1])
    tok = open(os.path.join(case, "step.dat")).read().split()
    nC, nIf, npch = int(tok[0]), int(tok[1]), int(tok[2])
    k = 3
    out["step_p"] = np.array(tok[k:k + nC], dtype=np.float64); k += nC
    out["step_U"] = np.array(tok[k:k + 3 * nC], dtype=np.float64).reshape(nC, 3); k += 3 * nC
    out["step_phi_internal"] = np.array(tok[k:k + nIf], dtype=np.float64); k += nIf
    for _ in range(npch):
        name, sz = tok[k], int(tok[k + 1]); k += 2
        out["step_phi_" + name] = np.array(tok[k:k + sz], dtype=np.float64); k += sz
    assert k == len(tok)
    np.savez_compressed(os.path.join(HERE, "simple_step.npz"), **out)


def kepsilon(ref):
    """validation/kEpsCase and validation/kEpsCorrect (pitzDaily, kEpsilon, nu 1e-5): OpenFOAM's step-282 fields
    U, k, epsilon, nut and phi with their patch types and values; keqn.dat/epseqn.dat (OpenFOAM's assembled
    k and epsilon equations from kEpsCase's fields, compared at 1e-11 by tests/test_kepsilon_eqn.cu) and
    correct.dat (k, epsilon, nut after one kEpsilon::correct() from kEpsCorrect's fields,
    tests/test_kepsilon_correct.cu)"""
    out = {}
    for tag, sub in (("eqn", "validation/kEpsCase"), ("corr", "validation/kEpsCorrect")):
        case = os.path.join(ref, sub)
        mesh = read_mesh(os.path.join(case, "constant/polyMesh"))
        names, sizes = mesh["patchNames"], mesh["patchSize"]
        if tag == "eqn":
            out.update({"mesh_" + k: v for k, v in mesh.items()})
        Ui, Ut, Uv = read_vector_field(os.path.join(case, "282/U"), names, sizes)
        out[tag + "_U0"], out[tag + "_U_types"], out[tag + "_U_values"] = Ui, np.array(Ut), np.concatenate(Uv)
        for f in ("k", "epsilon", "nut"):
            i0, t0, v0 = read_scalar_field(os.path.join(case, "282", f), names, sizes)
            out[tag + "_" + f + "0"], out[tag + "_" + f + "_types"] = i0, np.array(t0)
            out[tag + "_" + f + "_values"] = np.concatenate(v0)
        phiI, _, phiB = read_scalar_field(os.path.join(case, "282/phi"), names, sizes)
        out[tag + "_phi0_internal"], out[tag + "_phi0_boundary"] = phiI, np.concatenate(phiB)
    for eq in ("k", "eps"):
        d, u, l, src, pch = read_matrix_dat(os.path.join(ref, "validation/kEpsCase", eq + "eqn.dat"))
        out[eq + "eqn_diag"], out[eq + "eqn_upper"], out[eq + "eqn_lower"], out[eq + "eqn_source"] = d, u, l, src
        for name in out["mesh_patchNames"]:
            v = pch.get(str(name), np.zeros((0, 2)))
            out[eq + "eqn_iC_" + str(name)] = v[:, 0].copy()
            out[eq + "eqn_bC_" + str(name)] = v[:, 1].copy()
    tok = open(os.path.join(ref, "validation/kEpsCorrect/correct.dat")).read().split()
    nC = int(tok[0])
    out["correct_k"] = np.array(tok[1:1 + nC], dtype=np.float64)
    out["correct_epsilon"] = np.array(tok[1 + nC:1 + 2 * nC], dtype=np.float64)
    out["correct_nut"] = np.array(tok[1 + 2 * nC:1 + 3 * nC], dtype=np.float64)
    np.savez_compressed(os.path.join(HERE, "kepsilon.npz"), **out)


def case20(ref):
    base = os.path.join(ref, "demo/amgpcg/case20")
    out = dict(read_mesh(os.path.join(base, "constant/polyMesh")))
    tr = os.path.join(base, "trace")
    _, rows = read_csv(os.path.join(tr, "matrix.csv"))
    diag = [r for r in rows if r[0] == "diag"]
    upper = [r for r in rows if r[0] == "upper"]
    lower = [r for r in rows if r[0] == "lower"]
    out["m_diag"] = np.array([float(r[4]) for r in sorted(diag, key=lambda r: int(r[1]))])
    out["m_owner"] = np.array([int(r[2]) for r in sorted(upper, key=lambda r: int(r[1]))], np.int32)
    out["m_neighbour"] = np.array([int(r[3]) for r in sorted(upper, key=lambda r: int(r[1]))], np.int32)
    out["m_upper"] = np.array([float(r[4]) for r in sorted(upper, key=lambda r: int(r[1]))])
    out["m_lower"] = np.array([float(r[4]) for r in sorted(lower, key=lambda r: int(r[1]))])
    _, rows = read_csv(os.path.join(tr, "rhs.csv"))
    out["rhs"] = np.array([float(r[1]) for r in sorted(rows, key=lambda r: int(r[0]))])
    for tag in ("of_hierarchy", "brae_hierarchy"):
        _, rows = read_csv(os.path.join(tr, tag + ".csv"))
        lv = [r for r in rows if int(r[1]) >= 0]
        out[tag] = np.array([[int(r[0]), int(r[1]), int(r[2])] for r in lv], np.int32)
    _, rows = read_csv(os.path.join(tr, "brae_levels.csv"))
    out["brae_levels"] = np.array([[int(r[0]), {"diag": 0, "upper": 1, "lower": 2}[r[1]], int(r[2]), int(r[3]),
                                    int(r[4]), float(r[5])] for r in rows], np.float64)
    for tag in ("brae_cycles", "of_gamg_cycles"):
        _, rows = read_csv(os.path.join(tr, tag + ".csv"))
        out[tag] = np.array([[float(x) for x in r] for r in rows], np.float64)
    np.savez_compressed(os.path.join(HERE, "case20.npz"), **out)


def foam_files(ref):
    """OpenFOAM ASCII files kept as they are, for the readers' tests: case20's polyMesh, its 0/T and its
    system/fvSolution (a GAMG entry with directSolveCoarsest), and pitzDaily's system/fvSolution (a GAMG p
    entry, regex keys, relaxation factors) from validation/matrixDumpSimple"""
    import shutil
    src = os.path.join(ref, "demo/amgpcg/case20")
    dst = os.path.join(HERE, "foam", "case20")
    for sub in ("constant/polyMesh/points", "constant/polyMesh/faces", "constant/polyMesh/owner",
                "constant/polyMesh/neighbour", "constant/polyMesh/boundary", "0/T", "system/fvSolution"):
        os.makedirs(os.path.dirname(os.path.join(dst, sub)), exist_ok=True)
        shutil.copyfile(os.path.join(src, sub), os.path.join(dst, sub))
    os.makedirs(os.path.join(HERE, "foam", "pitzDaily"), exist_ok=True)
    shutil.copyfile(os.path.join(ref, "validation/matrixDumpSimple/system/fvSolution"),
                    os.path.join(HERE, "foam", "pitzDaily", "fvSolution"))


if __name__ == "__main__":
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    pitzdaily(ref)
    simple_step(ref)
    kepsilon(ref)
    case20(ref)
    foam_files(ref)
    for f in ("pitzdaily.npz", "simple_step.npz", "kepsilon.npz", "case20.npz"):
        print(f, os.path.getsize(os.path.join(HERE, f)), "bytes")
