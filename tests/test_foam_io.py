"""OpenFOAM's on-disk formats through include/brae_foam.hpp (host C++, no device): tests/cpp/foam_io.cpp
reads case20's polyMesh, its 0/T and two fvSolution files, writes fields and reads them back, and reads GAMG
entries with every refusal by name. Its mesh is compared here with the same files parsed independently
(tests/golden/make_fixtures.py -> case20.npz); the GAMG controls with the entry's own text."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "cpp", "foam_io.bin")
CASE = os.path.join(HERE, "golden", "foam", "case20")
FVS = os.path.join(HERE, "golden", "foam", "pitzDaily", "fvSolution")


@pytest.fixture(scope="module")
def out():
    if not os.path.exists(EXE):
        from brae_amd import build as B
        B.build_cpp_callers()
    with tempfile.TemporaryDirectory() as tmp:
        r = subprocess.run([EXE, CASE, FVS, tmp], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = {}
    for ln in r.stdout.splitlines():
        key, _, rest = ln.partition(" ")
        lines.setdefault(key, []).append(rest)
    return lines


def test_polymesh_matches_the_independent_parse(out):
    d = np.load(os.path.join(HERE, "golden", "case20.npz"))
    nC, nF, nI, nP = (int(x) for x in out["mesh"][0].split())
    assert (nF, nI, nP) == (d["owner"].size, d["neighbour"].size, d["points"].shape[0])
    assert nC == max(d["owner"].max(), d["neighbour"].max()) + 1
    ints = lambda k: np.array(out[k][0].split(), dtype=np.int64)
    assert np.array_equal(ints("owner"), d["owner"]) and np.array_equal(ints("neighbour"), d["neighbour"])
    assert np.array_equal(ints("faceOffsets"), d["faceOffsets"]) and np.array_equal(ints("faceVerts"), d["faceVerts"])
    assert np.array_equal(np.array(out["points"][0].split(), dtype=np.float64), d["points"].ravel())
    patches = [p.split() for p in out["patch"]]
    assert [p[0] for p in patches] == [str(x) for x in d["patchNames"]]
    assert [p[1] for p in patches] == [str(x) for x in d["patchTypes"]]
    assert [int(p[2]) for p in patches] == list(d["patchStart"]) and [int(p[3]) for p in patches] == list(d["patchSize"])


def test_field_read_and_round_trip(out):
    assert out["T"][0].split()[:2] == ["internal", "20"]
    rows = {r.split()[1]: r.split()[2:] for r in out["T"][1:]}
    assert rows["left"] == ["fixedValue", "1", "1"] and rows["top"] == ["fixedValue", "1", "0.5"]
    assert rows["bottom"][:2] == ["zeroGradient", "0"] and rows["frontAndBack"][:2] == ["empty", "0"]
    assert out["roundtrip"] == ["1"]


def test_gamg_entries(out):
    # case20's own entry asks for directSolveCoarsest: refused by name
    assert "directSolveCoarsest" in out["case_refusal"][0]
    # pitzDaily's p entry: GaussSeidel, 1e-10, relTol 0, nCellsInCoarsestLevel 10, V(0,2,2), the defaults
    g = out["gamg"][0].split()
    assert g[0] == "GaussSeidel" and float(g[1]) == 1e-10 and float(g[2]) == 0.0
    assert [int(x) for x in g[3:]] == [1000, 0, 0, 1, 4, 2, 1, 4, 2, 1, 10]
    # relaxationFactors/equations: literal U 0.7, the ".*" regex 0.9 for k
    assert [float(x) for x in out["relax"][0].split()] == [0.7, 0.9]
    assert all(r.startswith("1 ") for r in out["refusal"]) and len(out["refusal"]) == 13
    assert out["failures"] == ["0"]
