"""The C++ host path: tests/cpp/shim_pressure.cpp calls the library through include/brae_mfv.hpp with
brae's own signatures (assembly, fold, reference cell, buildAMG, amgGalerkin, deviceAMGPCG, deviceAmul,
deviceNormFactorInto, flux) and checks the solve, the recomputed residual and the refusals itself."""
import json
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "shim_pressure.bin")


@pytest.mark.parametrize("n", [16, 40])
def test_cpp_caller_solves_pressure(n):
    assert os.path.exists(EXE), "build the C++ callers first (__graft_entry__.build())"
    print("case: shim_pressure", n, flush=True)
    r = subprocess.run([EXE, str(n)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["ok"] and out["cells"] == n ** 3 and out["final"] < 1e-8
    assert out["cache_clone_same_bits"] and out["fvsolution_gamg_same_bits"]
