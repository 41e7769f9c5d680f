"""The binding's object lifetimes: a context closed before its meshes, or a reference cycle finalised by the
garbage collector in any order, releases what was built on a context or a mesh before the context or the mesh
itself (the library needs them to outlive their hierarchies and solvers)."""
import gc

import numpy as np
import pytest

import gpu_util as G
from brae_amd import hexmesh, mfv

pytestmark = pytest.mark.gpu


def build(ctx):
    m = hexmesh.box_mesh(12, 8, 6, (1.0, 1.0, 1.0), (1.0, 1.0, 1.0), 0.0, seed=0)
    nI = m["nInternal"]
    mesh = mfv.Mesh(ctx, m["nCells"], m["owner"][:nI], m["neighbour"], m["owner"][nI:])
    amg = mfv.Amg(mesh, np.asarray(m["magSf"][:nI]), mfv.AGG_BRAE, 16)
    return mesh, amg


def test_context_closed_first():
    G.torch_cuda()
    print("case: context closed before its mesh and hierarchy", flush=True)
    ctx = mfv.Context(0)
    mesh, amg = build(ctx)
    ctx.close()
    assert not mesh.h and not amg.h
    mesh.close()
    amg.close()


def test_cycle_collected_in_any_order():
    G.torch_cuda()
    print("case: reference cycle of context, mesh and hierarchy", flush=True)
    for _ in range(3):
        ctx = mfv.Context(0)
        mesh, amg = build(ctx)
        holder = {"ctx": ctx, "mesh": mesh, "amg": amg}
        holder["self"] = holder   # a cycle: only the collector frees it
        ctx.cycle = holder
        del ctx, mesh, amg, holder
        gc.collect()
    # the library still works afterwards
    ctx = mfv.Context(0)
    mesh, amg = build(ctx)
    assert amg.nGrids >= 2
    ctx.close()
