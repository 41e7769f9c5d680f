"""Helpers for the GPU parity tests: device tensors and a mesh + context pair."""
import numpy as np

from brae_amd import hexmesh, mfv


def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU tests need an MI355X"
    return torch


def dev(a, dtype=None):
    """a device copy, complete before return (the library's contexts run on their own streams)"""
    torch = torch_cuda()
    a = np.ascontiguousarray(a)
    t = torch.from_numpy(a.copy())
    if dtype is not None:
        t = t.to(dtype)
    t = t.to("cuda:0").contiguous()
    torch.cuda.synchronize()
    return t


def zeros(n, dtype="float64"):
    """zeros on the device, written before return"""
    torch = torch_cuda()
    t = torch.zeros(int(n), dtype=getattr(torch, dtype), device="cuda:0")
    torch.cuda.synchronize()
    return t


def host(t):
    torch_cuda().cuda.synchronize()
    return t.detach().cpu().numpy()


class Case:
    """Context + device mesh for a polyMesh given as numpy arrays."""

    def __init__(self, nCells, owner_internal, neighbour, bFaceCell, ctx=None):
        torch_cuda().cuda.synchronize()
        self.ctx = ctx or mfv.Context(0)
        self.mesh = mfv.Mesh(self.ctx, nCells, owner_internal, neighbour, bFaceCell)
        self.nCells = nCells
        self.nFaces = len(neighbour)
        self.nBFaces = len(bFaceCell)

    def call(self, name, *args):
        self.ctx.call(name, *args)

    def sync(self):
        self.ctx.sync()


def hex_case(nx, ny, nz, **kw):
    m = hexmesh.box_mesh(nx, ny, nz, **kw)
    nI = m["nInternal"]
    c = Case(m["nCells"], m["owner"][:nI], m["neighbour"], m["owner"][nI:])
    return m, c


_HIP = None


def d2h(dptr, n, dtype):
    """copy n elements of a raw device pointer the library returned to a new host array"""
    global _HIP
    import ctypes
    torch_cuda().cuda.synchronize()
    if _HIP is None:
        _HIP = ctypes.CDLL("libamdhip64.so")
        _HIP.hipMemcpy.restype = ctypes.c_int
        _HIP.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _HIP.hipDeviceSynchronize.restype = ctypes.c_int
    out = np.zeros(int(n), dtype)
    if n:
        assert _HIP.hipDeviceSynchronize() == 0
        assert _HIP.hipMemcpy(out.ctypes.data, dptr, out.nbytes, 2) == 0   # hipMemcpyDeviceToHost
    return out
