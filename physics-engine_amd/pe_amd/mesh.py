"""Host-side shape preparation, as the HIP library applies it to a body's mesh.

`prepare_mesh(mesh, concave)` runs ConvexMeshShape::setMesh (convex_mesh_shape.cpp:12-132)
or ConcaveMeshShape::setMesh (concave_mesh_shape.cpp:6-17) through pe_mesh_prepare and
returns the prepared shape in FP32: vertices (moved to the centre of mass for a convex
mesh), faces in the BVH build's order, face normals, unique edges, volume, unit-mass
inertia and the relocation vector setMesh returns.  No GPU is needed.
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import check, lib
from .scenes import MeshDesc, _dptr, _iptr

_F32P = ctypes.POINTER(ctypes.c_float)
_I32P = ctypes.POINTER(ctypes.c_int32)


def prepare_mesh(mesh, concave: bool = False) -> dict:
    v, st, ix, fn = mesh.flat()
    md = MeshDesc(v.shape[0], _dptr(v), len(st) - 1, _iptr(st), _iptr(ix), _dptr(fn))
    cnt = np.zeros(4, np.int32)
    f = lambda a: a.ctypes.data_as(_F32P)
    i = lambda a: a.ctypes.data_as(_I32P)
    check(lib().pe_mesh_prepare(ctypes.byref(md), ctypes.c_int32(1 if concave else 0), i(cnt),
                                None, None, None, None, None, None, None))
    nv, nf, ni, ne = (int(x) for x in cnt)
    verts, fs, fi = np.zeros((nv, 3), np.float32), np.zeros(nf + 1, np.int32), np.zeros(max(ni, 1), np.int32)
    fnorm, edges = np.zeros((nf, 3), np.float32), np.zeros((max(ne, 1), 2, 3), np.float32)
    vi, reloc = np.zeros(10, np.float32), np.zeros(3, np.float32)
    check(lib().pe_mesh_prepare(ctypes.byref(md), ctypes.c_int32(1 if concave else 0), i(cnt), f(verts), i(fs),
                                i(fi), f(fnorm), f(edges), f(vi), f(reloc)))
    return dict(v=verts, faces=[fi[fs[k]:fs[k + 1]].tolist() for k in range(nf)], fn=fnorm, edges=edges[:ne],
                volume=float(vi[0]), inertia=vi[1:].reshape(3, 3), relocation=reloc)
