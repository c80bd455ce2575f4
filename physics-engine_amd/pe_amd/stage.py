"""The stage entry points (include/pe_amd.h, "stage entry points"): one stage of the reference's
World on device buffers the caller owns -- here torch tensors on the GPU, whose memory and stream
are only the plumbing; the stages themselves are the library's HIP kernels.

    st = Stage(max_bodies, max_pairs)
    pa, pb, res = st.broadphase(aabb, axis)                  # BroadPhaseSweepAndPrune
    pts, npts, swap, res = st.narrowphase(body, pa, pb, prev_npairs, pts)   # SimpleNarrowPhase
    res = st.solve(body, solver, pa, pb, npts, pts, dt, iters)             # SequentialImpulseSolver

Each call is enqueued on torch's current stream and returns without waiting.
"""
from __future__ import annotations

import ctypes

from ._lib import check, lib


def _p(t):
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def _stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class Stage:
    def __init__(self, max_bodies: int, max_pairs: int):
        h = ctypes.c_void_p()
        check(lib().pe_stage_create(ctypes.c_int32(max_bodies), ctypes.c_int32(max_pairs), ctypes.byref(h)))
        self.h = h
        self.max_bodies, self.max_pairs = max_bodies, max_pairs

    def close(self):
        if getattr(self, "h", None):
            lib().pe_stage_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_meshes(self, meshes, concave, prepared: bool = False):
        """meshes: pe_amd.scenes.Mesh list; concave: one flag per mesh."""
        from .scenes import MeshDesc, _dptr, _iptr
        import numpy as np
        flats = [m.flat() for m in meshes]
        arr = (MeshDesc * max(1, len(flats)))()
        for k, (v, st, ix, fn) in enumerate(flats):
            arr[k] = MeshDesc(v.shape[0], _dptr(v), len(st) - 1, _iptr(st), _iptr(ix), _dptr(fn))
        cc = np.ascontiguousarray(concave, dtype=np.int32)
        check(lib().pe_stage_set_meshes(self.h, ctypes.c_int32(len(flats)), arr, _iptr(cc) if len(cc) else None,
                                        ctypes.c_int32(1 if prepared else 0)))

    def broadphase(self, aabb, axis: int, cap: int = 0, pair_a=None, pair_b=None, result=None):
        import torch
        n = aabb.shape[0]
        cap = cap or self.max_pairs
        dev = aabb.device
        pair_a = torch.empty(cap, dtype=torch.int32, device=dev) if pair_a is None else pair_a
        pair_b = torch.empty(cap, dtype=torch.int32, device=dev) if pair_b is None else pair_b
        result = torch.zeros(3, dtype=torch.int32, device=dev) if result is None else result
        check(lib().pe_stage_broadphase(self.h, ctypes.c_int32(n), _p(aabb), ctypes.c_int32(axis), _p(pair_a),
                                        _p(pair_b), ctypes.c_int32(cap), _p(result), _stream()))
        return pair_a, pair_b, result

    def narrowphase(self, body, pair_a, pair_b, npairs: int, prev_npairs: int, points):
        import torch
        dev = body.device
        npts = torch.zeros(max(npairs, 1), dtype=torch.int32, device=dev)
        swap = torch.zeros(max(npairs, 1), dtype=torch.int32, device=dev)
        result = torch.zeros(1, dtype=torch.int32, device=dev)
        check(lib().pe_stage_narrowphase(self.h, ctypes.c_int32(body.shape[0]), _p(body), ctypes.c_int32(npairs),
                                         _p(pair_a), _p(pair_b), ctypes.c_int32(prev_npairs), _p(points), _p(npts),
                                         _p(swap), _p(result), _stream()))
        return npts[:npairs], swap[:npairs], result

    def solve(self, body, solver, pair_a, pair_b, npairs: int, npts, points, dt: float, iterations: int,
              impulses=None):
        import torch
        result = torch.zeros(2, dtype=torch.int32, device=body.device)
        check(lib().pe_stage_solve(self.h, ctypes.c_int32(body.shape[0]), _p(body), _p(solver), ctypes.c_int32(npairs),
                                   _p(pair_a), _p(pair_b), _p(npts), _p(points), ctypes.c_float(dt),
                                   ctypes.c_int32(iterations), _p(impulses), _p(result), _stream()))
        return result
