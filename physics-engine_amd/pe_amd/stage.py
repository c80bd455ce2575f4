"""The stage entry points (include/pe_amd.h, "stage entry points"): one stage of the reference's
World on device buffers the caller owns (pe_amd.hip.DeviceArray here; any device pointer works),
enqueued on the caller's stream (pe_amd.hip.Stream, or None for the null stream) without waiting.

    st = Stage(max_bodies, max_pairs)
    pa, pb, res = st.broadphase(aabb, axis)                                   # BroadPhaseSweepAndPrune
    npts, swap, res = st.narrowphase(body, pa, pb, npairs, prev_npairs, pts)  # SimpleNarrowPhase
    res = st.solve(body, solver, pa, pb, npairs, npts, pts, dt, iterations)   # SequentialImpulseSolver
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import check, lib
from .hip import DeviceArray


def _p(t):
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def _s(stream):
    if stream is None:
        return ctypes.c_void_p(0)
    return stream.handle if hasattr(stream, "handle") else ctypes.c_void_p(int(stream))


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
        flats = [m.flat() for m in meshes]
        arr = (MeshDesc * max(1, len(flats)))()
        for k, (v, st, ix, fn) in enumerate(flats):
            arr[k] = MeshDesc(v.shape[0], _dptr(v), len(st) - 1, _iptr(st), _iptr(ix), _dptr(fn))
        cc = np.ascontiguousarray(concave, dtype=np.int32)
        self._keep = (flats, arr, cc)
        check(lib().pe_stage_set_meshes(self.h, ctypes.c_int32(len(flats)), arr, _iptr(cc) if len(cc) else None,
                                        ctypes.c_int32(1 if prepared else 0)))

    def broadphase(self, aabb, axis: int, cap: int = 0, pair_a=None, pair_b=None, result=None, stream=None):
        n = aabb.shape[0]
        cap = cap or self.max_pairs
        pair_a = DeviceArray(cap, np.int32) if pair_a is None else pair_a
        pair_b = DeviceArray(cap, np.int32) if pair_b is None else pair_b
        result = DeviceArray.from_numpy(np.zeros(3, np.int32)) if result is None else result
        check(lib().pe_stage_broadphase(self.h, ctypes.c_int32(n), _p(aabb), ctypes.c_int32(axis), _p(pair_a),
                                        _p(pair_b), ctypes.c_int32(cap), _p(result), _s(stream)))
        return pair_a, pair_b, result

    def narrowphase(self, body, pair_a, pair_b, npairs: int, prev_npairs: int, points, stream=None):
        npts = DeviceArray(max(npairs, 1), np.int32)
        swap = DeviceArray(max(npairs, 1), np.int32)
        result = DeviceArray.from_numpy(np.zeros(1, np.int32))
        check(lib().pe_stage_narrowphase(self.h, ctypes.c_int32(body.shape[0]), _p(body), ctypes.c_int32(npairs),
                                         _p(pair_a), _p(pair_b), ctypes.c_int32(prev_npairs), _p(points), _p(npts),
                                         _p(swap), _p(result), _s(stream)))
        return npts, swap, result

    def solve(self, body, solver, pair_a, pair_b, npairs: int, npts, points, dt: float, iterations: int,
              impulses=None, stream=None):
        result = DeviceArray.from_numpy(np.zeros(2, np.int32))
        check(lib().pe_stage_solve(self.h, ctypes.c_int32(body.shape[0]), _p(body), _p(solver), ctypes.c_int32(npairs),
                                   _p(pair_a), _p(pair_b), _p(npts), _p(points), ctypes.c_float(dt),
                                   ctypes.c_int32(iterations), _p(impulses), _p(result), _s(stream)))
        return result
