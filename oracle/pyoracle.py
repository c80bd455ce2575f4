"""CPU ORACLE -- test infrastructure only.

ctypes bindings for liboracle.so (the C++ restatement, FP32 `peo32_*` and FP64 `peo64_*`)
and for oracle/_ref/libperef{32,64}.so (the reference engine compiled from its own
sources, present only where the reference checkout was available at build time).
Only tests/, __graft_entry__.smoke() and bench.py's CPU baseline use this module.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_VP = ctypes.c_void_p
_I32P = ctypes.POINTER(ctypes.c_int32)
_F64P = ctypes.POINTER(ctypes.c_double)


def _d(a):
    return a.ctypes.data_as(_F64P)


def _i(a):
    return a.ctypes.data_as(_I32P)


_libs = {}


def _load(path):
    if path not in _libs:
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built (run __graft_entry__.build())")
        _libs[path] = ctypes.CDLL(path)
    return _libs[path]


def ref_available(bits: int = 32) -> bool:
    return os.path.exists(os.path.join(HERE, "_ref", f"libperef{bits}.so"))


def ref_amd_available() -> bool:
    """The reference's World with the library's stages swapped in (oracle/_ref/libperef32_amd.so)."""
    return os.path.exists(os.path.join(HERE, "_ref", "libperef32_amd.so"))


class _Base:
    prefix = ""

    def __init__(self, lib, scene):
        self.lib = lib
        self.scene = scene
        self.n = scene.n
        f = self._f
        f("world_create").restype = _VP
        f("world_create").argtypes = [_VP, _VP]
        self._desc = scene.desc()
        self._params = scene.params()
        self.h = f("world_create")(ctypes.byref(self._desc), ctypes.byref(self._params))
        if not self.h:
            raise ValueError("world_create failed (unsupported shape?)")
        for a, b in getattr(scene, "ignore_pairs", None) or []:
            f("add_ignore")(ctypes.c_void_p(self.h), ctypes.c_int(a), ctypes.c_int(b))
        for j in getattr(scene, "joints", None) or []:
            f("add_joint")(ctypes.c_void_p(self.h), ctypes.byref(j))
        if getattr(scene, "sleep", None):
            f("set_sleep")(ctypes.c_void_p(self.h), *[ctypes.c_double(x) for x in scene.sleep])
        for name in ("broadphase", "num_pairs", "get_axis", "num_results"):
            f(name).restype = ctypes.c_int
            f(name).argtypes = [_VP]

    def _f(self, name):
        return getattr(self.lib, f"{self.prefix}_{name}")

    def step(self, n=1):
        self._f("step")(ctypes.c_void_p(self.h), ctypes.c_int(n))

    def integrate(self):
        self._f("integrate")(ctypes.c_void_p(self.h))

    def broadphase(self):
        return self._f("broadphase")(ctypes.c_void_p(self.h))

    def narrowphase(self):
        self._f("narrowphase")(ctypes.c_void_p(self.h))

    def solve(self):
        self._f("solve")(ctypes.c_void_p(self.h))

    def axis(self):
        return self._f("get_axis")(ctypes.c_void_p(self.h))

    def set_axis(self, a):
        self._f("set_axis")(ctypes.c_void_p(self.h), ctypes.c_int(a))

    def state(self):
        n = self.n
        pos, rot, v, w = np.zeros((n, 3)), np.zeros((n, 3, 3)), np.zeros((n, 3)), np.zeros((n, 3))
        self._f("get_state")(ctypes.c_void_p(self.h), _d(pos), _d(rot), _d(v), _d(w))
        return pos, rot, v, w

    def set_state(self, pos=None, rot=None, v=None, w=None):
        c = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64)
        pos, rot, v, w = c(pos), c(rot), c(v), c(w)
        p = lambda a: None if a is None else _d(a)
        self._f("set_state")(ctypes.c_void_p(self.h), p(pos), p(rot), p(v), p(w))

    def aabbs(self):
        mn, mx = np.zeros((self.n, 3)), np.zeros((self.n, 3))
        self._f("get_aabbs")(ctypes.c_void_p(self.h), _d(mn), _d(mx))
        return mn, mx

    def world_inv_inertia(self):
        out = np.zeros((self.n, 3, 3))
        self._f("get_world_inv_inertia")(ctypes.c_void_p(self.h), _d(out))
        return out

    def mass_props(self):
        im, li = np.zeros(self.n), np.zeros((self.n, 3, 3))
        self._f("get_mass_props")(ctypes.c_void_p(self.h), _d(im), _d(li))
        return im, li

    def pairs(self):
        n = self._f("num_pairs")(ctypes.c_void_p(self.h))
        ab = np.zeros((n, 2), np.int32)
        if n:
            self._f("get_pairs")(ctypes.c_void_p(self.h), _i(ab))
        return ab

    def set_pairs(self, ab):
        ab = np.ascontiguousarray(ab, dtype=np.int32)
        self._f("set_pairs")(ctypes.c_void_p(self.h), ctypes.c_int(len(ab)), _i(ab))

    def status(self):
        """Per body: (status bits: 1 asleep, 2 removed; static and dynamic neighbour counts of the last
        narrowphase; sleep time)."""
        st, sc, dc, t = (np.zeros(self.n, np.int32), np.zeros(self.n, np.int32), np.zeros(self.n, np.int32),
                         np.zeros(self.n))
        self._f("get_status")(ctypes.c_void_p(self.h), _i(st), _i(sc), _i(dc), _d(t))
        return st, sc, dc, t

    def raycast(self, start, direction, length, ignore=()):
        """DefaultRaycastSolver::performRaycast: (hit, body (-1: none), distance, point [3], normal [3]) as
        the RaycastResult is left."""
        f = self._f("raycast")
        f.restype = ctypes.c_int
        ray = np.ascontiguousarray(list(start) + list(direction) + [length], np.float64)
        ig = np.ascontiguousarray(sorted(set(int(x) for x in ignore)) or [0], np.int32)
        out = np.zeros(8)
        hit = f(ctypes.c_void_p(self.h), _d(ray), ctypes.c_int(len(set(ignore))), _i(ig), _d(out))
        return bool(hit), int(out[0]), out[1], out[2:5].copy(), out[5:8].copy()

    def set_report(self, body, on=True):
        """RigidBody::addCollisionCallback (on) / clearCollisionCallbacks (off)."""
        self._f("set_report")(ctypes.c_void_p(self.h), ctypes.c_int(body), ctypes.c_int(1 if on else 0))

    def remove_body(self, body):
        """World::removeRigidBody."""
        self._f("remove_body")(ctypes.c_void_p(self.h), ctypes.c_int(body))

    def events(self):
        """The last narrowphase's collision reports: (bodies [E, 2] (A, B), data [E, 10]: mean position,
        normalised normal sum, mean depth, mean relative velocity)."""
        f = self._f("num_events")
        f.restype = ctypes.c_int
        n = f(ctypes.c_void_p(self.h))
        ab, data = np.zeros((n, 2), np.int32), np.zeros((n, 10))
        if n:
            self._f("get_events")(ctypes.c_void_p(self.h), _i(ab), _d(data))
        return ab, data

    def contacts(self):
        """(npts [P], swap [P], points [P, 8, 13]) for every contact result slot."""
        n = self._f("num_results")(ctypes.c_void_p(self.h))
        npts, swap = np.zeros(n, np.int32), np.zeros(n, np.int32)
        pts = np.zeros((n, 8, 13))
        if n:
            self._f("get_contacts")(ctypes.c_void_p(self.h), _i(npts), _i(swap), _d(pts))
        return npts, swap, pts


    def poly(self, body):
        """Polyhedral data of a body's shape after set-up, or None for a sphere:
        dict(v [nv,3], faces [list], fn [nf,3], edges [ne,2,3], volume, inertia [3,3])."""
        f = self._f("get_poly")
        f.restype = ctypes.c_int
        cnt = np.zeros(4, np.int32)
        if f(ctypes.c_void_p(self.h), ctypes.c_int(body), _i(cnt), None, None, None, None, None, None) != 0:
            return None
        nv, nf, ni, ne = (int(x) for x in cnt)
        v, st, ix = np.zeros((nv, 3)), np.zeros(nf + 1, np.int32), np.zeros(max(ni, 1), np.int32)
        fn, e, vi = np.zeros((nf, 3)), np.zeros((max(ne, 1), 2, 3)), np.zeros(10)
        f(ctypes.c_void_p(self.h), ctypes.c_int(body), _i(cnt), _d(v), _i(st), _i(ix), _d(fn), _d(e), _d(vi))
        faces = [ix[st[k]:st[k + 1]].tolist() for k in range(nf)]
        return dict(v=v, faces=faces, fn=fn, edges=e[:ne], volume=vi[0], inertia=vi[1:].reshape(3, 3))


class OracleWorld(_Base):
    """The C++ restatement (oracle/pe_oracle.hpp) in FP32 (bits=32) or FP64 (bits=64)."""

    def __init__(self, scene, bits: int = 32, color_order: bool = True):
        self.prefix = f"peo{bits}"
        lib = _load(os.path.join(HERE, "liboracle.so"))
        super().__init__(lib, scene)
        for name in ("num_colors", "get_solve_order"):
            self._f(name).restype = ctypes.c_int
        self._f("get_solve_order").argtypes = [_VP, _I32P]
        self._f("num_colors").argtypes = [_VP]
        self.set_color_order(color_order)

    def set_color_order(self, on: bool):
        self._f("set_color_order")(ctypes.c_void_p(self.h), ctypes.c_int(1 if on else 0))

    def solve_order(self):
        n = self._f("num_results")(ctypes.c_void_p(self.h))
        out = np.zeros(max(n, 1), np.int32)
        k = self._f("get_solve_order")(ctypes.c_void_p(self.h), _i(out))
        return out[:k].copy()

    def color_of(self):
        n = self._f("num_results")(ctypes.c_void_p(self.h))
        out = np.zeros(max(n, 1), np.int32)
        self._f("get_color_of")(ctypes.c_void_p(self.h), _i(out))
        return out[:n].copy()

    def num_colors(self):
        return self._f("num_colors")(ctypes.c_void_p(self.h))

    def joint_order(self):
        out = np.zeros(max(1, len(getattr(self.scene, "joints", None) or [])), np.int32)
        k = self._f("get_joint_order")(ctypes.c_void_p(self.h), _i(out))
        return out[:k].copy()

    def impulses(self):
        n = self._f("num_results")(ctypes.c_void_p(self.h))
        out = np.zeros((n, 8, 3))
        if n:
            self._f("get_impulses")(ctypes.c_void_p(self.h), _d(out))
        return out

    def __del__(self):
        try:
            self._f("world_destroy")(ctypes.c_void_p(self.h))
        except Exception:
            pass


class RefWorld(_Base):
    """The reference engine itself (oracle/_ref/libperef{bits}.so), stage by stage.  With
    amd_stages (bit 0: broadphase, 1: narrowphase, 2: solver) the World's stages are replaced by
    the library's through the product's reference-side adapter (FP32 build with the adapter,
    oracle/_ref/libperef32_amd.so); it then loads libpe_amd.so."""

    def __init__(self, scene, bits: int = 32, threads: int = 0, amd_stages: int = 0):
        self.prefix = "pref"
        if amd_stages:
            assert bits == 32
            import pe_amd  # noqa: F401  (the library, loaded from its package path first)
            pe_amd.lib()
        lib = _load(os.path.join(HERE, "_ref", "libperef32_amd.so" if amd_stages else f"libperef{bits}.so"))
        lib.pref_init_threads(ctypes.c_int(threads))
        if threads == 1:
            lib.pref_set_threads(ctypes.c_int(1))
        lib.pref_step_timed.restype = ctypes.c_double
        lib.pref_step_timed.argtypes = [_VP, ctypes.c_int]
        super().__init__(lib, scene)
        if amd_stages:
            lib.pref_use_amd_stages.restype = ctypes.c_int
            lib.pref_step_checked.restype = ctypes.c_int
            if lib.pref_use_amd_stages(ctypes.c_void_p(self.h), ctypes.c_int(amd_stages),
                                       ctypes.c_int(scene.iterations)) != 0:
                raise RuntimeError("the adapter could not replace the stages")

    def callbacks(self):
        """The collision callback invocations of the last narrowphase: (ids [C, 2] (self, other),
        data [C, 9] (position, normal, velocity as the callback received them)), in the order the
        reference's threads made them."""
        n = self.lib.pref_num_callbacks(ctypes.c_void_p(self.h))
        ids, data = np.zeros((n, 2), np.int32), np.zeros((n, 9))
        if n:
            self.lib.pref_get_callbacks(ctypes.c_void_p(self.h), _i(ids), _d(data))
        return ids, data

    def step_checked(self, n=1):
        """World::step through the installed stages; raises when an adapter stage throws."""
        if self.lib.pref_step_checked(ctypes.c_void_p(self.h), ctypes.c_int(n)) != 0:
            raise RuntimeError("World::step failed in a replaced stage")

    def _checked(self, r):
        if self.lib.pref_failed():
            raise RuntimeError("a stage of the reference's World threw (message on stderr)")
        return r

    def integrate(self):
        return self._checked(super().integrate())

    def broadphase(self):
        return self._checked(super().broadphase())

    def narrowphase(self):
        return self._checked(super().narrowphase())

    def solve(self):
        return self._checked(super().solve())

    def step_timed(self, n=1) -> float:
        return self.lib.pref_step_timed(ctypes.c_void_p(self.h), n)
