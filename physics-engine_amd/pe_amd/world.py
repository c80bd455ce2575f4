"""Python mirror of the reference's World for the HIP path.

`World` follows pe_intf::World (src/intf/world.h): it is built from a Scene (the bodies
World::addRigidBody would receive), `step()` runs World::step (world.cpp:267-300) on the GPU,
and the four stage methods run the pipeline one stage at a time:

    integrate()    -> World::updateObjectStatus + applyExternalForce
    broadphase()   -> BroadPhaseSweepAndPrune::calcCollisionPairs
    narrowphase()  -> SimpleNarrowPhase::calcContactResults
    solve()        -> SequentialImpulseSolver::setupSolver + solve

All state stays on the device; the getters copy it back.
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import check, lib

_I32P = ctypes.POINTER(ctypes.c_int32)
_F32P = ctypes.POINTER(ctypes.c_float)
_F64P = ctypes.POINTER(ctypes.c_double)
_I64P = ctypes.POINTER(ctypes.c_int64)


def _f(a):
    return a.ctypes.data_as(_F32P)


def _i(a):
    return a.ctypes.data_as(_I32P)


class World:
    def __init__(self, scene, max_pairs: int = 0):
        self.scene = scene
        self.n = scene.n
        self._desc = scene.desc()
        self._params = scene.params()
        h = ctypes.c_void_p()
        check(lib().pe_world_create(ctypes.byref(self._desc), ctypes.byref(self._params), ctypes.c_int32(max_pairs),
                                    ctypes.byref(h)))
        self.h = h
        if scene.ignore_pairs:
            self.add_ignore_pairs(scene.ignore_pairs)
        if scene.joints:
            self.add_joints(scene.joints)
        if getattr(scene, "sleep", None):
            self.set_sleep(*scene.sleep)
        self._callbacks = {}   # body -> [callable]

    def add_joints(self, joints):
        """World::addConstraint for each JointDesc, in order (pe_world_add_joints)."""
        from .scenes import JointDesc
        arr = (JointDesc * len(joints))(*joints)
        check(lib().pe_world_add_joints(self.h, ctypes.c_int32(len(joints)), arr))

    def add_ignore_pairs(self, pairs):
        """RigidBody::addIgnoreCollisionId on both bodies of each (a, b) (pe_world_add_ignore_pairs)."""
        p = np.ascontiguousarray(np.asarray(pairs, np.int32).reshape(-1, 2))
        check(lib().pe_world_add_ignore_pairs(self.h, ctypes.c_int32(p.shape[0]), _i(p)))

    def close(self):
        if getattr(self, "h", None):
            lib().pe_world_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- World::step and its stages
    def step(self, n: int = 1):
        """World::step n times.  With collision callbacks registered the steps run one at a time and
        each step's callbacks are called after it (World::execCollisionCallbacks, world.cpp:129-192)."""
        if not self._callbacks:
            check(lib().pe_world_step(self.h, ctypes.c_int32(n)))
            return
        for _ in range(n):
            check(lib().pe_world_step(self.h, ctypes.c_int32(1)))
            self.dispatch_callbacks()

    # ---- raycast
    def raycast(self, starts, directions, lengths, ignore=()):
        """DefaultRaycastSolver::performRaycast for each ray (pe_world_raycast): starts, directions [R, 3],
        lengths [R]; ignore: body indices skipped by every ray.  Returns (hit [R] bool, body [R] int32,
        distance [R], point [R, 3], normal [R, 3]) as each RaycastResult is left."""
        from .scenes import RayHit
        st = np.asarray(starts, np.float32).reshape(-1, 3)
        rays = np.ascontiguousarray(np.concatenate([st, np.asarray(directions, np.float32).reshape(-1, 3),
                                                    np.asarray(lengths, np.float32).reshape(-1, 1)], axis=1))
        ig = np.ascontiguousarray(np.asarray(list(ignore) or [0], np.int32))
        out = (RayHit * max(len(rays), 1))()
        check(lib().pe_world_raycast(self.h, ctypes.c_int32(len(rays)), _f(rays), ctypes.c_int32(len(ignore)), _i(ig), out))
        raw = np.frombuffer(out, dtype=np.float32, count=9 * len(rays)).reshape(-1, 9)
        ints = raw[:, :2].copy().view(np.int32)
        return ints[:, 0] != 0, ints[:, 1].copy(), raw[:, 2].copy(), raw[:, 3:6].copy(), raw[:, 6:9].copy()

    # ---- sleeping, lifetimes, removal, collision callbacks
    def set_sleep(self, lin_vel2: float, ang_vel2: float, time: float):
        """World::setSleepLinVel2Threshold / setSleepAngVel2Threshold / setSleepTimeThreshold."""
        check(lib().pe_world_set_sleep(self.h, ctypes.c_double(lin_vel2), ctypes.c_double(ang_vel2),
                                       ctypes.c_double(time)))

    def remove_body(self, body: int):
        """World::removeRigidBody."""
        check(lib().pe_world_remove_body(self.h, ctypes.c_int32(body)))

    def body_status(self):
        """(flags [n]: 1 asleep, 2 removed; static and dynamic neighbour counts [n] of the last
        narrowphase; sleep times [n])."""
        fl, sc, dc = (np.zeros(self.n, np.int32) for _ in range(3))
        t = np.zeros(self.n, np.float32)
        check(lib().pe_world_get_body_status(self.h, _i(fl), _i(sc), _i(dc), _f(t)))
        return fl, sc, dc, t

    def set_collision_reports(self, bodies, on: bool = True):
        b = np.ascontiguousarray(np.asarray(bodies, np.int32).ravel())
        check(lib().pe_world_set_collision_reports(self.h, ctypes.c_int32(b.size), _i(b), ctypes.c_int32(1 if on else 0)))

    def collision_events(self):
        """The last narrowphase's reports: (bodies [E, 2] (A, B), data [E, 10] float32: position,
        normal, depth, velocity)."""
        from .scenes import CollisionEvent
        cnt = ctypes.c_int32()
        check(lib().pe_world_get_collision_events(self.h, ctypes.c_int32(0), None, ctypes.byref(cnt)))
        arr = (CollisionEvent * max(cnt.value, 1))()
        check(lib().pe_world_get_collision_events(self.h, ctypes.c_int32(cnt.value), arr, ctypes.byref(cnt)))
        raw = np.frombuffer(arr, dtype=np.float32, count=12 * cnt.value).reshape(-1, 12)
        return raw[:, :2].copy().view(np.int32), raw[:, 2:].copy()

    def add_collision_callback(self, body: int, fn):
        """RigidBody::addCollisionCallback: fn(self_body, other_body, position, normal, velocity) is
        called after each step for each contact result of the body with points."""
        self._callbacks.setdefault(int(body), []).append(fn)
        self.set_collision_reports([body], True)

    def clear_collision_callbacks(self, body: int):
        self._callbacks.pop(int(body), None)
        self.set_collision_reports([body], False)

    def dispatch_callbacks(self):
        """Calls the registered callbacks with the last step's reports, as World::execCollisionCallbacks
        calls them: A's with (A, B, position + normal * depth, -normal, -velocity), B's with
        (B, A, position, normal, velocity) (FP32)."""
        ab, ev = self.collision_events()
        for (a, b), e in zip(ab, ev):
            pos, nor, dep, vel = e[0:3], e[3:6], e[6], e[7:10]
            for fn in self._callbacks.get(int(a), ()):
                fn(int(a), int(b), pos + nor * dep, -nor, -vel)
            for fn in self._callbacks.get(int(b), ()):
                fn(int(b), int(a), pos, nor, vel)

    def step_async(self, n: int, stream):
        """n steps enqueued on `stream` (a hipStream_t handle as an int, e.g.
        pe_amd.hip.Stream().handle.value; 0 is the null stream); ordered after the world's pending
        work, and its later calls wait for them."""
        check(lib().pe_world_step_async(self.h, ctypes.c_int32(n), ctypes.c_void_p(stream)))

    def cache(self):
        """(prev_npairs, distances [prev_npairs, 8]): the contact cache the next narrowphase starts from."""
        k = ctypes.c_int32()
        check(lib().pe_world_get_cache(self.h, None, 0, ctypes.byref(k)))
        d = np.zeros((max(k.value, 1), 8), np.float32)
        check(lib().pe_world_get_cache(self.h, _f(d), ctypes.c_int32(k.value), ctypes.byref(k)))
        return k.value, d[:k.value].copy()

    def set_cache(self, distances):
        d = np.ascontiguousarray(distances, dtype=np.float32).reshape(-1, 8)
        check(lib().pe_world_set_cache(self.h, _f(d) if len(d) else None, ctypes.c_int32(len(d))))

    def integrate(self):
        check(lib().pe_world_integrate(self.h))

    def broadphase(self) -> int:
        n = ctypes.c_int32()
        check(lib().pe_world_broadphase(self.h, ctypes.byref(n)))
        return n.value

    def narrowphase(self):
        check(lib().pe_world_narrowphase(self.h))

    def solve(self):
        check(lib().pe_world_solve(self.h))

    def sync(self):
        check(lib().pe_world_sync(self.h))

    # ---- state
    def state(self):
        n = self.n
        pos, rot = np.zeros((n, 3), np.float32), np.zeros((n, 3, 3), np.float32)
        v, w = np.zeros((n, 3), np.float32), np.zeros((n, 3), np.float32)
        check(lib().pe_world_get_state(self.h, _f(pos), _f(rot), _f(v), _f(w)))
        return pos, rot, v, w

    def set_state(self, pos=None, rot=None, v=None, w=None):
        c = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float32)
        pos, rot, v, w = c(pos), c(rot), c(v), c(w)
        p = lambda a: None if a is None else _f(a)
        check(lib().pe_world_set_state(self.h, p(pos), p(rot), p(v), p(w)))

    def aabbs(self):
        mn, mx = np.zeros((self.n, 3), np.float32), np.zeros((self.n, 3), np.float32)
        check(lib().pe_world_get_aabbs(self.h, _f(mn), _f(mx)))
        return mn, mx

    def pairs(self):
        cnt = ctypes.c_int32()
        check(lib().pe_world_get_pairs(self.h, None, 0, ctypes.byref(cnt)))
        ab = np.zeros((max(cnt.value, 1), 2), np.int32)
        check(lib().pe_world_get_pairs(self.h, _i(ab), ctypes.c_int32(cnt.value), ctypes.byref(cnt)))
        return ab[:cnt.value].copy()

    def contacts(self):
        """(npts [P], points [P, 8, 13]) for every pair of the last narrowphase."""
        cnt = ctypes.c_int32()
        check(lib().pe_world_get_pairs(self.h, None, 0, ctypes.byref(cnt)))
        P = cnt.value
        npts = np.zeros(max(P, 1), np.int32)
        pts = np.zeros((max(P, 1), 8, 13), np.float32)
        check(lib().pe_world_get_contacts(self.h, _i(npts), _f(pts), ctypes.c_int32(P)))
        return npts[:P].copy(), pts[:P].copy()

    def impulses(self):
        """[P, 8, 3] accumulated (normal, tangent 0, tangent 1) impulses of the last solve."""
        cnt = ctypes.c_int32()
        check(lib().pe_world_get_pairs(self.h, None, 0, ctypes.byref(cnt)))
        P = cnt.value
        out = np.zeros((max(P, 1), 8, 3), np.float32)
        check(lib().pe_world_get_impulses(self.h, _f(out), ctypes.c_int32(P)))
        return out[:P].copy()

    def solve_order(self):
        """(order, color_start): pair indices grouped by colour."""
        no, nc = ctypes.c_int32(), ctypes.c_int32()
        check(lib().pe_world_get_solve_order(self.h, None, 0, ctypes.byref(no), None, 0, ctypes.byref(nc)))
        order = np.zeros(max(no.value, 1), np.int32)
        cs = np.zeros(nc.value + 1, np.int32)
        check(lib().pe_world_get_solve_order(self.h, _i(order), ctypes.c_int32(no.value), ctypes.byref(no), _i(cs),
                                             ctypes.c_int32(nc.value + 1), ctypes.byref(nc)))
        return order[:no.value].copy(), cs

    def axis(self) -> int:
        a = ctypes.c_int32()
        check(lib().pe_world_get_axis(self.h, ctypes.byref(a)))
        return a.value

    def set_axis(self, a: int):
        check(lib().pe_world_set_axis(self.h, ctypes.c_int32(a)))

    def set_timing(self, on: bool):
        check(lib().pe_world_set_timing(self.h, ctypes.c_int32(1 if on else 0)))

    def stage_times(self):
        ms = (ctypes.c_double * 4)()
        check(lib().pe_world_get_stage_times(self.h, ms))
        return list(ms)

    def set_solver_blocks(self, blocks: int):
        check(lib().pe_world_set_solver_blocks(self.h, ctypes.c_int32(blocks)))

    def set_colouring(self, mode: int):
        """0: by body count, 1: one workgroup's claims (at most 12 288 bodies), 2: rounds over the GPU;
        the colour classes are the same."""
        check(lib().pe_world_set_colouring(self.h, ctypes.c_int32(mode)))

    def records(self):
        """(AABB records [n, 8], collision records [n, 20], solver records [n, 24]) as the last
        integrate stage wrote them: the inputs of the stage entry points (pe_amd.stage)."""
        a = np.zeros((self.n, 8), np.float32)
        b = np.zeros((self.n, 20), np.float32)
        c = np.zeros((self.n, 24), np.float32)
        check(lib().pe_world_get_records(self.h, _f(a), _f(b), _f(c)))
        return a, b, c

    def counts(self):
        c = (ctypes.c_int64 * 6)()
        check(lib().pe_world_get_counts(self.h, c))
        return dict(zip(("bodies", "dynamic", "pairs", "points", "colors", "colour_rounds"), list(c)))
