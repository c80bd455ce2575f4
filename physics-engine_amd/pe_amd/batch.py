"""Batched stepping of many independent worlds (include/pe_amd_batch.h).

    b = Batch([scene0, scene1, ...]); b.capture_graph(); b.step(120)

Each world is stepped by the same algorithm as World.step, one workgroup per world.
"""
from __future__ import annotations

import ctypes
import time

import numpy as np

from ._lib import check, lib

_F32P = ctypes.POINTER(ctypes.c_float)
_I32P = ctypes.POINTER(ctypes.c_int32)


class Batch:
    def __init__(self, scenes):
        assert len(scenes) > 0
        self.scenes = scenes
        self._descs = (type(scenes[0].desc()) * len(scenes))(*[s.desc() for s in scenes])
        self._params = scenes[0].params()
        h = ctypes.c_void_p()
        check(lib().pe_batch_create(ctypes.c_int32(len(scenes)), self._descs, ctypes.byref(self._params), ctypes.byref(h)))
        self.h = h
        self.n_bodies = sum(s.n for s in scenes)
        self.n_dynamic = sum(s.n_dynamic() for s in scenes)

    def close(self):
        if getattr(self, "h", None):
            lib().pe_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, n: int = 1):
        check(lib().pe_batch_step(self.h, ctypes.c_int32(n)))

    def sync(self):
        check(lib().pe_batch_sync(self.h))

    def step_timed(self, n: int) -> float:
        """Runs n steps; returns their device time in ms (hipEvents on the batch stream)."""
        ms = ctypes.c_double()
        check(lib().pe_batch_step_timed(self.h, ctypes.c_int32(n), ctypes.byref(ms)))
        return ms.value

    def stage_times(self, n: int):
        """n eager steps with events between stages: summed ms of (broadphase, narrowphase, solver)."""
        ms = (ctypes.c_double * 3)()
        check(lib().pe_batch_stage_times(self.h, ctypes.c_int32(n), ms))
        return list(ms)

    def phase_profile(self):
        """One step with in-kernel phase timestamps: summed microseconds per phase."""
        us = (ctypes.c_double * 8)()
        check(lib().pe_batch_phase_profile(self.h, us))
        names = ("integrate", "sort", "sweep", "narrowphase", "setup", "colouring", "gauss_seidel")
        return dict(zip(names, list(us)[:7]))

    def capture_graph(self):
        check(lib().pe_batch_capture_graph(self.h))

    def set_config(self, upper, block):
        u = (ctypes.c_int32 * len(upper))(*upper)
        bl = (ctypes.c_int32 * len(block))(*block)
        check(lib().pe_batch_set_config(self.h, ctypes.c_int32(len(upper)), u, bl))

    def config(self):
        nc = ctypes.c_int32()
        u, bl, cnt = (ctypes.c_int32 * 32)(), (ctypes.c_int32 * 32)(), (ctypes.c_int32 * 32)()
        check(lib().pe_batch_get_config(self.h, ctypes.byref(nc), u, bl, cnt, 32))
        return [(u[i], bl[i], cnt[i]) for i in range(nc.value)]

    def set_solver_groups(self, group_bodies):
        r = (ctypes.c_int32 * len(group_bodies))(*group_bodies)
        check(lib().pe_batch_set_solver_groups(self.h, ctypes.c_int32(len(group_bodies)), r))

    def solver_groups(self):
        nc = ctypes.c_int32()
        r = (ctypes.c_int32 * 32)()
        check(lib().pe_batch_get_solver_groups(self.h, ctypes.byref(nc), r, 32))
        return [r[i] for i in range(nc.value)]

    def world_state(self, w: int):
        n = self.scenes[w].n
        pos, rot = np.zeros((n, 3), np.float32), np.zeros((n, 3, 3), np.float32)
        v, av = np.zeros((n, 3), np.float32), np.zeros((n, 3), np.float32)
        f = lambda a: a.ctypes.data_as(_F32P)
        check(lib().pe_batch_get_world_state(self.h, ctypes.c_int32(w), f(pos), f(rot), f(v), f(av)))
        return pos, rot, v, av

    def counts(self):
        c = (ctypes.c_int64 * 7)()
        check(lib().pe_batch_get_counts(self.h, c))
        return dict(zip(("bodies", "dynamic", "pairs", "points", "max_colors", "overflow_worlds", "worlds"), list(c)))

    def _time_steps(self, steps):
        self.step(1)
        self.sync()
        t0 = time.perf_counter()
        self.step(steps)
        self.sync()
        return time.perf_counter() - t0

    def autotune(self, candidates=(64, 128, 256, 512), steps: int = 3, classes=None,
                 group_candidates=(0, 128, 256, 512, 1024)):
        """Per size class (other classes at their current setting), times every candidate block
        size, then every candidate solver group size (bodies per solve workgroup), and keeps the
        fastest of each; returns {class upper bound: (block, group bodies)}.  Stepping during
        tuning advances the worlds (it is real simulation work); results never depend on the
        choice."""
        cur = self.config()
        uppers = [c[0] for c in cur] if classes is None else list(classes)
        blocks = [c[1] for c in cur] if classes is None else [256] * len(uppers)
        self.set_config(uppers, blocks)
        rows = self.solver_groups()
        counts = [c[2] for c in self.config()]
        for ci in range(len(uppers)):
            if counts[ci] == 0:
                continue
            best, best_t = blocks[ci], None
            for cand in candidates:
                trial = list(blocks)
                trial[ci] = cand
                self.set_config(uppers, trial)
                self.set_solver_groups(rows)
                t = self._time_steps(steps)
                if best_t is None or t < best_t:
                    best, best_t = cand, t
            blocks[ci] = best
            self.set_config(uppers, blocks)
            best, best_t = rows[ci], None
            for cand in group_candidates:
                trial = list(rows)
                trial[ci] = cand
                self.set_solver_groups(trial)
                t = self._time_steps(steps)
                if best_t is None or t < best_t:
                    best, best_t = cand, t
            rows[ci] = best
        self.set_config(uppers, blocks)
        self.set_solver_groups(rows)
        return {u: (b, r) for u, b, r in zip(uppers, blocks, rows)}
