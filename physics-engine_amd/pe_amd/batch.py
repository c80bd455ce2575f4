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


def load_schedule(name: str) -> dict:
    """A launch schedule from the tuning table shipped beside this module (schedules.json): the
    size-class upper bounds, block size and solver group size per class that Batch.autotune chose
    for a named workload, applied with Batch.apply_schedule."""
    import json
    import os
    with open(os.path.join(os.path.dirname(__file__), "schedules.json")) as f:
        return json.load(f)[name]


def _plain_scene(scene):
    """Batched worlds carry no joints, ignore-pair lists, sleep thresholds or lifetimes (their entry
    points are single-world ones)."""
    if getattr(scene, "joints", None) or getattr(scene, "ignore_pairs", None):
        raise ValueError(f"scene {scene.name!r}: joints and ignore pairs need a World, not a batch")
    sleep = getattr(scene, "sleep", None)
    if (sleep and (sleep[0] != 0.0 or sleep[1] != 0.0)) or (
            getattr(scene, "lifetime", None) is not None and bool(np.any(np.asarray(scene.lifetime) < 1e30))):
        raise ValueError(f"scene {scene.name!r}: sleeping and lifetimes need a World, not a batch")
    if getattr(scene, "compounds", None):
        raise ValueError(f"scene {scene.name!r}: compound bodies need a World, not a batch")


class Batch:
    def __init__(self, scenes):
        assert len(scenes) > 0
        for s in scenes:
            _plain_scene(s)
        self.scenes = scenes
        self._descs = (type(scenes[0].desc()) * len(scenes))(*[s.desc() for s in scenes])
        self._params = scenes[0].params()
        h = ctypes.c_void_p()
        check(lib().pe_batch_create(ctypes.c_int32(len(scenes)), self._descs, ctypes.byref(self._params), ctypes.byref(h)))
        self.h = h
        self.n_bodies = sum(s.n for s in scenes)
        self.n_dynamic = sum(s.n_dynamic() for s in scenes)
        self.steps_done = 0   # World::step calls made on every world so far

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
        self.steps_done += n

    def sync(self):
        check(lib().pe_batch_sync(self.h))

    def step_timed(self, n: int) -> float:
        """Runs n steps; returns their device time in ms (hipEvents on the batch stream)."""
        ms = ctypes.c_double()
        check(lib().pe_batch_step_timed(self.h, ctypes.c_int32(n), ctypes.byref(ms)))
        self.steps_done += n
        return ms.value

    def stage_times(self, n: int):
        """n eager steps with events between stages: summed ms of (broadphase, narrowphase, solver)."""
        ms = (ctypes.c_double * 3)()
        check(lib().pe_batch_stage_times(self.h, ctypes.c_int32(n), ms))
        self.steps_done += n
        return list(ms)

    def phase_profile(self):
        """One step with in-kernel phase timestamps: summed microseconds per phase."""
        us = (ctypes.c_double * 8)()
        check(lib().pe_batch_phase_profile(self.h, us))
        self.steps_done += 1
        names = ("integrate", "sort", "sweep", "narrowphase", "setup", "colouring", "gauss_seidel")
        return dict(zip(names, list(us)[:7]))

    def phase_trace(self):
        """One step with in-kernel phase timestamps: [nworlds, 8] array of 10 ns ticks, 0 where not
        recorded (K1: start, integrate, sort, sweep; K3: start, rows, colouring, sweeps, on the first
        world of each solver group)."""
        ts = np.zeros((len(self.scenes), 8), np.uint64)
        check(lib().pe_batch_phase_trace(self.h, ts.ctypes.data_as(ctypes.c_void_p)))
        self.steps_done += 1
        return ts

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

    def apply_schedule(self, sched: dict):
        """Sets the size classes, block sizes and solver group sizes of a schedule (load_schedule, or
        {"class_upper_bodies", "block", "solver_group_bodies"} from autotune's result)."""
        self.set_config(list(sched["class_upper_bodies"]), list(sched["block"]))
        self.set_solver_groups(list(sched["solver_group_bodies"]))

    def schedule(self) -> dict:
        cfg = self.config()
        return {"class_upper_bodies": [c[0] for c in cfg], "block": [c[1] for c in cfg],
                "solver_group_bodies": self.solver_groups()}

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

    def _whole_step_ms(self, steps: int) -> float:
        """Device ms per step of the current configuration as a captured graph replays it, from
        the snapshot, best of three runs of steps steps, after dealing the classes onto streams
        by their cost.  (Eager steps rank configurations differently: their launch gaps differ.)"""
        self.restore()
        self.class_times(2)
        self.capture_graph()
        best = float("inf")
        for _ in range(3):
            self.restore()
            best = min(best, self.step_timed(steps) / steps)
        return best

    def snapshot(self):
        """Copies the device state of every world aside (pe_batch_snapshot)."""
        check(lib().pe_batch_snapshot(self.h))
        self._snap_steps = self.steps_done

    def restore(self):
        """Puts the last snapshot back (pe_batch_restore): the next step is the one that followed it."""
        check(lib().pe_batch_restore(self.h))
        self.steps_done = self._snap_steps

    def _refine(self, uppers, blocks, rows, candidates, group_candidates, steps, passes=2):
        def apply(bl, rw):
            self.set_config(uppers, bl)
            self.set_solver_groups(rw)
        apply(blocks, rows)
        self.restore()
        cost = self.class_times(steps)
        counts = [c[2] for c in self.config()]
        steps = max(steps, 5)
        best_t = self._whole_step_ms(steps)
        for _ in range(passes):
            changed = False
            for c in sorted(range(len(uppers)), key=lambda k: -(cost[k][0] + cost[k][1])):
                if counts[c] == 0:
                    continue
                for cb in candidates:
                    for cg in group_candidates:
                        if (cb, cg) == (blocks[c], rows[c]):
                            continue
                        tb, tr = list(blocks), list(rows)
                        tb[c], tr[c] = cb, cg
                        apply(tb, tr)
                        t = self._whole_step_ms(steps)
                        if t < best_t * 0.995:
                            best_t, blocks, rows, changed = t, tb, tr, True
                apply(blocks, rows)
            if not changed:
                break
        return blocks, rows

    def class_times(self, n: int):
        """n eager steps with every size class launched alone, in turn: per class the summed ms of
        (integrate + broadphase + narrowphase, solver).  Also deals the classes onto streams by these
        costs."""
        nc = len(self.config())
        ms = (ctypes.c_double * (2 * nc))()
        check(lib().pe_batch_class_times(self.h, ctypes.c_int32(n), ms, ctypes.c_int32(2 * nc)))
        self.steps_done += n
        return [(ms[2 * c], ms[2 * c + 1]) for c in range(nc)]

    def autotune(self, candidates=(64, 128, 256, 512), steps: int = 3, classes=None,
                 group_candidates=(0, 128, 256, 512, 1024), refine: bool = True):
        """Picks each size class's block size and solver group size (bodies per solve workgroup).
        Every timing starts from one snapshot of the worlds (snapshot / restore), so all
        candidates are timed on the same state, and the worlds are left as they were: tuning
        does not advance them.  Results never depend on the choice.

        First every candidate block size is timed on all size classes at once, each class
        launched alone and timed by its own events so the classes do not disturb each other's
        figures, and each class keeps its fastest (integrate + broadphase + solver); then the same
        for the group size (solver time).  With refine, the choice is revised on the whole step,
        where the classes run concurrently and compete for the CUs: class by class, costliest
        first, every (block, group) pair is timed in whole graph-replayed steps and kept when it
        shortens them (at most two passes over the classes).  Finally the chosen configuration is
        timed once more, which deals the classes onto streams by cost.  Returns {class upper
        bound: (block, group bodies)}."""
        cur = self.config()
        uppers = [c[0] for c in cur] if classes is None else list(classes)
        nc = len(uppers)
        self.snapshot()
        self.set_config(uppers, [256] * nc)
        rows = self.solver_groups()

        def class_times():
            self.restore()
            self.class_times(1)
            self.restore()
            return self.class_times(steps)

        best = [None] * nc
        for cand in candidates:
            self.set_config(uppers, [cand] * nc)
            self.set_solver_groups(rows)
            t = class_times()
            for c in range(nc):
                tot = t[c][0] + t[c][1]
                if best[c] is None or tot < best[c][1]:
                    best[c] = (cand, tot)
        blocks = [b[0] for b in best]
        self.set_config(uppers, blocks)
        best = [None] * nc
        for cand in group_candidates:
            self.set_solver_groups([cand] * nc)
            t = class_times()
            for c in range(nc):
                if best[c] is None or t[c][1] < best[c][1]:
                    best[c] = (cand, t[c][1])
        rows = [b[0] for b in best]
        self.set_solver_groups(rows)
        if refine:
            blocks, rows = self._refine(uppers, blocks, rows, candidates, group_candidates, steps)
        class_times()
        self.restore()
        return {u: (b, r) for u, b, r in zip(uppers, blocks, rows)}
