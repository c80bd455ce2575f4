"""Continuous-batching server of world-simulation requests (include/pe_amd_batch.h)."""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import check, lib
from .batch import _plain_scene
from .scenes import WorldParams

_F32P = ctypes.POINTER(ctypes.c_float)


class ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_int64), ("max_worlds", ctypes.c_int32),
                ("max_world_bodies", ctypes.c_int32), ("params", WorldParams)]


class Server:
    def __init__(self, hbm_budget_bytes: int, max_worlds: int = 8192, max_world_bodies: int = 4096,
                 dt: float = 1.0 / 60.0, gravity=(0.0, -9.8, 0.0), iterations: int = 10):
        self.cfg = ServerConfig(int(hbm_budget_bytes), max_worlds, max_world_bodies,
                                WorldParams(dt, (ctypes.c_double * 3)(*gravity), iterations))
        h = ctypes.c_void_p()
        check(lib().pe_server_create(ctypes.byref(self.cfg), ctypes.byref(h)))
        self.h = h
        self.sizes = {}
        self._keep = {}

    def close(self):
        if getattr(self, "h", None):
            lib().pe_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def footprint(scene) -> int:
        b = ctypes.c_int64()
        d = scene.desc()
        check(lib().pe_server_footprint(ctypes.byref(d), ctypes.byref(b)))
        return b.value

    def submit(self, scene, horizon: int) -> int:
        _plain_scene(scene)
        d = scene.desc()
        rid = ctypes.c_int64()
        check(lib().pe_server_submit(self.h, ctypes.byref(d), ctypes.c_int32(horizon), ctypes.byref(rid)))
        self.sizes[rid.value] = scene.n
        return rid.value

    def tick(self):
        lw, lb = ctypes.c_int32(), ctypes.c_int64()
        check(lib().pe_server_tick(self.h, ctypes.byref(lw), ctypes.byref(lb)))
        return lw.value, lb.value

    def poll(self):
        ids = (ctypes.c_int64 * 4096)()
        cnt = ctypes.c_int32()
        out = []
        while True:
            check(lib().pe_server_poll(self.h, ids, 4096, ctypes.byref(cnt)))
            out.extend(ids[i] for i in range(cnt.value))
            if cnt.value < 4096:
                return out

    def result(self, rid: int):
        n = self.sizes.pop(rid)
        pos, rot = np.zeros((n, 3), np.float32), np.zeros((n, 3, 3), np.float32)
        v, w = np.zeros((n, 3), np.float32), np.zeros((n, 3), np.float32)
        f = lambda a: a.ctypes.data_as(_F32P)
        check(lib().pe_server_result(self.h, ctypes.c_int64(rid), f(pos), f(rot), f(v), f(w)))
        return pos, rot, v, w

    def cancel(self, rid: int):
        check(lib().pe_server_cancel(self.h, ctypes.c_int64(rid)))

    def request_stats(self):
        """Per-request figures: admission wait and submission-to-result latency (p50 / p99, ms), queue
        depth over ticks (mean / max), requests finished, cancelled, refused for capacity."""
        out = (ctypes.c_double * 9)()
        check(lib().pe_server_request_stats(self.h, out))
        keys = ("wait_p50_ms", "wait_p99_ms", "latency_p50_ms", "latency_p99_ms", "queue_mean", "queue_max",
                "finished", "cancelled", "capacity_errors")
        return dict(zip(keys, list(out)))

    def stats(self):
        out = (ctypes.c_double * 8)()
        check(lib().pe_server_stats(self.h, out))
        keys = ("ticks", "p50_ms", "p99_ms", "mean_ms", "body_steps", "live", "queued", "admitted")
        return dict(zip(keys, list(out)))
