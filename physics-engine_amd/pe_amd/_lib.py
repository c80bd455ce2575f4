"""Loader for libpe_amd.so (the HIP library behind include/pe_amd.h).

The library is built in-tree by __graft_entry__.build() (physics-engine_amd/build.py).
There is no fallback: if the shared object is missing, loading raises.
"""
from __future__ import annotations

import ctypes
import os

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "lib", "libpe_amd.so")

PE_OK = 0
ERRORS = {-1: "PE_ERR_INVALID", -2: "PE_ERR_DEVICE", -3: "PE_ERR_CAPACITY", -4: "PE_ERR_UNSUPPORTED",
          -5: "PE_ERR_BUSY", -6: "PE_ERR_NOT_FOUND"}

# Every symbol include/pe_amd.h declares (checked by tests/test_abi.py).
EXPORTS = [
    "pe_last_error", "pe_device_count", "pe_mesh_prepare",
    "pe_world_create", "pe_world_destroy", "pe_world_step", "pe_world_integrate", "pe_world_broadphase",
    "pe_world_narrowphase", "pe_world_solve", "pe_world_sync", "pe_world_get_state", "pe_world_set_state",
    "pe_world_get_aabbs", "pe_world_get_pairs", "pe_world_get_contacts", "pe_world_get_impulses", "pe_world_get_solve_order",
    "pe_world_get_axis", "pe_world_set_axis", "pe_world_set_timing", "pe_world_get_stage_times",
    "pe_world_get_counts", "pe_world_set_solver_blocks", "pe_world_set_colouring", "pe_world_get_records", "pe_world_step_async",
    "pe_world_get_cache", "pe_world_set_cache", "pe_world_add_joints", "pe_world_add_ignore_pairs",
    "pe_world_set_sleep", "pe_world_remove_body", "pe_world_get_body_status", "pe_world_set_collision_reports",
    "pe_world_get_collision_events", "pe_world_raycast",
    "pe_stage_create", "pe_stage_destroy", "pe_stage_set_meshes", "pe_stage_broadphase", "pe_stage_narrowphase",
    "pe_stage_solve", "pe_stage_broadphase_noalloc",
    "pe_batch_create", "pe_batch_destroy", "pe_batch_step", "pe_batch_sync", "pe_batch_get_world_state",
    "pe_batch_get_counts", "pe_batch_set_config", "pe_batch_get_config", "pe_batch_capture_graph",
    "pe_batch_world_footprint", "pe_batch_step_timed", "pe_batch_set_solver_groups", "pe_batch_get_solver_groups", "pe_batch_phase_profile", "pe_batch_phase_trace", "pe_batch_stage_times",
    "pe_batch_class_times", "pe_batch_snapshot", "pe_batch_restore",
    "pe_server_create", "pe_server_destroy", "pe_server_footprint", "pe_server_submit", "pe_server_tick",
    "pe_server_poll", "pe_server_result", "pe_server_stats", "pe_server_cancel", "pe_server_request_stats",
]


class PEError(RuntimeError):
    def __init__(self, rc, msg):
        super().__init__(f"{ERRORS.get(rc, rc)}: {msg}")
        self.rc = rc


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} missing: build the HIP extension first (__graft_entry__.build())")
        _lib = ctypes.CDLL(LIB_PATH)
        _lib.pe_last_error.restype = ctypes.c_char_p
    return _lib


def check(rc: int):
    if rc != PE_OK:
        raise PEError(rc, lib().pe_last_error().decode())
    return rc
