"""ctypes binding of the C ABI in include/ropephysics.h.

Loads ``libropephysics.so`` built in-tree by ``ropephysics_amd.build``. There is
no fallback: if the library is missing or no GPU is usable, calls raise.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libropephysics.so")

RP_OK, RP_ERR_INVALID, RP_ERR_DEVICE, RP_ERR_MEMORY, RP_ERR_CAPACITY, RP_ERR_STATE = 0, -1, -2, -3, -4, -5
STRATEGY = {"plain": 0, "tiled": 1, "auto": 2, "global": 3}
STRATEGY_NAME = {v: k for k, v in STRATEGY.items()}

_fp = C.POINTER(C.c_float)
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def _batch_structs(real, suffix):
    """The primitive, world, batch and step-parameter structs over one real
    type: rp_primitive ... rp_step_params (FP32) or their *64 twins."""
    rp = C.POINTER(real)

    class Primitive(C.Structure):
        _fields_ = [("type", C.c_int), ("a", real * 3), ("b", real * 3), ("normal", real * 3),
                    ("half_extents", real * 3), ("rotation", real * 4), ("radius", real), ("friction", real)]

    class WorldDesc(C.Structure):
        _fields_ = [("num_prims", C.c_int), ("prims", C.POINTER(Primitive)), ("contact_margin_radii", real),
                    ("self_collision", C.c_int), ("self_friction", real), ("self_skip", C.c_int),
                    ("hash_table_size", C.c_int), ("self_capacity", C.c_int),
                    ("self_capacity_per_segment", real), ("self_order", C.c_int)]

    class BatchDesc(C.Structure):
        _fields_ = [("num_rods", C.c_int), ("particle_start", _ip), ("frame_start", _ip), ("stretch_start", _ip),
                    ("bend_start", _ip), ("inv_mass", rp), ("inv_inertia", rp), ("inertia", rp),
                    ("stretch_p0", _ip), ("stretch_p1", _ip), ("stretch_seg", _ip), ("stretch_rest", rp),
                    ("stretch_compliance", rp), ("bend_a", _ip), ("bend_b", _ip), ("bend_rest", rp),
                    ("bend_compliance", rp), ("bend_rest_darboux", rp), ("radius", rp), ("num_worlds", C.c_int),
                    ("worlds", C.POINTER(WorldDesc)), ("world_of_rod", _ip), ("strategy", C.c_int),
                    ("tile_config", C.c_int)]

    class StepParams(C.Structure):
        _fields_ = [("dt", real), ("substeps", C.c_int), ("iterations", C.c_int), ("gravity", real * 3),
                    ("linear_damping", real), ("angular_damping", real), ("contact_interval", C.c_int)]

    for cls, name in ((Primitive, "RpPrimitive"), (WorldDesc, "RpWorldDesc"), (BatchDesc, "RpBatchDesc"),
                      (StepParams, "RpStepParams")):
        cls.__name__ = cls.__qualname__ = name + suffix
    return Primitive, WorldDesc, BatchDesc, StepParams


RpPrimitive, RpWorldDesc, RpBatchDesc, RpStepParams = _batch_structs(C.c_float, "")
RpPrimitive64, RpWorldDesc64, RpBatchDesc64, RpStepParams64 = _batch_structs(C.c_double, "64")


class RpPathDesc(C.Structure):
    _fields_ = [("particle", C.c_int), ("num_waypoints", C.c_int), ("waypoints", C.POINTER(C.c_double)),
                ("leg_seconds", C.POINTER(C.c_double)), ("start", C.c_double * 2), ("height", C.c_double),
                ("control_period", C.c_double), ("steps_per_control", C.c_int)]


class RpObserverDesc(C.Structure):
    _fields_ = [("particle", C.c_int), ("every", C.c_int), ("z0", C.c_double)]


class RpDeviceInfo(C.Structure):
    _fields_ = [("name", C.c_char * 256), ("arch", C.c_char * 64), ("compute_units", C.c_int),
                ("wavefront_size", C.c_int), ("lds_per_workgroup", C.c_size_t), ("total_global_mem", C.c_size_t),
                ("clock_khz", C.c_int), ("max_tiled_segments", C.c_int)]


class RpCounters(C.Structure):
    _fields_ = [("self_contact_overflow", C.c_longlong), ("strategy", C.c_int), ("tile_config", C.c_int),
                ("workgroups", C.c_int), ("lanes_busy", C.c_int), ("lanes_total", C.c_int),
                ("lds_bytes_per_workgroup", C.c_size_t), ("launches_per_step", C.c_int),
                ("steps_taken", C.c_longlong), ("device_bytes", C.c_size_t), ("last_step_gpu_ms", C.c_float),
                ("last_step_launches", C.c_int), ("motion_margin", C.c_double), ("low_margin_steps", C.c_longlong),
                ("scratch_bytes", C.c_size_t)]


class RpServerDesc(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_longlong), ("max_rods", C.c_int), ("params", RpStepParams),
                ("tile_config", C.c_int), ("steps_per_launch", C.c_int), ("self_tile_config", C.c_int),
                ("time_slice_us", C.c_int)]


class RpRequestState(C.Structure):
    _fields_ = [(n, _fp) for n in ("x", "v", "q", "omega", "force", "kin_vel", "torque", "youngs_scale",
                                   "friction_scale")]


class RpRequestInfo(C.Structure):
    _fields_ = [("num_rods", C.c_int), ("num_particles", C.c_int), ("num_frames", C.c_int), ("steps", C.c_int),
                ("steps_done", C.c_int), ("device_bytes", C.c_longlong)]


class RpServerStats(C.Structure):
    _fields_ = [("resident_requests", C.c_int), ("running_requests", C.c_int), ("running_rods", C.c_int),
                ("bytes_in_use", C.c_longlong), ("device_bytes", C.c_longlong),
                ("capacity_particles", C.c_longlong), ("used_particles", C.c_longlong),
                ("capacity_rods", C.c_longlong), ("used_rods", C.c_longlong), ("launches", C.c_longlong),
                ("workgroups", C.c_int), ("tile_config", C.c_int), ("steps_per_launch", C.c_int)]


# Every symbol include/ropephysics.h declares, with its signature.
SIGNATURES = {
    "rp_device_query": (C.c_int, [C.POINTER(RpDeviceInfo)]),
    "rp_batch_create": (C.c_int, [C.POINTER(RpBatchDesc), C.POINTER(C.c_void_p)]),
    "rp_batch_destroy": (None, [C.c_void_p]),
    "rp_batch_set_state": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _fp, _fp, _fp, _fp]),
    "rp_batch_set_loads": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _fp, _fp, _fp, _fp]),
    "rp_batch_set_scales": (C.c_int, [C.c_void_p, _fp, _fp]),
    "rp_batch_step": (C.c_int, [C.c_void_p, C.POINTER(RpStepParams), C.c_int]),
    "rp_batch_step_noalloc": (C.c_int, [C.c_void_p, C.POINTER(RpStepParams), C.c_int]),
    "rp_batch_sync": (C.c_int, [C.c_void_p]),
    "rp_batch_get_state": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _fp, _fp, _fp, _fp]),
    "rp_batch_counters": (C.c_int, [C.c_void_p, C.POINTER(RpCounters)]),
    "rp_batch_set_timing": (C.c_int, [C.c_void_p, C.c_int]),
    "rp_batch_tune_report": (C.c_int, [C.c_void_p, _ip, _fp, C.c_int]),
    "rp_batch_set_paths": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.POINTER(RpPathDesc)]),
    "rp_batch_set_clock": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _ip]),
    "rp_batch_get_clock": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _ip]),
    "rp_batch_set_observers": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.POINTER(RpObserverDesc)]),
    "rp_batch_get_observations": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_double)]),
    "rp_batch_record": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "rp_batch_recorded_frames": (C.c_int, [C.c_void_p, _ip]),
    "rp_batch_get_frames": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _fp]),
    "rp_batch_create64": (C.c_int, [C.POINTER(RpBatchDesc64), C.POINTER(C.c_void_p)]),
    "rp_batch_set_state64": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _dp, _dp, _dp, _dp]),
    "rp_batch_set_loads64": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _dp, _dp, _dp, _dp]),
    "rp_batch_set_scales64": (C.c_int, [C.c_void_p, _dp, _dp]),
    "rp_batch_step64": (C.c_int, [C.c_void_p, C.POINTER(RpStepParams64), C.c_int]),
    "rp_batch_get_state64": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _dp, _dp, _dp, _dp]),
    "rp_batch_precision": (C.c_int, [C.c_void_p]),
    "rp_tile_config_name": (C.c_char_p, [C.c_int]),
    "rp_tile_config_count": (C.c_int, []),
    "rp_last_error": (C.c_char_p, []),
    "rp_server_create": (C.c_int, [C.POINTER(RpServerDesc), C.POINTER(C.c_void_p)]),
    "rp_server_destroy": (None, [C.c_void_p]),
    "rp_server_submit": (C.c_int, [C.c_void_p, C.POINTER(RpBatchDesc), C.POINTER(RpRequestState), C.c_int, _ip]),
    "rp_server_step": (C.c_int, [C.c_void_p, _ip]),
    "rp_server_sync": (C.c_int, [C.c_void_p]),
    "rp_server_finished": (C.c_int, [C.c_void_p, _ip, C.c_int]),
    "rp_server_request_info": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(RpRequestInfo)]),
    "rp_server_result": (C.c_int, [C.c_void_p, C.c_int, _fp, _fp, _fp, _fp]),
    "rp_server_release": (C.c_int, [C.c_void_p, C.c_int]),
    "rp_server_get_stats": (C.c_int, [C.c_void_p, C.POINTER(RpServerStats)]),
}

_LIB = None


class RopePhysicsError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"ropephysics error {code}: {msg}")
        self.code = code


def lib():
    """The loaded HIP library. Raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RopePhysicsError(RP_ERR_DEVICE, f"{LIB_PATH} is not built; run ropephysics_amd.build.build()")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def check(rc: int) -> None:
    if rc != RP_OK:
        raise RopePhysicsError(rc, lib().rp_last_error().decode())


def fptr(a):
    return None if a is None else a.ctypes.data_as(_fp)


def dptr(a):
    return None if a is None else a.ctypes.data_as(_dp)


def iptr(a):
    return None if a is None else a.ctypes.data_as(_ip)


def device_info() -> dict:
    info = RpDeviceInfo()
    check(lib().rp_device_query(C.byref(info)))
    return {"name": info.name.decode(), "arch": info.arch.decode(), "compute_units": info.compute_units,
            "wavefront_size": info.wavefront_size, "lds_per_workgroup": info.lds_per_workgroup,
            "total_global_mem": info.total_global_mem, "clock_khz": info.clock_khz,
            "max_tiled_segments": info.max_tiled_segments}


def tile_configs() -> dict:
    L = lib()
    return {i: L.rp_tile_config_name(i).decode() for i in range(1, L.rp_tile_config_count() + 1)}
