"""MI355X-native batched XPBD Cosserat-rod stepper (RopePhysics' batched GPU
step, rebuilt for gfx950).

The compute path is ``libropephysics.so`` (HIP kernels plus the C ABI in
``include/ropephysics.h``); this package holds the host-side rod/scene model
a ``Batch`` mirroring the reference's ``crs::gpu::Batch``, and a continuous-batching
``Server`` that keeps requests resident on the device.
"""
from ._abi import RopePhysicsError, device_info, tile_configs
from .batch import Batch
from .server import Server, ServerFull
from .rod import (BOX, CAPSULE, PLANE, SPHERE, CollisionWorld, Primitive, Rod, RodMaterial, SolverParams,
                  add_bend_constraint, add_fixed_frame, clamp_end_exact, clamp_root_exact, make_rod_from_centerline,
                  make_straight_rod, quat_axis_angle, set_rest_from_current)

__all__ = [
    "Batch", "Server", "ServerFull", "RopePhysicsError", "device_info", "tile_configs", "Rod", "RodMaterial", "CollisionWorld", "Primitive",
    "SolverParams", "PLANE", "SPHERE", "CAPSULE", "BOX", "make_straight_rod", "make_rod_from_centerline",
    "clamp_end_exact", "clamp_root_exact", "add_fixed_frame", "add_bend_constraint", "set_rest_from_current",
    "quat_axis_angle",
]
