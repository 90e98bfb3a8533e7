"""Python mirror of the reference's ``crs::gpu::Batch`` over the C ABI.

Reference interface: ``src/gpu/gpu_solver.h:66-145``. Method names follow it
(``create``/``step``/``synchronize``/``download``/``download_positions``/
``upload``/``set_loads``/``upload_rod``/``set_kinematic_velocities``/
``set_material_scales``/``self_contact_overflow``/``launches_per_step``), with
two extensions: a batch may hold rods of different shapes (a list of rods
instead of one replicated prototype), and each rod may collide with its own
world. ``precision=64`` makes an FP64 batch (``rp_batch_create64``): the
reference CPU solver's own precision, computed on the GPU.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi
from .rod import CollisionWorld, Rod, SolverParams


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


class _Prec:
    """The ABI pieces of one precision."""

    def __init__(self, bits):
        self.bits = bits
        self.dtype = np.float32 if bits == 32 else np.float64
        self.suffix = "" if bits == 32 else "64"
        self.ptr = _abi.fptr if bits == 32 else _abi.dptr
        self.real = C.c_float if bits == 32 else C.c_double
        self.Primitive = getattr(_abi, "RpPrimitive" + self.suffix)
        self.WorldDesc = getattr(_abi, "RpWorldDesc" + self.suffix)
        self.BatchDesc = getattr(_abi, "RpBatchDesc" + self.suffix)
        self.StepParams = getattr(_abi, "RpStepParams" + self.suffix)

    def arr(self, a):
        return np.ascontiguousarray(a, dtype=self.dtype)

    def fn(self, name):
        return getattr(_abi.lib(), name + self.suffix)


_PREC = {32: _Prec(32), 64: _Prec(64)}


def _prec(bits) -> _Prec:
    if bits not in _PREC:
        raise ValueError("precision must be 32 or 64")
    return _PREC[bits]


def _world_desc(w: CollisionWorld, precision: int = 32):
    pr = _prec(precision)
    prims = (pr.Primitive * max(1, len(w.primitives)))()
    for k, p in enumerate(w.primitives):
        c = prims[k]
        c.type = int(p.type)
        c.a[:] = list(map(float, p.a))
        c.b[:] = list(map(float, p.b))
        c.normal[:] = list(map(float, p.normal))
        c.half_extents[:] = list(map(float, p.half_extents))
        c.rotation[:] = list(map(float, p.rotation))
        c.radius = float(p.radius)
        c.friction = float(p.friction)
    d = pr.WorldDesc(len(w.primitives), C.cast(prims, C.POINTER(pr.Primitive)),
                         float(w.contact_margin_radii), int(w.self_collision), float(w.self_friction),
                         int(w.self_skip), int(w.hash_table_size), int(w.self_capacity),
                         float(w.self_capacity_per_segment), int(w.self_order))
    return d, prims


def describe(rods: list, world=None, world_of_rod=None, precision: int = 32):
    """An ``rp_batch_desc`` (``rp_batch_desc64`` for precision 64) for
    ``rods`` (ragged CSR over the list) and their world(s): one
    ``CollisionWorld`` for every rod, or a list chosen per rod by
    ``world_of_rod``. Returns the descriptor, the arrays it points into (keep
    them alive while it is used) and the four CSR start arrays."""
    pr = _prec(precision)
    _r = pr.arr
    R = len(rods)
    if R == 0:
        raise ValueError("a batch needs at least one rod")
    nP = np.array([r.num_particles for r in rods])
    nF = np.array([r.num_frames for r in rods])
    nS = np.array([r.num_segments for r in rods])
    nB = np.array([len(r.bA) for r in rods])
    starts = tuple(_i32(np.concatenate([[0], np.cumsum(n)])) for n in (nP, nF, nS, nB))
    same = len({id(r) for r in rods}) == 1
    proto = rods[0]

    def cat(name):
        if same:
            a = getattr(proto, name)
            return np.tile(a, (R, 1)) if a.ndim == 2 else np.tile(a, R)
        return np.concatenate([getattr(r, name) for r in rods])

    keep = [starts]
    arr = {
        "inv_mass": _r(cat("invMass")), "inv_inertia": _r(cat("invInertia")), "inertia": _r(cat("inertia")),
        "stretch_p0": _i32(cat("sP0")), "stretch_p1": _i32(cat("sP1")), "stretch_seg": _i32(cat("sSeg")),
        "stretch_rest": _r(cat("sRest")), "stretch_compliance": _r(cat("sComp")),
        "bend_a": _i32(cat("bA")), "bend_b": _i32(cat("bB")), "bend_rest": _r(cat("bRest")),
        "bend_compliance": _r(cat("bComp")), "bend_rest_darboux": _r(cat("bRestD")),
        "radius": _r([r.material.radius for r in rods]),
    }
    keep.append(arr)
    if world is None:
        worlds = []
    elif isinstance(world, CollisionWorld):
        worlds = [world]
    else:
        worlds = list(world)
    wdescs = (pr.WorldDesc * max(1, len(worlds)))()
    for i, w in enumerate(worlds):
        d, prims = _world_desc(w, precision)
        keep.append(prims)
        wdescs[i] = d
    keep.append(wdescs)
    wor = None
    if world_of_rod is not None:
        wor = _i32(world_of_rod)
        keep.append(wor)
    desc = pr.BatchDesc()
    desc.num_rods = R
    desc.particle_start = _abi.iptr(starts[0])
    desc.frame_start = _abi.iptr(starts[1])
    desc.stretch_start = _abi.iptr(starts[2])
    desc.bend_start = _abi.iptr(starts[3])
    for k, a in arr.items():
        setattr(desc, k, pr.ptr(a) if a.dtype == pr.dtype else _abi.iptr(a))
    desc.num_worlds = len(worlds)
    desc.worlds = C.cast(wdescs, C.POINTER(pr.WorldDesc))
    desc.world_of_rod = _abi.iptr(wor)
    return desc, keep, starts


def step_params(p: SolverParams, precision: int = 32):
    """``rp_step_params`` (or ``rp_step_params64``) for ``p``."""
    pr = _prec(precision)
    g = (pr.real * 3)(*map(float, p.gravity))
    return pr.StepParams(float(p.dt), int(p.substeps), int(p.iterations), g, float(p.linear_damping),
                         float(p.angular_damping), int(p.contact_interval))


class Batch:
    """A batch of independent rods stepped on the GPU.

    ``Batch(prototype, num_rods)`` replicates one rod as the reference's
    ``Batch::create`` does; ``Batch([rod_a, rod_b, ...])`` packs rods of any
    shapes. ``worlds`` is one ``CollisionWorld`` for every rod or a list, with
    ``world_of_rod`` choosing per rod. ``precision`` is 32 (the reference GPU
    path's FP32) or 64 (its CPU path's FP64); state arrays come back in it.
    """

    def __init__(self, rods, num_rods: int | None = None, strategy: str = "auto", world=None,
                 world_of_rod=None, tile_config: int = 0, precision: int = 32):
        self._h = None
        self._p = _prec(precision)
        self.precision = precision
        if isinstance(rods, Rod):
            rods = [rods] * int(num_rods if num_rods is not None else 1)
        self.rods = list(rods)
        R = len(self.rods)
        if R == 0:
            raise ValueError("a batch needs at least one rod")
        desc, keep, starts = describe(self.rods, world, world_of_rod, precision)
        self._keep = keep
        self.particle_start, self.frame_start, self.stretch_start, self.bend_start = starts
        desc.strategy = _abi.STRATEGY[strategy]
        desc.tile_config = int(tile_config)
        h = C.c_void_p()
        _abi.check(self._p.fn("rp_batch_create")(C.byref(desc), C.byref(h)))
        self._h = h
        self.num_rods = R
        self.upload_all(self.rods)

    def __del__(self):
        self.destroy()

    def destroy(self):
        if getattr(self, "_h", None):
            _abi.lib().rp_batch_destroy(self._h)
            self._h = None

    # ------------------------------------------------------------------ state and loads

    def _range_sizes(self, first, count):
        if first < 0 or count < 0 or first + count > self.num_rods:
            raise ValueError(f"rod range [{first}, {first + count}) outside a batch of {self.num_rods}")
        p0, p1 = self.particle_start[first], self.particle_start[first + count]
        f0, f1 = self.frame_start[first], self.frame_start[first + count]
        return int(p1 - p0), int(f1 - f0)

    def set_state(self, first: int, count: int, x=None, v=None, q=None, omega=None):
        """Bulk state upload for rods [first, first + count), packed per rod."""
        P, F = self._range_sizes(first, count)
        arrs = [None if a is None else self._p.arr(a).reshape(-1) for a in (x, v, q, omega)]
        for a, n in zip(arrs, (P * 3, P * 3, F * 4, F * 3)):
            if a is not None and a.size != n:
                raise ValueError(f"state array has {a.size} values, expected {n}")
        _abi.check(self._p.fn("rp_batch_set_state")(self._h, first, count, *[self._p.ptr(a) for a in arrs]))

    def get_state(self, first: int = 0, count: int | None = None) -> dict:
        """Bulk read-back of x, v, q, omega for rods [first, first + count)."""
        count = self.num_rods - first if count is None else count
        P, F = self._range_sizes(first, count)
        dt = self._p.dtype
        out = {"x": np.empty((P, 3), dt), "v": np.empty((P, 3), dt), "q": np.empty((F, 4), dt),
               "omega": np.empty((F, 3), dt)}
        _abi.check(self._p.fn("rp_batch_get_state")(self._h, first, count,
                                                    *[self._p.ptr(out[k]) for k in ("x", "v", "q", "omega")]))
        return out

    def set_loads_range(self, first: int, count: int, force=None, kin_vel=None, torque=None, fixed_q=None):
        P, F = self._range_sizes(first, count)
        arrs = [None if a is None else self._p.arr(a).reshape(-1) for a in (force, kin_vel, torque, fixed_q)]
        for a, n in zip(arrs, (P * 3, P * 3, F * 3, F * 4)):
            if a is not None and a.size != n:
                raise ValueError(f"load array has {a.size} values, expected {n}")
        _abi.check(self._p.fn("rp_batch_set_loads")(self._h, first, count, *[self._p.ptr(a) for a in arrs]))

    def upload_all(self, rods):
        """Every rod's full state and loads, one bulk copy per field."""
        if len(rods) != self.num_rods:
            raise ValueError("need one rod per batch entry")
        same = len({id(r) for r in rods}) == 1

        def cat(name):
            if same:
                a = getattr(rods[0], name)
                return np.tile(a, (self.num_rods, 1))
            return np.concatenate([getattr(r, name) for r in rods])

        self.set_state(0, self.num_rods, cat("x"), cat("v"), cat("q"), cat("omega"))
        self.set_loads_range(0, self.num_rods, cat("extForce"), cat("kinVel"), cat("extTorque"))

    # Reference method set (gpu_solver.h:74-105).
    def upload(self, rod: Rod):
        self.upload_all([rod] * self.num_rods)

    def upload_rod(self, index: int, rod: Rod):
        self.set_state(index, 1, rod.x, rod.v, rod.q, rod.omega)
        self.set_loads(index, rod)

    def set_loads(self, index: int, rod: Rod):
        self.set_loads_range(index, 1, rod.extForce, rod.kinVel, rod.extTorque, rod.q)

    def set_kinematic_velocities(self, rod_major):
        self.set_loads_range(0, self.num_rods, kin_vel=rod_major)

    def set_material_scales(self, youngs_scale=None, friction_scale=None):
        ys = None if youngs_scale is None or len(youngs_scale) == 0 else self._p.arr(youngs_scale)
        fs = None if friction_scale is None or len(friction_scale) == 0 else self._p.arr(friction_scale)
        for a in (ys, fs):
            if a is not None and a.size != self.num_rods:
                raise ValueError(f"scale array has {a.size} values for {self.num_rods} rods")
        _abi.check(self._p.fn("rp_batch_set_scales")(self._h, self._p.ptr(ys), self._p.ptr(fs)))

    def step(self, params: SolverParams, steps: int = 1):
        sp = step_params(params, self.precision)
        _abi.check(self._p.fn("rp_batch_step")(self._h, C.byref(sp), int(steps)))

    def step_noalloc(self, params: SolverParams, steps: int = 1):
        """rp_batch_step_noalloc: a GLOBAL FP32 batch stepped with its working
        arrays in the library's fixed per-device region; allocates nothing."""
        sp = step_params(params, 32)
        _abi.check(_abi.lib().rp_batch_step_noalloc(self._h, C.byref(sp), int(steps)))

    # ------------------------------------------------------------------ drivers and observers

    def set_paths(self, first: int, paths: list):
        """Gripper paths for rods first.. (one `GripperPath` or None per rod)."""
        descs = (_abi.RpPathDesc * max(1, len(paths)))()
        keep = []
        for i, p in enumerate(paths):
            if p is None:
                descs[i].particle = -1
                continue
            wp = np.ascontiguousarray(np.asarray(p.waypoints, np.float64).reshape(-1, 2))
            legs = np.ascontiguousarray(np.asarray(p.leg_seconds, np.float64))
            keep += [wp, legs]
            descs[i] = _abi.RpPathDesc(int(p.particle), len(legs), wp.ctypes.data_as(C.POINTER(C.c_double)),
                                       legs.ctypes.data_as(C.POINTER(C.c_double)),
                                       (C.c_double * 2)(*map(float, p.start)), float(p.height),
                                       float(p.control_period), int(p.steps_per_control))
        _abi.check(_abi.lib().rp_batch_set_paths(self._h, first, len(paths), descs))

    def set_clock(self, first: int = 0, count: int | None = None, clock=None):
        count = self.num_rods - first if count is None else count
        c = None if clock is None else np.ascontiguousarray(clock, np.int32)
        _abi.check(_abi.lib().rp_batch_set_clock(self._h, first, count, _abi.iptr(c)))

    def get_clock(self, first: int = 0, count: int | None = None) -> np.ndarray:
        count = self.num_rods - first if count is None else count
        out = np.empty(count, np.int32)
        _abi.check(_abi.lib().rp_batch_get_clock(self._h, first, count, _abi.iptr(out)))
        return out

    def set_observers(self, first: int, observers: list):
        """(particle, every, z0) per rod, or None."""
        descs = (_abi.RpObserverDesc * max(1, len(observers)))()
        for i, o in enumerate(observers):
            descs[i] = _abi.RpObserverDesc(-1, 1, 0.0) if o is None else _abi.RpObserverDesc(int(o[0]), int(o[1]),
                                                                                                   float(o[2]))
        _abi.check(_abi.lib().rp_batch_set_observers(self._h, first, len(observers), descs))

    def get_observations(self, first: int = 0, count: int | None = None) -> np.ndarray:
        count = self.num_rods - first if count is None else count
        out = np.empty(count, np.float64)
        _abi.check(_abi.lib().rp_batch_get_observations(self._h, first, count,
                                                        out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    def record(self, every: int, max_frames: int):
        """Record every rod's positions on the device after each `every` steps
        from the rods' current clocks, up to `max_frames` frames (FP32, rod-major);
        every = 0 stops recording and frees the frames."""
        _abi.check(_abi.lib().rp_batch_record(self._h, int(every), int(max_frames)))

    def recorded_frames(self) -> int:
        """Frames that every rod has completed."""
        n = np.zeros(1, np.int32)
        _abi.check(_abi.lib().rp_batch_recorded_frames(self._h, _abi.iptr(n)))
        return int(n[0])

    def frames(self, first: int = 0, count: int | None = None) -> np.ndarray:
        """Frames [first, first + count) as float32 (count, particles, 3); count
        defaults to the frames completed so far."""
        count = self.recorded_frames() - first if count is None else count
        out = np.empty((max(count, 0), int(self.particle_start[-1]), 3), np.float32)
        _abi.check(_abi.lib().rp_batch_get_frames(self._h, int(first), int(count), _abi.fptr(out)))
        return out

    def tune_report(self) -> list:
        """(tile configuration name, trial ms) for each shape the autotuner timed."""
        ids = np.zeros(64, np.int32)
        ms = np.zeros(64, np.float32)
        n = _abi.lib().rp_batch_tune_report(self._h, _abi.iptr(ids), _abi.fptr(ms), 64)
        if n < 0:
            _abi.check(n)
        return [(_abi.lib().rp_tile_config_name(int(i)).decode(), float(m)) for i, m in zip(ids[:n], ms[:n])]

    def set_timing(self, enable: bool = True):
        """Time every step call with HIP events on the batch's stream."""
        _abi.check(_abi.lib().rp_batch_set_timing(self._h, int(enable)))

    def synchronize(self):
        _abi.check(_abi.lib().rp_batch_sync(self._h))

    def download(self, index: int) -> dict:
        return self.get_state(index, 1)

    def download_positions(self) -> np.ndarray:
        P = int(self.particle_start[-1])
        x = np.empty((P, 3), self._p.dtype)
        _abi.check(self._p.fn("rp_batch_get_state")(self._h, 0, self.num_rods, self._p.ptr(x), None, None, None))
        return x

    def counters(self) -> dict:
        c = _abi.RpCounters()
        _abi.check(_abi.lib().rp_batch_counters(self._h, C.byref(c)))
        d = {k: getattr(c, k) for k, _ in c._fields_}
        d["strategy"] = _abi.STRATEGY_NAME.get(d["strategy"], d["strategy"])
        if d["tile_config"]:
            d["tile_config_name"] = _abi.lib().rp_tile_config_name(d["tile_config"]).decode()
        return d

    def self_contact_overflow(self) -> int:
        return int(self.counters()["self_contact_overflow"])

    def launches_per_step(self) -> int:
        return int(self.counters()["launches_per_step"])

    def rod_slice(self, arr, index, kind="particle"):
        s = self.particle_start if kind == "particle" else self.frame_start
        return arr[s[index]:s[index + 1]]
