"""Continuous-batching server over the C ABI (``rp_server_*``).

Requests (rods with their own collision world, an initial state, per-rod
stiffness and friction scales and a horizon in steps) are admitted while the
server's HBM budget has room, stay resident on the device between calls, and
advance together: every ``step()`` launches one chunk of steps over the rods
of every running request (those without self-collision and those with it in
two launches that share the device), each rod stopping at its own horizon.
There is no counterpart in the reference, whose callers build one ``Batch``
per workload; the arithmetic is the batch's, so a request's result equals the
one it gets alone, bit for bit.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi
from .batch import _f32, describe, step_params
from .rod import SolverParams

# Tile shapes the server uses by default: three particles per lane in 256-lane
# workgroups for rods without self-collision; for the others one per lane in
# 512-lane workgroups, or two per lane in 256-lane workgroups when the server
# is time-sliced.
DEFAULT_TILE_CONFIG = 13
DEFAULT_SLICED_TILE_CONFIG = 9
DEFAULT_LEAN_TILE_CONFIG = 10


class ServerFull(RuntimeError):
    """The request does not fit the free part of the budget now."""


class Server:
    def __init__(self, params: SolverParams, hbm_budget_bytes: int, max_rods: int = 0, tile_config: int = 0,
                 steps_per_launch: int = 0, self_tile_config: int = 0, time_slice_us: int = 0):
        self._h = None
        d = _abi.RpServerDesc()
        d.hbm_budget_bytes = int(hbm_budget_bytes)
        d.max_rods = int(max_rods)
        d.params = step_params(params)
        d.tile_config = int(tile_config)
        d.steps_per_launch = int(steps_per_launch)
        d.self_tile_config = int(self_tile_config)
        d.time_slice_us = int(time_slice_us)
        h = C.c_void_p()
        _abi.check(_abi.lib().rp_server_create(C.byref(d), C.byref(h)))
        self._h = h
        self.params = params
        self._sizes = {}  # id -> (particles, frames)

    def __del__(self):
        self.destroy()

    def destroy(self):
        if getattr(self, "_h", None):
            _abi.lib().rp_server_destroy(self._h)
            self._h = None

    def submit(self, rods: list, world, steps: int, youngs_scale=None, friction_scale=None, x=None,
               world_of_rod=None) -> int:
        """Admit a request; returns its id. ``x`` (packed positions) overrides
        the rods' own. Raises ServerFull when it does not fit now. The uploads
        are queued behind the launch in flight, not waited for."""
        desc, keep, starts = describe(list(rods), world, world_of_rod)
        R = len(rods)

        def cat(name):
            return _f32(np.concatenate([getattr(r, name) for r in rods]))

        st = _abi.RpRequestState()
        arrays = {"x": _f32(x).reshape(-1) if x is not None else cat("x"), "v": cat("v"), "q": cat("q"),
                  "omega": cat("omega"), "force": cat("extForce"), "kin_vel": cat("kinVel"),
                  "torque": cat("extTorque")}
        if arrays["x"].size != int(starts[0][-1]) * 3:
            raise ValueError("x has the wrong size for these rods")
        for name, scale in (("youngs_scale", youngs_scale), ("friction_scale", friction_scale)):
            if scale is not None:
                a = _f32(scale).reshape(-1)
                if a.size != R:
                    raise ValueError(f"{name} needs one value per rod")
                arrays[name] = a
        for k, a in arrays.items():
            setattr(st, k, _abi.fptr(a))
        rid = C.c_int(0)
        rc = _abi.lib().rp_server_submit(self._h, C.byref(desc), C.byref(st), int(steps), C.byref(rid))
        del keep
        if rc == _abi.RP_ERR_CAPACITY:
            raise ServerFull(_abi.lib().rp_last_error().decode())
        _abi.check(rc)
        self._sizes[rid.value] = (int(starts[0][-1]), int(starts[1][-1]))
        return rid.value

    def step(self) -> int:
        """Launch one chunk of steps for every running request; returns how
        many requests reach their horizon in it."""
        n = C.c_int(0)
        _abi.check(_abi.lib().rp_server_step(self._h, C.byref(n)))
        return n.value

    def synchronize(self):
        _abi.check(_abi.lib().rp_server_sync(self._h))

    def finished(self) -> list:
        n = _abi.lib().rp_server_finished(self._h, None, 0)
        if n <= 0:
            return []
        ids = np.zeros(n, np.int32)
        _abi.lib().rp_server_finished(self._h, _abi.iptr(ids), n)
        return [int(i) for i in ids]

    def info(self, rid: int) -> dict:
        out = _abi.RpRequestInfo()
        _abi.check(_abi.lib().rp_server_request_info(self._h, int(rid), C.byref(out)))
        return {k: getattr(out, k) for k, _ in out._fields_}

    def result(self, rid: int) -> dict:
        P, F = self._sizes[rid]
        x = np.zeros((P, 3), np.float32)
        v = np.zeros((P, 3), np.float32)
        q = np.zeros((F, 4), np.float32)
        om = np.zeros((F, 3), np.float32)
        _abi.check(_abi.lib().rp_server_result(self._h, int(rid), _abi.fptr(x), _abi.fptr(q), _abi.fptr(v),
                                               _abi.fptr(om)))
        return {"x": x, "v": v, "q": q, "omega": om}

    def release(self, rid: int):
        _abi.check(_abi.lib().rp_server_release(self._h, int(rid)))
        self._sizes.pop(rid, None)

    def stats(self) -> dict:
        out = _abi.RpServerStats()
        _abi.check(_abi.lib().rp_server_get_stats(self._h, C.byref(out)))
        return {k: getattr(out, k) for k, _ in out._fields_}
