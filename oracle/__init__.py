"""Python binding of the CPU oracle -- TEST INFRASTRUCTURE.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg import
this package, and only as the checker. The product (``ropephysics_amd``)
never imports it.

Three libraries share one C interface (``oracle_capi.h``):

* ``oracle``  -- ``liboracle.so``, the restatement in ``rp_oracle.hpp``;
* ``ref64``   -- ``_ref/libref64.so``, the reference's own CPU sources (FP64);
* ``ref32``   -- ``_ref/libref32.so``, the same built with ``CRS_REAL_FLOAT``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_PATHS = {
    "oracle": os.path.join(_HERE, "liboracle.so"),
    "ref64": os.path.join(_HERE, "_ref", "libref64.so"),
    "ref32": os.path.join(_HERE, "_ref", "libref32.so"),
}
_LIBS: dict = {}

PLANE, SPHERE, CAPSULE, BOX = 0, 1, 2, 3

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class RodView(C.Structure):
    _fields_ = [("nP", C.c_int), ("nF", C.c_int), ("nStretch", C.c_int), ("nBend", C.c_int)] + [
        (n, _dp) for n in ("x", "xPrev", "v", "invMass", "mass", "extForce", "kinVel",
                           "q", "qPrev", "omega", "invInertia", "inertia", "extTorque")
    ] + [(n, _ip) for n in ("sP0", "sP1", "sSeg")] + [
        (n, _dp) for n in ("sRest", "sComp", "sLam")
    ] + [(n, _ip) for n in ("bA", "bB")] + [
        (n, _dp) for n in ("bRest", "bRestD", "bComp", "bLam")
    ] + [("material", C.c_double * 4)]


class Prim(C.Structure):
    _fields_ = [("type", C.c_int), ("a", C.c_double * 3), ("b", C.c_double * 3),
                ("normal", C.c_double * 3), ("half", C.c_double * 3), ("rot", C.c_double * 4),
                ("radius", C.c_double), ("friction", C.c_double)]


class WorldC(C.Structure):
    _fields_ = [("numPrims", C.c_int), ("prims", C.POINTER(Prim)), ("selfCollision", C.c_int),
                ("selfFriction", C.c_double), ("selfSkip", C.c_int), ("marginRadii", C.c_double),
                ("hashTableSize", C.c_int), ("selfOrder", C.c_int)]


class ParamsC(C.Structure):
    _fields_ = [("dt", C.c_double), ("substeps", C.c_int), ("iterations", C.c_int),
                ("gravity", C.c_double * 3), ("linDamp", C.c_double), ("angDamp", C.c_double),
                ("contactInterval", C.c_int), ("coloured", C.c_int),
                ("complianceScale", C.c_double), ("frictionScale", C.c_double)]


# Field name -> (size key, components, dtype)
FIELDS = {
    "x": ("nP", 3, float), "xPrev": ("nP", 3, float), "v": ("nP", 3, float),
    "invMass": ("nP", 1, float), "mass": ("nP", 1, float), "extForce": ("nP", 3, float),
    "kinVel": ("nP", 3, float), "q": ("nF", 4, float), "qPrev": ("nF", 4, float),
    "omega": ("nF", 3, float), "invInertia": ("nF", 3, float), "inertia": ("nF", 3, float),
    "extTorque": ("nF", 3, float), "sP0": ("nStretch", 1, int), "sP1": ("nStretch", 1, int),
    "sSeg": ("nStretch", 1, int), "sRest": ("nStretch", 1, float), "sComp": ("nStretch", 3, float),
    "sLam": ("nStretch", 3, float), "bA": ("nBend", 1, int), "bB": ("nBend", 1, int),
    "bRest": ("nBend", 1, float), "bRestD": ("nBend", 3, float), "bComp": ("nBend", 3, float),
    "bLam": ("nBend", 3, float),
}


def build(reference: str | None = "/root/reference") -> None:
    """Compile liboracle.so and, when the reference checkout exists, _ref/."""
    targets = ["all"]
    if reference and os.path.isdir(os.path.join(reference, "src", "core")):
        targets.append("ref")
    subprocess.run(["make", "-s", "-C", _HERE, f"REFERENCE={reference}"] + targets, check=True)


def available(kind: str) -> bool:
    return os.path.exists(_PATHS[kind])


def lib(kind: str = "oracle"):
    if kind in _LIBS:
        return _LIBS[kind]
    path = _PATHS[kind]
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is not built (run oracle.build())")
    L = C.CDLL(path)
    vp = C.c_void_p
    L.orc_rod_straight.restype = vp
    L.orc_rod_straight.argtypes = [C.c_int, C.c_int, C.c_double, _dp, _dp, _dp]
    L.orc_rod_centerline.restype = vp
    L.orc_rod_centerline.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, C.c_double]
    L.orc_rod_clamp_end.argtypes = [vp, C.c_int]
    L.orc_rod_pin.argtypes = [vp, C.c_int]
    L.orc_rod_set_rest_from_current.argtypes = [vp]
    L.orc_rod_sizes.argtypes = [vp, _ip]
    L.orc_rod_get.argtypes = [vp, C.POINTER(RodView)]
    L.orc_rod_set.argtypes = [vp, C.POINTER(RodView)]
    L.orc_rod_free.argtypes = [vp]
    L.orc_step.argtypes = [vp, C.POINTER(ParamsC), C.POINTER(WorldC), C.c_int]
    L.orc_contact_counts.argtypes = [vp, _ip, _ip]
    L.orc_step_many.argtypes = [C.POINTER(vp), C.c_int, C.POINTER(ParamsC),
                                C.POINTER(C.POINTER(WorldC)), C.c_int, C.c_int]
    L.orc_chol_solve3.argtypes = [C.c_int, _dp, _dp, _dp]
    L.orc_signed_distance.restype = C.c_double
    L.orc_signed_distance.argtypes = [C.c_int, C.POINTER(Prim), _dp, _dp]
    L.orc_seg_seg.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp]
    L.orc_hash_cell.restype = C.c_uint
    L.orc_hash_cell.argtypes = [C.c_int] * 4
    L.orc_cell_coord.argtypes = [C.c_int, C.c_double, C.c_double]
    L.orc_self_gap.argtypes = [vp, C.POINTER(WorldC)]
    L.orc_colour.argtypes = [vp, C.c_int, _ip, _ip, C.c_int]
    if hasattr(L, "orc_self_contacts"):
        L.orc_self_contacts.argtypes = [vp, _ip, _ip, _ip, C.c_int]
    if hasattr(L, "orc_count_ops"):
        L.orc_count_ops.argtypes = [vp, C.POINTER(ParamsC), C.POINTER(WorldC), C.c_int,
                                    C.POINTER(C.c_longlong)]
    _LIBS[kind] = L
    return L


def _d(a) -> C.Array:
    a = np.ascontiguousarray(a, dtype=np.float64).ravel()
    return (C.c_double * len(a))(*a)


# ------------------------------------------------------------------ worlds and params

def make_prim(type: int, a=(0, 0, 0), b=(0, 0, 0), normal=(0, 0, 1), half=(0, 0, 0),
              rot=(1, 0, 0, 0), radius=0.0, friction=0.0) -> Prim:
    p = Prim()
    p.type = type
    p.a[:] = list(map(float, a))
    p.b[:] = list(map(float, b))
    p.normal[:] = list(map(float, normal))
    p.half[:] = list(map(float, half))
    p.rot[:] = list(map(float, rot))
    p.radius = float(radius)
    p.friction = float(friction)
    return p


@dataclass
class World:
    """Collision world (reference src/core/collision.h:83-106)."""
    prims: list = field(default_factory=list)
    self_collision: bool = False
    self_friction: float = 0.0
    self_skip: int = 2
    margin_radii: float = 0.25
    hash_table_size: int = 4096
    self_order: int = 0  # 0: list order (the reference's); 1: coloured order (this build's opt-in)

    def c(self) -> WorldC:
        arr = (Prim * max(1, len(self.prims)))(*self.prims)
        w = WorldC(len(self.prims), C.cast(arr, C.POINTER(Prim)), int(self.self_collision),
                   float(self.self_friction), int(self.self_skip), float(self.margin_radii),
                   int(self.hash_table_size), int(self.self_order))
        w._keep = arr
        return w


@dataclass
class Params:
    """Step parameters (reference src/core/solver.h:15-37) plus per-rod scales."""
    dt: float = 1e-3
    substeps: int = 10
    iterations: int = 1
    gravity: tuple = (0.0, 0.0, -9.81)
    lin_damp: float = 0.0
    ang_damp: float = 0.0
    contact_interval: int = 1
    coloured: bool = True
    compliance_scale: float = 1.0
    friction_scale: float = 1.0

    def c(self) -> ParamsC:
        g = (C.c_double * 3)(*map(float, self.gravity))
        return ParamsC(float(self.dt), int(self.substeps), int(self.iterations), g,
                       float(self.lin_damp), float(self.ang_damp), int(self.contact_interval),
                       int(self.coloured), float(self.compliance_scale), float(self.friction_scale))


# ------------------------------------------------------------------ rods

class Rod:
    """One oracle rod (FP64 or FP32), owning a handle in one of the libraries."""

    def __init__(self, handle, kind: str, precision: int):
        if not handle:
            raise ValueError(f"{kind} refused precision {precision}")
        self.h = C.c_void_p(handle)
        self.kind = kind
        self.precision = precision
        self.L = lib(kind)

    def __del__(self):
        try:
            if self.h:
                self.L.orc_rod_free(self.h)
                self.h = None
        except Exception:
            pass

    @classmethod
    def straight(cls, n, length, material=(1e7, 0.35, 1000.0, 5e-3), origin=(0, 0, 0),
                 direction=(1, 0, 0), precision=64, kind="oracle"):
        L = lib(kind)
        h = L.orc_rod_straight(precision, int(n), float(length), _d(material), _d(origin), _d(direction))
        return cls(h, kind, precision)

    @classmethod
    def centerline(cls, points, material=(1e7, 0.35, 1000.0, 5e-3), seed=(0, 0, 1), twist=0.0,
                   precision=64, kind="oracle"):
        L = lib(kind)
        pts = np.asarray(points, dtype=np.float64)
        h = L.orc_rod_centerline(precision, len(pts), _d(pts), _d(material), _d(seed), float(twist))
        return cls(h, kind, precision)

    def clamp_end(self, tip: bool = False) -> int:
        return self.L.orc_rod_clamp_end(self.h, int(tip))

    def pin(self, i: int) -> None:
        self.L.orc_rod_pin(self.h, int(i))

    def set_rest_from_current(self) -> None:
        self.L.orc_rod_set_rest_from_current(self.h)

    def sizes(self) -> dict:
        out = (C.c_int * 4)()
        self.L.orc_rod_sizes(self.h, out)
        return {"nP": out[0], "nF": out[1], "nStretch": out[2], "nBend": out[3]}

    def get(self) -> dict:
        sz = self.sizes()
        arrays = {}
        view = RodView()
        for name, (key, comps, dt) in FIELDS.items():
            shape = (sz[key], comps) if comps > 1 else (sz[key],)
            a = np.zeros(shape, dtype=np.float64 if dt is float else np.int32)
            arrays[name] = a
            setattr(view, name, a.ctypes.data_as(_dp if dt is float else _ip))
        self.L.orc_rod_get(self.h, C.byref(view))
        arrays["material"] = np.array(list(view.material))
        return arrays

    def set(self, arrays: dict) -> None:
        cur = self.get()
        cur.update({k: v for k, v in arrays.items()})
        view = RodView()
        view.nP, view.nF = len(cur["x"]), len(cur["q"])
        view.nStretch, view.nBend = len(cur["sSeg"]), len(cur["bA"])
        keep = []
        for name, (key, comps, dt) in FIELDS.items():
            a = np.ascontiguousarray(cur[name], dtype=np.float64 if dt is float else np.int32)
            keep.append(a)
            setattr(view, name, a.ctypes.data_as(_dp if dt is float else _ip))
        view.material[:] = list(map(float, cur["material"]))
        self.L.orc_rod_set(self.h, C.byref(view))

    def step(self, params: Params, world: World | None = None, steps: int = 1) -> None:
        pc = params.c()
        wc = world.c() if world is not None else None
        rc = self.L.orc_step(self.h, C.byref(pc), C.byref(wc) if wc is not None else None, int(steps))
        if rc != 0:
            raise RuntimeError(f"{self.kind} refused the step (rc={rc})")

    def contact_counts(self) -> tuple:
        w, s = C.c_int(), C.c_int()
        self.L.orc_contact_counts(self.h, C.byref(w), C.byref(s))
        return w.value, s.value

    def self_contacts(self) -> dict:
        """The self contacts of the last contact generation, in list order: their
        particle indices (n, 4), and under the coloured order their colours (32:
        the sequential tail) and positions in the projection order (-1 otherwise)."""
        n = self.L.orc_self_contacts(self.h, None, None, None, 0)
        idx = np.zeros((max(n, 1), 4), np.int32)
        col = np.zeros(max(n, 1), np.int32)
        rank = np.zeros(max(n, 1), np.int32)
        self.L.orc_self_contacts(self.h, idx.ctypes.data_as(_ip), col.ctypes.data_as(_ip),
                                 rank.ctypes.data_as(_ip), n)
        return {"idx": idx[:n], "colour": col[:n], "rank": rank[:n]}

    def count_ops(self, params: Params, world: World | None = None, steps: int = 1) -> dict:
        """Floating-point operations of `steps` steps of a copy of this FP64 rod."""
        out = (C.c_longlong * 5)()
        pc = params.c()
        wc = world.c() if world is not None else None
        rc = self.L.orc_count_ops(self.h, C.byref(pc), C.byref(wc) if wc is not None else None, int(steps), out)
        if rc != 0:
            raise RuntimeError("op counting needs an FP64 oracle rod")
        return dict(zip(("add", "mul", "div", "sqrt", "cmp"), list(out)))

    def self_gap(self, world: World) -> int:
        wc = world.c()
        return self.L.orc_self_gap(self.h, C.byref(wc))

    def colouring(self, family: int):
        sz = self.sizes()
        count = sz["nStretch"] if family == 0 else sz["nBend"]
        order = (C.c_int * max(1, count))()
        start = (C.c_int * 65)()
        n = self.L.orc_colour(self.h, family, order, start, 64)
        return np.array(order[:count], dtype=np.int32), np.array(start[: n + 1], dtype=np.int32)


def step_many(rods: list, params: Params, worlds: list | None, steps: int, threads: int) -> None:
    """Step rods of one library on `threads` host threads (static partition)."""
    if not rods:
        return
    L = rods[0].L
    handles = (C.c_void_p * len(rods))(*[r.h.value for r in rods])
    pc = params.c()
    if worlds is not None:
        wcs = [w.c() if w is not None else None for w in worlds]
        ptrs = (C.POINTER(WorldC) * len(rods))(*[C.pointer(w) if w is not None else None for w in wcs])
    else:
        wcs, ptrs = None, None
    L.orc_step_many(handles, len(rods), C.byref(pc), ptrs, int(steps), int(threads))


def step_many_driven(rods: list, params: Params, worlds: list | None, steps: int, threads: int,
                     driven: list, kin_vel: np.ndarray) -> None:
    """step_many with a per-step kinematic velocity for one particle of each
    rod: driven[r] (-1: none) and kin_vel[r, s] (shape rods x steps x 3)."""
    if not rods:
        return
    L = rods[0].L
    handles = (C.c_void_p * len(rods))(*[r.h.value for r in rods])
    pc = params.c()
    if worlds is not None:
        wcs = [w.c() if w is not None else None for w in worlds]
        ptrs = (C.POINTER(WorldC) * len(rods))(*[C.pointer(w) if w is not None else None for w in wcs])
    else:
        wcs, ptrs = None, None
    kv = np.ascontiguousarray(kin_vel, dtype=np.float64)
    assert kv.shape == (len(rods), steps, 3)
    dr = (C.c_int * len(rods))(*[int(d) for d in driven])
    L.orc_step_many_driven(handles, len(rods), C.byref(pc), ptrs, int(steps), int(threads), dr,
                           kv.ctypes.data_as(C.POINTER(C.c_double)))


# ------------------------------------------------------------------ pinned pieces

def chol_solve3(A, b, precision=64, kind="oracle"):
    y = (C.c_double * 3)()
    ok = lib(kind).orc_chol_solve3(precision, _d(A), _d(b), y)
    return bool(ok), np.array(list(y))


def signed_distance(prim: Prim, p, precision=64, kind="oracle"):
    n = (C.c_double * 3)()
    d = lib(kind).orc_signed_distance(precision, C.byref(prim), _d(p), n)
    return d, np.array(list(n))


def seg_seg(p1, q1, p2, q2, precision=64, kind="oracle"):
    st = (C.c_double * 2)()
    lib(kind).orc_seg_seg(precision, _d(p1), _d(q1), _d(p2), _d(q2), st)
    return st[0], st[1]


def hash_cell(ix, iy, iz, table, kind="oracle") -> int:
    return lib(kind).orc_hash_cell(int(ix), int(iy), int(iz), int(table))


def cell_coord(v, cell, precision=64, kind="oracle") -> int:
    return lib(kind).orc_cell_coord(precision, float(v), float(cell))
