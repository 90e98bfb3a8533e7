"""Host-side rod, material and scene model (FP64), the reference's rod API.

Reference: ``src/core/rod.h:26-170`` and ``src/core/rod.cpp:11-319`` (material
constants, lumped masses, cylinder inertia, compliance, straight and
centre-line builders, the exact end clamp through a ghost frame) and
``src/core/geometry.h:17-65`` / ``src/core/collision.h:83-106`` (primitives
and collision worlds).

Every formula evaluates in the reference's operation order with IEEE double
scalars (explicit component arithmetic, no BLAS), so a rod built here equals
the reference builder's rod bit for bit; the batch API rounds it to FP32 on
upload, as the reference's ``Batch`` does.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

PI = 3.14159265358979323846

# ------------------------------------------------------------------ scalar helpers


def _dot(a, b):
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


def _cross(a, b):
    return (a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0])


def _scale(a, s):
    return (a[0] * s, a[1] * s, a[2] * s)


def _add(a, b):
    return (a[0] + b[0], a[1] + b[1], a[2] + b[2])


def _sub(a, b):
    return (a[0] - b[0], a[1] - b[1], a[2] - b[2])


def _norm(a):
    return math.sqrt(_dot(a, a))


def _normalize(a):
    return _scale(a, 1.0 / _norm(a))


def _qmul(a, b):
    av, bv = a[1:], b[1:]
    w = a[0] * b[0] - _dot(av, bv)
    v = _add(_add(_scale(bv, a[0]), _scale(av, b[0])), _cross(av, bv))
    return (w, v[0], v[1], v[2])


def _qnormalize(q):
    n = math.sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3])
    s = 1.0 / n
    return (q[0] * s, q[1] * s, q[2] * s, q[3] * s)


def _qconj(q):
    return (q[0], -q[1], -q[2], -q[3])


def _hemi_identity(q):
    d = q[0] * 1.0 + q[1] * 0.0 + q[2] * 0.0 + q[3] * 0.0
    return (q[0] * -1.0, q[1] * -1.0, q[2] * -1.0, q[3] * -1.0) if d < 0.0 else q


def quat_from_to(frm, to):
    """Shortest rotation taking unit ``frm`` onto unit ``to`` (math3.h:299-314)."""
    d = _dot(frm, to)
    if d > 1.0 - 1e-12:
        return (1.0, 0.0, 0.0, 0.0)
    if d < -1.0 + 1e-12:
        axis = _cross(frm, (1.0, 0.0, 0.0))
        if _dot(axis, axis) < 1e-12:
            axis = _cross(frm, (0.0, 1.0, 0.0))
        axis = _normalize(axis)
        return (0.0, axis[0], axis[1], axis[2])
    c = _cross(frm, to)
    s = math.sqrt((1.0 + d) * 2.0)
    v = _scale(c, 1.0 / s)
    return _qnormalize((s * 0.5, v[0], v[1], v[2]))


def quat_from_mat(R):
    """Shepperd's method (math3.h:316-340); R is row-major 3x3."""
    t = R[0][0] + R[1][1] + R[2][2]
    if t > 0.0:
        s = math.sqrt(t + 1.0) * 2.0
        q = (0.25 * s, (R[2][1] - R[1][2]) / s, (R[0][2] - R[2][0]) / s, (R[1][0] - R[0][1]) / s)
    elif R[0][0] > R[1][1] and R[0][0] > R[2][2]:
        s = math.sqrt(1.0 + R[0][0] - R[1][1] - R[2][2]) * 2.0
        q = ((R[2][1] - R[1][2]) / s, 0.25 * s, (R[0][1] + R[1][0]) / s, (R[0][2] + R[2][0]) / s)
    elif R[1][1] > R[2][2]:
        s = math.sqrt(1.0 + R[1][1] - R[0][0] - R[2][2]) * 2.0
        q = ((R[0][2] - R[2][0]) / s, (R[0][1] + R[1][0]) / s, 0.25 * s, (R[1][2] + R[2][1]) / s)
    else:
        s = math.sqrt(1.0 + R[2][2] - R[0][0] - R[1][1]) * 2.0
        q = ((R[1][0] - R[0][1]) / s, (R[0][2] + R[2][0]) / s, (R[1][2] + R[2][1]) / s, 0.25 * s)
    return _qnormalize(q)


def quat_axis_angle(axis, angle):
    h = angle * 0.5
    a = _scale(_normalize(axis), math.sin(h))
    return (math.cos(h), a[0], a[1], a[2])


def darboux(qa, qb, rest_len):
    """Discrete Darboux vector of a joint (rod.cpp:41-47)."""
    rel = _hemi_identity(_qmul(_qconj(qa), qb))
    return _scale(rel[1:], 2.0 / rest_len)


# ------------------------------------------------------------------ material


@dataclass
class RodMaterial:
    """Isotropic circular section (rod.h:26-45)."""
    youngs: float = 1e7
    poisson: float = 0.35
    density: float = 1000.0
    radius: float = 5e-3

    def shear(self):
        return self.youngs / (2.0 * (1.0 + self.poisson))

    def area(self):
        return PI * self.radius * self.radius

    def second_moment(self):
        r2 = self.radius * self.radius
        return PI * r2 * r2 / 4.0

    def polar_moment(self):
        return 2.0 * self.second_moment()

    def shear_correction(self):
        return 6.0 * (1.0 + self.poisson) / (7.0 + 6.0 * self.poisson)

    def bend_stiffness(self):
        return self.youngs * self.second_moment()

    def twist_stiffness(self):
        return self.shear() * self.polar_moment()

    def as_tuple(self):
        return (self.youngs, self.poisson, self.density, self.radius)


def stretch_compliance(m: RodMaterial, l):
    """diag(1/kGA, 1/kGA, 1/EA)/l (rod.cpp:135-140); l may be an array."""
    A = m.area()
    EA = m.youngs * A
    GA = m.shear_correction() * m.shear() * A
    l = np.asarray(l, dtype=np.float64)
    return np.stack([1.0 / (GA * l), 1.0 / (GA * l), 1.0 / (EA * l)], axis=-1)


def bend_compliance(m: RodMaterial, l):
    """diag(1/EI, 1/EI, 1/GJ)/lbar (rod.cpp:142-146); l may be an array."""
    EI = m.bend_stiffness()
    GJ = m.twist_stiffness()
    l = np.asarray(l, dtype=np.float64)
    return np.stack([1.0 / (EI * l), 1.0 / (EI * l), 1.0 / (GJ * l)], axis=-1)


# ------------------------------------------------------------------ rod


@dataclass
class Rod:
    """Discrete Cosserat rod in SoA form (rod.h:49-121). Vectors are (n, 3),
    quaternions (n, 4) as (w, x, y, z)."""
    x: np.ndarray
    xPrev: np.ndarray
    v: np.ndarray
    invMass: np.ndarray
    mass: np.ndarray
    extForce: np.ndarray
    kinVel: np.ndarray
    q: np.ndarray
    qPrev: np.ndarray
    omega: np.ndarray
    invInertia: np.ndarray
    inertia: np.ndarray
    extTorque: np.ndarray
    sP0: np.ndarray
    sP1: np.ndarray
    sSeg: np.ndarray
    sRest: np.ndarray
    sComp: np.ndarray
    bA: np.ndarray
    bB: np.ndarray
    bRest: np.ndarray
    bRestD: np.ndarray
    bComp: np.ndarray
    material: RodMaterial = field(default_factory=RodMaterial)

    @property
    def num_particles(self):
        return len(self.x)

    @property
    def num_frames(self):
        return len(self.q)

    @property
    def num_segments(self):
        return len(self.sSeg)

    def copy(self) -> "Rod":
        kw = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in self.__dict__.items()}
        kw["material"] = RodMaterial(*self.material.as_tuple())
        return Rod(**kw)

    # Boundary conditions (rod.cpp:51-64).
    def pin_particle(self, i: int) -> None:
        self.invMass[i] = 0.0
        self.v[i] = 0.0

    def clamp_segment(self, j: int) -> None:
        self.invInertia[j] = 0.0
        self.omega[j] = 0.0

    def rest_length(self) -> float:
        total = 0.0
        for l in self.sRest:
            total += float(l)
        return total

    def arrays(self) -> dict:
        """The fields the oracle's rod view carries, for handing the same rod to
        a checker."""
        return {k: v for k, v in self.__dict__.items() if isinstance(v, np.ndarray)} | {
            "material": np.array(self.material.as_tuple())}


def set_rest_from_current(rod: Rod) -> None:
    """Rest lengths and rest Darboux vectors from the current pose
    (rod.cpp:118-133)."""
    x = rod.x
    d = x[rod.sP1] - x[rod.sP0]
    rod.sRest = np.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2])
    n_seg = len(rod.sSeg)
    for k in range(len(rod.bA)):
        a, b = int(rod.bA[k]), int(rod.bB[k])
        if a < n_seg and b < n_seg:
            rod.bRest[k] = 0.5 * (rod.sRest[a] + rod.sRest[b])
        rod.bRestD[k] = darboux(tuple(rod.q[a]), tuple(rod.q[b]), float(rod.bRest[k]))


def _finish_build(pts: np.ndarray, frames: np.ndarray, mat: RodMaterial) -> Rod:
    """Masses, inertias, chain topology, rest state and compliance
    (rod.cpp:219-283)."""
    nP = len(pts)
    nS = nP - 1
    d = pts[1:] - pts[:-1]
    l = np.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2])
    rho, A, r = mat.density, mat.area(), mat.radius
    seg_mass = rho * A * l
    mass = np.zeros(nP)
    half = 0.5 * seg_mass
    # mass[j] += half[j]; mass[j+1] += half[j], in segment order.
    for j in range(nS):
        mass[j] += half[j]
        mass[j + 1] += half[j]
    tr = seg_mass * (3.0 * r * r + l * l) / 12.0
    ax = 0.5 * seg_mass * r * r
    inertia = np.stack([tr, tr, ax], axis=-1)
    with np.errstate(divide="ignore"):
        inv_inertia = np.where(inertia > 0.0, 1.0 / np.where(inertia > 0.0, inertia, 1.0), 0.0)
    inv_mass = 1.0 / mass
    nB = nS - 1
    rod = Rod(
        x=pts.copy(), xPrev=pts.copy(), v=np.zeros((nP, 3)), invMass=inv_mass, mass=mass,
        extForce=np.zeros((nP, 3)), kinVel=np.zeros((nP, 3)), q=frames.copy(), qPrev=frames.copy(),
        omega=np.zeros((nS, 3)), invInertia=inv_inertia, inertia=inertia, extTorque=np.zeros((nS, 3)),
        sP0=np.arange(nS, dtype=np.int32), sP1=np.arange(1, nS + 1, dtype=np.int32),
        sSeg=np.arange(nS, dtype=np.int32), sRest=np.zeros(nS), sComp=np.zeros((nS, 3)),
        bA=np.arange(nB, dtype=np.int32), bB=np.arange(1, nB + 1, dtype=np.int32), bRest=np.zeros(nB),
        bRestD=np.zeros((nB, 3)), bComp=np.zeros((nB, 3)), material=RodMaterial(*mat.as_tuple()),
    )
    set_rest_from_current(rod)
    rod.sComp = stretch_compliance(mat, rod.sRest)
    rod.bComp = bend_compliance(mat, rod.bRest)
    return rod


def make_straight_rod(num_segments: int, length: float, material: RodMaterial, origin=(0.0, 0.0, 0.0),
                      direction=(1.0, 0.0, 0.0)) -> Rod:
    """makeStraightRod (rod.cpp:285-299)."""
    dirn = _normalize(tuple(map(float, direction)))
    l = length / num_segments
    i = np.arange(num_segments + 1, dtype=np.float64)
    li = l * i
    o = tuple(map(float, origin))
    pts = np.stack([o[0] + dirn[0] * li, o[1] + dirn[1] * li, o[2] + dirn[2] * li], axis=-1)
    frame = quat_from_to((0.0, 0.0, 1.0), dirn)
    frames = np.tile(np.array(frame, dtype=np.float64), (num_segments, 1))
    return _finish_build(pts, frames, material)


def make_rod_from_centerline(points, material: RodMaterial, frame_seed=(0.0, 0.0, 1.0),
                             twist_per_length: float = 0.0) -> Rod:
    """makeRodFromCenterline with parallel-transported frames (rod.cpp:301-337)."""
    pts = np.asarray(points, dtype=np.float64)
    nS = len(pts) - 1
    tang = [_normalize(_sub(tuple(pts[j + 1]), tuple(pts[j]))) for j in range(nS)]
    seed = tuple(map(float, frame_seed))
    t0 = tang[0]
    d1 = _sub(seed, _scale(t0, _dot(seed, t0)))
    if _dot(d1, d1) < 1e-20:
        d1 = _sub((1.0, 0.0, 0.0), _scale(t0, _dot((1.0, 0.0, 0.0), t0)))
        if _dot(d1, d1) < 1e-20:
            d1 = _sub((0.0, 1.0, 0.0), _scale(t0, _dot((0.0, 1.0, 0.0), t0)))
    d1 = _normalize(d1)
    c1 = _cross(t0, d1)
    R = [[d1[i], c1[i], t0[i]] for i in range(3)]
    frames = [quat_from_mat(R)]
    for j in range(1, nS):
        frames.append(_qnormalize(_qmul(quat_from_to(tang[j - 1], tang[j]), frames[j - 1])))
    if twist_per_length != 0.0:
        s = 0.0
        for j in range(nS):
            l = _norm(_sub(tuple(pts[j + 1]), tuple(pts[j])))
            s_mid = s + 0.5 * l
            frames[j] = _qnormalize(_qmul(frames[j], quat_axis_angle((0.0, 0.0, 1.0), twist_per_length * s_mid)))
            s += l
    return _finish_build(pts, np.array(frames, dtype=np.float64), material)


def add_fixed_frame(rod: Rod, orientation) -> int:
    """A zero-inertia frame that never moves (rod.cpp:149-158)."""
    q = np.array(_qnormalize(tuple(map(float, orientation))))
    rod.q = np.vstack([rod.q, q])
    rod.qPrev = np.vstack([rod.qPrev, q])
    z = np.zeros((1, 3))
    rod.omega = np.vstack([rod.omega, z])
    rod.extTorque = np.vstack([rod.extTorque, z])
    rod.inertia = np.vstack([rod.inertia, z])
    rod.invInertia = np.vstack([rod.invInertia, z])
    return len(rod.q) - 1


def add_bend_constraint(rod: Rod, seg_a: int, seg_b: int, element_length: float, rest_darboux=(0.0, 0.0, 0.0)):
    rod.bA = np.append(rod.bA, np.int32(seg_a)).astype(np.int32)
    rod.bB = np.append(rod.bB, np.int32(seg_b)).astype(np.int32)
    rod.bRest = np.append(rod.bRest, element_length)
    rod.bRestD = np.vstack([rod.bRestD, np.array(rest_darboux, dtype=np.float64)])
    rod.bComp = np.vstack([rod.bComp, bend_compliance(rod.material, element_length)])


def clamp_end_exact(rod: Rod, tip: bool) -> int:
    """Pin the end particle and tie the end segment to a fixed ghost frame
    through a half-length bend element (rod.cpp:169-184). Returns the ghost."""
    last = len(rod.sSeg) - 1
    seg = last if tip else 0
    particle = rod.num_particles - 1 if tip else 0
    half = 0.5 * rod.sRest[seg]
    rod.pin_particle(particle)
    ghost = add_fixed_frame(rod, tuple(rod.q[seg]))
    if tip:
        add_bend_constraint(rod, seg, ghost, half)
    else:
        add_bend_constraint(rod, ghost, seg, half)
    return ghost


def clamp_root_exact(rod: Rod) -> int:
    return clamp_end_exact(rod, False)


# ------------------------------------------------------------------ primitives and worlds

PLANE, SPHERE, CAPSULE, BOX = 0, 1, 2, 3


@dataclass
class Primitive:
    """TPrimitive (geometry.h:17-65)."""
    type: int = PLANE
    a: tuple = (0.0, 0.0, 0.0)
    b: tuple = (0.0, 0.0, 0.0)
    normal: tuple = (0.0, 0.0, 0.0)
    half_extents: tuple = (0.0, 0.0, 0.0)
    rotation: tuple = (1.0, 0.0, 0.0, 0.0)
    radius: float = 0.0
    friction: float = 0.0

    @staticmethod
    def plane(point, unit_normal, friction):
        return Primitive(PLANE, a=tuple(map(float, point)), normal=_normalize(tuple(map(float, unit_normal))),
                         friction=float(friction))

    @staticmethod
    def sphere(centre, radius, friction):
        return Primitive(SPHERE, a=tuple(map(float, centre)), radius=float(radius), friction=float(friction))

    @staticmethod
    def capsule(end_a, end_b, radius, friction):
        return Primitive(CAPSULE, a=tuple(map(float, end_a)), b=tuple(map(float, end_b)), radius=float(radius),
                         friction=float(friction))

    @staticmethod
    def box(centre, half_extents, rotation, friction):
        return Primitive(BOX, a=tuple(map(float, centre)), half_extents=tuple(map(float, half_extents)),
                         rotation=tuple(map(float, rotation)), friction=float(friction))


@dataclass
class CollisionWorld:
    """CollisionWorld (collision.h:83-106)."""
    primitives: list = field(default_factory=list)
    self_collision: bool = False
    self_friction: float = 0.0
    self_skip: int = 2
    contact_margin_radii: float = 0.25
    hash_table_size: int = 4096
    # Self contacts stored per rod by the batch: self_capacity when positive,
    # else max(16, ceil(self_capacity_per_segment * segments)) when that is
    # positive, else max(16, segments / 2), the reference's pool. Contacts
    # beyond it are counted as overflow.
    self_capacity: int = 0
    self_capacity_per_segment: float = 0.0
    # Order of the self-contact projection: 0, one at a time in list order (the
    # reference's); 1, the coloured order (an opt-in with no counterpart in the
    # reference: greedy colours over the list, a colour's contacts at once).
    self_order: int = 0


@dataclass
class SolverParams:
    """SolverParams / BatchParams (solver.h:15-37, gpu_solver.h:52-60)."""
    dt: float = 1e-3
    substeps: int = 10
    iterations: int = 1
    gravity: tuple = (0.0, 0.0, -9.81)
    linear_damping: float = 0.0
    angular_damping: float = 0.0
    contact_interval: int = 1
