"""Application tasks on top of the batch: gripper paths, cable routing and
cable hanging, with their verdicts.

Reference: ``src/apps/harness_routing.{h,cpp}`` (gripper schedule, board,
routing verdict) and ``src/apps/cable_hanging.{h,cpp}`` (drape builder, bar
and floor, hold/slide outcome). The routing geometry here is the benchmark's:
capsule pegs of 15 mm radius, a box clip (a bar the cable must pass beneath)
and eight waypoints; the reference's own harness uses 12 mm pegs, a two-post
clip and five waypoints. Verdict rules are stated on each function.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .rod import (PI, CollisionWorld, Primitive, Rod, RodMaterial, SolverParams, make_rod_from_centerline,
                  make_straight_rod)


# ------------------------------------------------------------------ gripper paths


@dataclass
class GripperPath:
    """A driven (pinned) particle following waypoints on a fixed schedule.

    Position: starting at ``start``, a smoothstep blend u^2 (3 - 2u) to each
    waypoint over its leg, then holding at the last one; ``height`` is the
    gripper's z (HarnessTask::gripperAt, harness_routing.cpp:69-84). Velocity
    during step s: (P(t1) - P(t0)) / period with t0 = floor(s / k) * period,
    t1 = t0 + period, k = steps_per_control (gripperVelocity,
    harness_routing.cpp:86-92). The device evaluates the same in FP64.
    """
    particle: int
    waypoints: np.ndarray          # (W, 2)
    leg_seconds: np.ndarray        # (W,)
    start: tuple
    height: float
    control_period: float
    steps_per_control: int = 10

    def at(self, t: float):
        fx, fy = float(self.start[0]), float(self.start[1])
        t0 = 0.0
        for k in range(len(self.leg_seconds)):
            leg = float(self.leg_seconds[k])
            t1 = t0 + leg
            if t < t1:
                u = (t - t0) / leg
                u = u * u * (3.0 - 2.0 * u)
                wx, wy = float(self.waypoints[k][0]), float(self.waypoints[k][1])
                return (fx + (wx - fx) * u, fy + (wy - fy) * u, self.height)
            fx, fy = float(self.waypoints[k][0]), float(self.waypoints[k][1])
            t0 = t1
        return (fx, fy, self.height)

    def velocity(self, step: int):
        tick = step // self.steps_per_control
        a = self.at(float(tick + 1) * self.control_period)
        b = self.at(float(tick) * self.control_period)
        inv = 1.0 / self.control_period
        return ((a[0] - b[0]) * inv, (a[1] - b[1]) * inv, (a[2] - b[2]) * inv)


def _crossing_side(cable: np.ndarray, px: float, py: float):
    """Signed y-offset from (px, py) of the cable's crossing of the line x = px
    nearest py (0 when it never crosses) -- HarnessTask::evaluate's sideOf
    (harness_routing.cpp:103-117)."""
    best, side = 1e9, 0.0
    for i in range(len(cable) - 1):
        p, q = cable[i], cable[i + 1]
        if (p[0] - px) * (q[0] - px) > 0:
            continue
        u = (px - p[0]) / (q[0] - p[0]) if abs(q[0] - p[0]) > 1e-9 else 0.0
        y = p[1] + (q[1] - p[1]) * u
        if abs(y - py) < best:
            best = abs(y - py)
            side = y - py
    return side


# ------------------------------------------------------------------ cable routing


@dataclass
class RoutingTask:
    """A cable plugged into a connector at the origin, its other end held by a
    gripper that must lay it below peg A, above peg B and beneath the clip bar.
    Board z = 0. Reference structure: HarnessTask (harness_routing.h:29-91)."""
    connector: tuple = (0.0, 0.0, 0.0)
    peg_a: tuple = (0.25, -0.06, 0.0)
    peg_b: tuple = (0.55, 0.06, 0.0)
    peg_radius: float = 0.015
    peg_height: float = 0.06
    clip: tuple = (0.85, 0.0, 0.0)
    clip_half: tuple = (0.012, 0.06, 0.006)   # box half extents
    clip_bottom: float = 0.010                 # bar underside above the board
    board_friction: float = 0.5
    cable_length: float = 1.2
    cable_youngs: float = 5e6
    cable_radius: float = 0.004
    cable_density: float = 1100.0
    self_friction: float = 0.3
    hash_table_size: int = 256
    # Self-contact pool per cable: 1.5 slots per segment. A cable lying in
    # coils holds up to about 1.13 contacts per segment (measured over the
    # routing workload); the reference's max(16, n / 2) drops contacts there,
    # where the CPU path keeps every one.
    self_capacity_per_segment: float = 1.5
    slack: float = 0.015
    leg_seconds: tuple = (1.2, 0.6, 0.5, 0.5, 0.5, 0.6, 0.6, 0.5)
    hold_seconds: float = 1.0
    steps_per_control: int = 10

    def material(self) -> RodMaterial:
        return RodMaterial(self.cable_youngs, 0.35, self.cable_density, self.cable_radius)

    def build(self, segments: int) -> Rod:
        """Straight from the connector towards -y, both ends pinned
        (HarnessTask::build, harness_routing.cpp:20-26)."""
        mat = self.material()
        start = (self.connector[0], self.connector[1], self.connector[2] + mat.radius)
        rod = make_straight_rod(segments, self.cable_length, mat, start, (0.0, -1.0, 0.0))
        rod.pin_particle(0)
        rod.pin_particle(segments)
        return rod

    def world(self) -> CollisionWorld:
        mu = self.board_friction
        w = CollisionWorld(self_collision=True, self_friction=self.self_friction,
                           hash_table_size=self.hash_table_size,
                           self_capacity_per_segment=self.self_capacity_per_segment)
        w.primitives.append(Primitive.plane((0.0, 0.0, 0.0), (0.0, 0.0, 1.0), mu))
        for peg in (self.peg_a, self.peg_b):
            w.primitives.append(Primitive.capsule(peg, (peg[0], peg[1], peg[2] + self.peg_height), self.peg_radius, mu))
        cz = self.clip_bottom + self.clip_half[2]
        w.primitives.append(Primitive.box((self.clip[0], self.clip[1], cz), self.clip_half, (1.0, 0.0, 0.0, 0.0), mu))
        return w

    def params(self) -> SolverParams:
        return SolverParams(dt=1e-3, substeps=10, iterations=1, gravity=(0.0, 0.0, -9.81))

    def steps(self) -> int:
        return int(round((sum(self.leg_seconds) + self.hold_seconds) / self.params().dt))

    def base_waypoints(self) -> np.ndarray:
        """Under peg A, over peg B, then past the clip to an end point that
        leaves the cable nearly taut, on the -y side so that the stretch from
        peg B crosses the clip's line 2-3 cm from its centre, well inside the
        6 cm half-width: a routed cable is routed by a clear margin."""
        a, b, c = self.peg_a, self.peg_b, self.clip
        return np.array([(0.10, -0.30), (a[0], a[1] - 0.06), (0.33, -0.03), (0.40, 0.0), (0.48, 0.05),
                         (b[0], b[1] + 0.06), (c[0] - 0.10, 0.0), (c[0] + 0.28, -0.03)], dtype=np.float64)

    def hand_written_waypoints(self) -> np.ndarray:
        """Straight at the targets, as HarnessTask::handWrittenPolicy
        (src/apps/harness_routing.cpp:62-67) writes it: the base path's first
        seven waypoints, then an end point on the clip's line 18 cm past it
        (y = 0) instead of 28 cm past it. The cable is left slack and lies
        several centimetres off peg A, so it is not routed. This is the
        starting point of harness learning and the motion that the learned
        one is compared with."""
        w = self.base_waypoints()
        w[-1] = (self.clip[0] + 0.18, 0.0)
        return w

    def wrong_side_waypoints(self) -> np.ndarray:
        """The base path with its second waypoint above peg A instead of
        below: the cable ends on the wrong side of A, by centimetres."""
        w = self.base_waypoints()
        w[1] = (self.peg_a[0], self.peg_a[1] + 0.06)
        return w

    def path(self, segments: int, waypoints: np.ndarray) -> GripperPath:
        mat = self.material()
        return GripperPath(particle=segments, waypoints=np.asarray(waypoints, np.float64),
                           leg_seconds=np.array(self.leg_seconds, np.float64),
                           start=(self.connector[0], self.connector[1] - self.cable_length), height=mat.radius,
                           control_period=self.params().dt * self.steps_per_control,
                           steps_per_control=self.steps_per_control)

    def evaluate(self, cable: np.ndarray) -> dict:
        """Routed = below peg A and above peg B, each laid against the peg
        (the crossing of the peg's x-line within peg radius + cable radius +
        1.5 cm slack), and through the clip: some segment crossing the clip's
        x-line between its ends (|y - clip.y| < clip half-width) and below the
        bar's underside. Rules as HarnessTask::evaluate
        (harness_routing.cpp:94-154) with the box clip in place of two posts."""
        cable = np.asarray(cable, np.float64)
        touching = self.peg_radius + self.cable_radius
        a = _crossing_side(cable, self.peg_a[0], self.peg_a[1])
        b = _crossing_side(cable, self.peg_b[0], self.peg_b[1])
        below_a = a < 0 and -a <= touching + self.slack
        above_b = b > 0 and b <= touching + self.slack
        miss = 1e9
        cx = self.clip[0]
        for i in range(len(cable) - 1):
            p, q = cable[i], cable[i + 1]
            if (p[0] - cx) * (q[0] - cx) > 0:
                continue
            u = (cx - p[0]) / (q[0] - p[0]) if abs(q[0] - p[0]) > 1e-9 else 0.0
            cz = p[2] + (q[2] - p[2]) * u
            if cz > self.clip_bottom:
                continue
            cy = p[1] + (q[1] - p[1]) * u
            miss = min(miss, abs(cy - self.clip[1]))
        through = miss < self.clip_half[1]

        # Smooth score for learning, 3 when routed: each peg 1 when the cable
        # lies against its correct side, fading to 0.5 further out, 0 on the
        # wrong side; the clip 1 when threaded, fading with the miss distance
        # (HarnessTask::evaluate's score, harness_routing.cpp:147-153).
        def clamp01(v):
            return max(0.0, min(1.0, v))

        def peg_score(offset):
            if offset <= 0:
                return 0.0
            return 1.0 - 0.5 * clamp01((offset - (touching + self.slack)) / 0.1)

        half = self.clip_half[1]
        score = peg_score(-a) + peg_score(b) + (1.0 if through else clamp01(1.0 - (miss - half) / 0.1))
        return {"below_a": bool(below_a), "above_b": bool(above_b), "through_clip": bool(through),
                "routed": bool(below_a and above_b and through), "side_a": a, "side_b": b, "clip_miss": miss,
                "score": score}

    def clamp_reach(self, waypoints: np.ndarray) -> np.ndarray:
        """Waypoints pulled back to the cable's reach from the connector
        (toPolicy, application_cases.cpp:303-312)."""
        w = np.array(waypoints, np.float64).reshape(-1, 2)
        c = np.array(self.connector[:2], np.float64)
        for k in range(len(w)):
            reach = float(np.linalg.norm(w[k] - c))
            if reach > self.cable_length:
                w[k] = c + (w[k] - c) * (self.cable_length / reach)
        return w


# ------------------------------------------------------------------ cable hanging


@dataclass
class HangingCable:
    """A cable draped over a horizontal bar and released (HangingCable,
    cable_hanging.h:28-63) at the benchmark's sizes: 1 m, 120 segments,
    r = 3 mm, bar radius 10 mm."""
    length: float = 1.0
    segments: int = 120
    bar_radius: float = 0.010
    bar_height: float = 0.8
    bar_half_length: float = 0.3
    duration: float = 8.0
    radius: float = 0.003
    density: float = 1100.0
    poisson: float = 0.35
    youngs: float = 5e6
    dt: float = 5e-4
    substeps: int = 10
    iterations: int = 1

    def material(self, youngs: float | None = None) -> RodMaterial:
        return RodMaterial(self.youngs if youngs is None else youngs, self.poisson, self.density, self.radius)

    def build(self, leg_ratio: float, youngs: float | None = None, bar_centre=(0.0, 0.0, 0.0)) -> Rod:
        """Short leg on -x, half a turn over the bar, long leg on +x with
        long/short = leg_ratio; straight rest shape (cable_hanging.cpp:18-47)."""
        mat = self.material(youngs)
        wrap = self.bar_radius + mat.radius
        arc = PI * wrap
        legs = self.length - arc
        long_leg = legs * leg_ratio / (1.0 + leg_ratio)
        short_leg = legs - long_leg
        top = bar_centre[2] + self.bar_height
        h = self.length / self.segments
        pts = []
        for i in range(self.segments + 1):
            s = h * i
            if s < short_leg:
                p = (-wrap, 0.0, top - (short_leg - s))
            elif s < short_leg + arc:
                phi = (s - short_leg) / wrap
                p = (-wrap * math.cos(phi), 0.0, top + wrap * math.sin(phi))
            else:
                p = (wrap, 0.0, top - (s - short_leg - arc))
            pts.append((bar_centre[0] + p[0], bar_centre[1] + p[1], 0.0 + p[2]))
        rod = make_rod_from_centerline(pts, mat, (0.0, 1.0, 0.0))
        rod.bRestD[:] = 0.0
        return rod

    def world(self, friction: float, bar_centre=(0.0, 0.0, 0.0)) -> CollisionWorld:
        axis = (bar_centre[0], bar_centre[1], bar_centre[2] + self.bar_height)
        w = CollisionWorld()
        w.primitives.append(Primitive.capsule((axis[0], axis[1] - self.bar_half_length, axis[2]),
                                              (axis[0], axis[1] + self.bar_half_length, axis[2]), self.bar_radius,
                                              friction))
        w.primitives.append(Primitive.plane((0.0, 0.0, bar_centre[2]), (0.0, 0.0, 1.0), friction))
        return w

    def params(self) -> SolverParams:
        return SolverParams(dt=self.dt, substeps=self.substeps, iterations=self.iterations, gravity=(0.0, 0.0, -9.81))

    def steps(self) -> int:
        return int(round(self.duration / self.dt))

    def float_motion_margin(self) -> float:
        """Gravity's per-substep displacement over the FP32 spacing at the
        scene's highest coordinate (HangingCable::floatMotionMargin,
        cable_hanging.cpp:75-81). Below about 1 a resting FP32 cable cannot
        start to slide."""
        h = self.dt / self.substeps
        highest = np.float32(self.bar_height + self.bar_radius + 0.05)
        spacing = float(np.nextafter(highest, np.float32(2.0)) - highest)
        return 9.81 * h * h / spacing

    @staticmethod
    def capstan_friction(leg_ratio: float) -> float:
        return math.log(leg_ratio) / PI

    SLIP_TOLERANCE = 0.05

    def held_from_rise(self, max_rise: float) -> bool:
        """HangOutcome::held: the short end never rose more than 5 cm
        (cable_hanging.h:68-77)."""
        return max_rise <= self.SLIP_TOLERANCE

    def held_from_final(self, cable: np.ndarray, bar_centre=(0.0, 0.0, 0.0)) -> bool:
        """Verdict from the final shape: the cable still hangs on the bar when
        its highest particle is within 5 cm of the bar's top."""
        top = bar_centre[2] + self.bar_height + self.bar_radius
        return float(np.max(np.asarray(cable)[:, 2])) >= top - self.SLIP_TOLERANCE
