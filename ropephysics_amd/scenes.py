"""Synthetic workloads of the benchmark configurations, built on the host.

Each builder returns plain host data (rods, worlds, per-rod scales, step
parameters); nothing here touches the device. Random draws use numpy's PCG64
generator with the stated seed.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .rod import CollisionWorld, Primitive, Rod, RodMaterial, SolverParams, clamp_root_exact, make_straight_rod

REFERENCE_MATERIAL = RodMaterial(1e7, 0.35, 1000.0, 5e-3)


@dataclass
class Workload:
    name: str
    prototype: Rod | None            # set when every rod shares one topology
    rods: list | None                # set for ragged batches
    num_rods: int
    x0: np.ndarray | None            # packed initial positions (None: the rods' own)
    world: object                    # CollisionWorld or list of them
    world_of_rod: np.ndarray | None
    youngs_scale: np.ndarray | None
    friction_scale: np.ndarray | None
    params: SolverParams
    steps: int
    segments_total: int              # sum of real segments over rods
    info: dict = field(default_factory=dict)

    @property
    def segment_substeps(self) -> float:
        return float(self.segments_total) * self.params.substeps * self.steps


def cantilever_tip_load():
    """Single cantilever, FP64 CPU workload: 1 m, 64 segments, r = 5 mm,
    E = 1e7 Pa, Poisson 0.5, rho = 1000, clamped at x = 0, no gravity, tip load
    0.01 N in -z ramped over 0.5 s, dt = 1e-3, 20 substeps, 2 s."""
    mat = RodMaterial(1e7, 0.5, 1000.0, 5e-3)
    rod = make_straight_rod(64, 1.0, mat, (0, 0, 0), (1, 0, 0))
    clamp_root_exact(rod)
    params = SolverParams(dt=1e-3, substeps=20, iterations=1, gravity=(0.0, 0.0, 0.0))

    def tip_force(step: int) -> float:
        t = step * 1e-3
        return -0.01 * min(t / 0.5, 1.0)

    return rod, params, 2000, tip_force


def uniform_floor_batch(num_rods: int = 4096, segments: int = 64, seed: int = 20260917, height: float = 0.1,
                        steps: int = 1000) -> Workload:
    """Uniform batch: 4096 rods x 64 segments, 0.5 m, root clamped at
    `height` above a floor plane z = 0 with Coulomb friction, gravity. Per rod:
    Young's modulus scale 0.4 * 10^U(0,1) (log-uniform in [0.4, 4] x 1e7 Pa) and
    friction mu ~ U[0.3, 0.9] (plane friction 1 times the per-rod scale). Free
    particles start with N(0, 1 mm) jitter about the horizontal pose, which is
    also the rest pose. dt = 1e-3, 10 substeps, 1000 steps."""
    rng = np.random.default_rng(seed)
    proto = make_straight_rod(segments, 0.5, REFERENCE_MATERIAL, (0.0, 0.0, height), (1.0, 0.0, 0.0))
    clamp_root_exact(proto)
    youngs = 0.4 * 10.0 ** rng.uniform(0.0, 1.0, num_rods)
    mu = rng.uniform(0.3, 0.9, num_rods)
    nP = proto.num_particles
    jitter = rng.normal(0.0, 1e-3, (num_rods, nP, 3))
    jitter[:, proto.invMass == 0.0, :] = 0.0
    x0 = (proto.x[None, :, :] + jitter).reshape(-1, 3)
    world = CollisionWorld(primitives=[Primitive.plane((0.0, 0.0, 0.0), (0.0, 0.0, 1.0), 1.0)])
    params = SolverParams(dt=1e-3, substeps=10, iterations=1, gravity=(0.0, 0.0, -9.81))
    return Workload("uniform-4096x64-floor" if num_rods == 4096 and segments == 64 else
                    f"uniform-{num_rods}x{segments}-floor", proto, None, num_rods, x0, world, None, youngs, mu,
                    params, steps, num_rods * segments,
                    {"rods": num_rods, "segments": segments, "length_m": 0.5, "height_m": height, "seed": seed,
                     "dt": 1e-3, "substeps": 10, "steps": steps, "contacts": "floor plane, Coulomb friction"})


def routing_rollouts(num_rods: int = 1024, seed: int = 20260917, steps: int | None = None) -> Workload:
    """Ragged batch of routing rollouts: cables of 48-160 segments (uniform
    integer), 1.2 m, connector pinned, the other end driven along a seeded
    8-waypoint smoothstep (piecewise-cubic) path; floor, two capsule pegs
    (r = 15 mm), one box clip; friction 0.5 x U[0.6, 1.4]; stiffness
    0.4-4x log-uniform; self-collision on; dt = 1e-3, 10 substeps, 6 s.
    Paths: the task's base waypoints, or for a quarter of the cables (drawn
    from the seed) its wrong-side waypoints (over peg A), each waypoint
    jittered by N(0, 15 mm)."""
    from .apps import RoutingTask

    task = RoutingTask()
    rng = np.random.default_rng(seed)
    n = rng.integers(48, 161, num_rods)
    youngs = 0.4 * 10.0 ** rng.uniform(0.0, 1.0, num_rods)
    fric = rng.uniform(0.6, 1.4, num_rods)
    wrong = rng.uniform(0.0, 1.0, num_rods) < 0.25
    base = np.where(wrong[:, None, None], task.wrong_side_waypoints()[None], task.base_waypoints()[None])
    wps = base + rng.normal(0.0, 0.015, (num_rods, 8, 2))
    cache = {}
    rods = []
    for k in n:
        k = int(k)
        if k not in cache:
            cache[k] = task.build(k)
        rods.append(cache[k])
    paths = [task.path(int(n[r]), wps[r]) for r in range(num_rods)]
    st = task.steps() if steps is None else steps
    w = Workload("routing-1024-ragged" if num_rods == 1024 else f"routing-{num_rods}-ragged", None, rods, num_rods,
                 None, task.world(), None, youngs, fric, task.params(), st, int(n.sum()),
                 {"rods": num_rods, "segments": "48-160 uniform", "segments_total": int(n.sum()), "length_m": 1.2,
                  "seed": seed, "dt": 1e-3, "substeps": 10, "steps": st, "self_collision": True,
                  "contacts": "floor, 2 capsule pegs r=15 mm, box clip"})
    w.info["paths"] = paths
    w.info["task"] = task
    return w


def drape_sweep(seed: int = 20260917, steps: int | None = None, ratios: int = 60, frictions: int = 30) -> Workload:
    """Hold/slide sweep: 3600 cables (60 leg ratios 1:1-4:1 x 30 friction
    values 0.05-0.5 x stiffness E and 3E, E = 5e6 Pa), 120 segments, 1 m,
    r = 3 mm, draped over a horizontal capsule bar (r = 10 mm); dt = 5e-4,
    10 substeps, 8 s after release."""
    from .apps import HangingCable

    task = HangingCable()
    ratio_vals = np.linspace(1.0, 4.0, ratios)
    mu_vals = np.linspace(0.05, 0.5, frictions)
    protos = [task.build(float(r)) for r in ratio_vals]
    rods, ys, fs, meta = [], [], [], []
    for stiff in (1.0, 3.0):
        for i, r in enumerate(ratio_vals):
            for mu in mu_vals:
                rods.append(protos[i])
                ys.append(stiff)
                fs.append(mu)
                meta.append((float(r), float(mu), stiff))
    world = task.world(1.0)
    st = task.steps() if steps is None else steps
    w = Workload("drape-3600" if ratios * frictions == 1800 else f"drape-{len(rods)}", None, rods, len(rods), None,
                 world, None, np.array(ys), np.array(fs), task.params(), st, len(rods) * task.segments,
                 {"rods": len(rods), "segments": task.segments, "length_m": task.length, "dt": task.dt,
                  "substeps": task.substeps, "steps": st, "bar_radius_m": task.bar_radius,
                  "contacts": "capsule bar and floor"})
    w.info["task"] = task
    w.info["meta"] = meta
    return w
