"""GPU parity on seeded random scenes: curved rods of random length and
material, clamped or free, a floor with random spheres, capsules and boxes,
self-collision on or off (power-of-two and other hash table sizes), random
substeps, iterations, contact intervals, damping and per-rod scales, through
PLAIN and TILED at several tile shapes. Each rod's final state must equal the
FP32 oracle's bit for bit (the reference's step, src/core/solver.cpp, stepped
per rod in the GPU's colour order). The scenes cross paths the fixed
scenarios take one at a time: world contacts kept in slots or formed in the
sweep, far tests of every primitive type, self contacts on candidate lists.
Half of the seeds run again as FP64 batches against the FP64 oracle."""
import numpy as np
import pytest

import ropephysics_amd as rp
from helpers import compare_bitwise, compliance_scale_f32, oracle_params, oracle_state, oracle_twin, oracle_world
from test_gpu_fp64 import FP64_TILE_CONFIGS, diff_words

pytestmark = pytest.mark.gpu

SEEDS = range(24)


def random_scene(seed: int):
    rng = np.random.default_rng(7000 + seed)
    n_rods = int(rng.integers(1, 5))
    mat = rp.RodMaterial(float(10 ** rng.uniform(6.0, 7.5)), 0.35, float(rng.uniform(800.0, 1500.0)),
                         float(rng.uniform(3e-3, 6e-3)))
    self_collision = bool(rng.integers(0, 2))
    rods = []
    for r in range(n_rods):
        n = int(rng.integers(8, 48))
        coiled = (self_collision and r == 0) or rng.uniform() < 0.3
        if coiled:
            # A coil whose turns nearly touch: self contacts from the start.
            turns = rng.uniform(1.5, 3.0)
            n = int(turns * rng.integers(12, 17))
        t = np.linspace(0.0, 1.0, n + 1)
        if coiled:
            rc = rng.uniform(0.02, 0.04)
            pitch = rng.uniform(1.3, 1.95) * mat.radius
            a = 2 * np.pi * turns * t
            pts = np.stack([0.3 + rc * np.cos(a), 0.06 * r + rc * np.sin(a), 0.01 + pitch * turns * t], -1)
        else:
            length = rng.uniform(0.3, 0.8)
            pts = np.stack([length * t, np.full_like(t, 0.06 * r), np.full_like(t, rng.uniform(0.01, 0.05))], -1)
            for axis in range(3):
                pts[:, axis] += rng.uniform(0.0, 0.04) * np.sin(2 * np.pi * rng.uniform(0.5, 3.0) * t + rng.uniform(0, 6.3))
        rod = rp.make_rod_from_centerline(pts, mat, (0.0, 0.0, 1.0))
        if coiled:
            # Both ends pinned: the turns keep pressing on each other.
            rod.pin_particle(0)
            rod.pin_particle(rod.num_segments)
        elif rng.uniform() < 0.5:
            rp.clamp_root_exact(rod)
        rods.append(rod)
    prims = [rp.Primitive.plane((0.0, 0.0, float(rng.uniform(-0.02, 0.0))), (0.0, 0.0, 1.0),
                                float(rng.uniform(0.1, 0.9)))]
    for kind in rng.permutation(["sphere", "capsule", "box"])[: int(rng.integers(0, 4))]:
        c = (float(rng.uniform(0.1, 0.6)), float(rng.uniform(-0.05, 0.2)), float(rng.uniform(-0.06, 0.0)))
        mu = float(rng.uniform(0.2, 0.8))
        if kind == "sphere":
            prims.append(rp.Primitive.sphere(c, float(rng.uniform(0.02, 0.08)), mu))
        elif kind == "capsule":
            prims.append(rp.Primitive.capsule(c, (c[0] + 0.05, c[1] + 0.1, c[2]), float(rng.uniform(0.005, 0.02)), mu))
        else:
            ang = float(rng.uniform(0.0, 1.0))
            prims.append(rp.Primitive.box(c, (0.04, 0.03, 0.02), (np.cos(ang), 0.0, 0.0, np.sin(ang)), mu))
    world = rp.CollisionWorld(primitives=prims, self_collision=self_collision,
                              self_friction=float(rng.uniform(0.0, 0.5)),
                              hash_table_size=int(rng.choice([64, 97, 256])), self_capacity=512)
    params = rp.SolverParams(dt=1e-3, substeps=int(rng.choice([4, 8, 12])), iterations=int(rng.integers(1, 4)),
                             gravity=(0.0, 0.0, -9.81), linear_damping=float(rng.choice([0.0, 0.5])),
                             angular_damping=float(rng.choice([0.0, 0.5])), contact_interval=int(rng.integers(1, 4)))
    scales = (rng.uniform(0.5, 2.0, n_rods), rng.uniform(0.5, 1.5, n_rods))
    shape = int(rng.choice([0, 9, 10, 12, 13]))
    return rods, world, params, scales, shape


def run_scene(seed: int, strategy: str, steps: int = 60, precision: int = 32):
    rods, world, params, (ys, fs), shape = random_scene(seed)
    if precision == 64 and shape not in FP64_TILE_CONFIGS:
        shape = 0
    try:
        b = rp.Batch(rods, strategy=strategy, world=world, tile_config=shape if strategy == "tiled" else 0,
                     precision=precision)
    except rp.RopePhysicsError:
        b = rp.Batch(rods, strategy=strategy, world=world, precision=precision)  # the shape does not fit
    b.set_material_scales(ys, fs)
    b.step(params, steps)
    b.synchronize()
    ow = oracle_world(world)
    diffs = {"x": 0, "v": 0, "q": 0, "omega": 0}
    for i, r in enumerate(rods):
        o = oracle_twin(r, precision)
        if precision == 32:
            op = oracle_params(params, compliance_scale=compliance_scale_f32(ys[i]),
                               friction_scale=float(np.float32(fs[i])))
        else:
            op = oracle_params(params, compliance_scale=1.0 / float(ys[i]), friction_scale=float(fs[i]))
        o.step(op, ow, steps)
        d = (compare_bitwise if precision == 32 else diff_words)(b.download(i), oracle_state(o))
        for k in diffs:
            diffs[k] += d[k]
    moved = max(float(np.max(np.abs(b.download(i)["x"] - r.x))) for i, r in enumerate(rods))
    b.destroy()
    return diffs, moved


@pytest.mark.parametrize("strategy", ["plain", "tiled"])
@pytest.mark.parametrize("seed", SEEDS)
def test_random_scene_bitwise(seed, strategy):
    diffs, moved = run_scene(seed, strategy)
    assert all(v == 0 for v in diffs.values()), (seed, strategy, diffs)
    assert moved > 1e-4  # the rods really moved


@pytest.mark.parametrize("strategy", ["plain", "tiled"])
@pytest.mark.parametrize("seed", SEEDS[::2])
def test_random_scene_bitwise_fp64(seed, strategy):
    diffs, moved = run_scene(seed, strategy, precision=64)
    assert all(v == 0 for v in diffs.values()), (seed, strategy, diffs)
    assert moved > 1e-4
