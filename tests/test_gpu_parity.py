"""GPU parity: the HIP path through the C ABI against the CPU oracle.

The bar, per scenario:
  * bit for bit equal to the oracle's FP32 instantiation run in the GPU's
    colour order (the arithmetic order is the reference's, contraction off on
    both sides), every field, every step checked;
  * within tolerance of the FP64 oracle: 1e-7 m after one step (the
    reference's own one-step bar, src/gpu/gpu_cases.cpp:285) and within the
    float envelope over long horizons;
  * both strategies, every tile shape, and a rod alone or packed among others
    give identical bits; repeated runs are bitwise identical.
Reference scenarios follow src/gpu/gpu_cases.cpp:126-548.
"""
import numpy as np
import pytest

import oracle
import ropephysics_amd as rp
from helpers import (REFERENCE_MATERIAL, benchmark_params, benchmark_rod, bits, compare_bitwise,
                     compliance_scale_f32, max_pos_err, oracle_params, oracle_state, oracle_twin, oracle_world)

pytestmark = pytest.mark.gpu

STRATEGIES = ["plain", "tiled"]

CONTACT_WORLD = rp.CollisionWorld(primitives=[
    rp.Primitive.plane((0, 0, -0.0051), (0, 0, 1), 0.5),
    rp.Primitive.sphere((0.5, 0, -0.0851), 0.08, 0.3),
])

ALL_PRIMS_WORLD = rp.CollisionWorld(primitives=[
    rp.Primitive.plane((0, 0, -0.0051), (0, 0, 1), 0.5),
    rp.Primitive.sphere((0.5, 0, -0.0851), 0.08, 0.3),
    rp.Primitive.capsule((0.25, -0.1, -0.016), (0.25, 0.1, -0.016), 0.01, 0.4),
    rp.Primitive.box((0.8, 0.0, -0.026), (0.05, 0.05, 0.02), (0.9659258262890683, 0, 0, 0.25881904510252074), 0.6),
])


def run_pair(rods, world, params, steps, strategy, checkpoints=None, scales=None, tile_config=0):
    """Step rods on the GPU and each rod on the FP32 oracle; return the number
    of differing FP32 words per field, summed over rods and checkpoints, plus
    the worst FP64 error at the end."""
    b = rp.Batch(rods, strategy=strategy, world=world, tile_config=tile_config)
    ys, fs = (None, None) if scales is None else scales
    if scales is not None:
        b.set_material_scales(ys, fs)
    checkpoints = checkpoints or [steps]
    twins = []
    for i, r in enumerate(rods):
        cs = compliance_scale_f32(ys[i]) if ys is not None else 1.0
        f = float(np.float32(fs[i])) if fs is not None else 1.0
        twins.append((oracle_twin(r, 32), oracle_twin(r, 64), oracle_params(params, compliance_scale=cs,
                                                                             friction_scale=f)))
    ow = oracle_world(world)
    diffs = {"x": 0, "v": 0, "q": 0, "omega": 0}
    done = 0
    for cp in checkpoints:
        b.step(params, cp - done)
        for o32, o64, op in twins:
            o32.step(op, ow, cp - done)
            o64.step(op, ow, cp - done)
        done = cp
        b.synchronize()
        for i, (o32, _, _) in enumerate(twins):
            d = compare_bitwise(b.download(i), oracle_state(o32))
            for k in diffs:
                diffs[k] += d[k]
    err64 = max(max_pos_err(b.download(i)["x"], o64.get()["x"]) for i, (_, o64, _) in enumerate(twins))
    return diffs, err64, b


def assert_bitwise(diffs):
    assert all(v == 0 for v in diffs.values()), diffs


@pytest.mark.parametrize("strategy", STRATEGIES)
@pytest.mark.parametrize("n", [16, 48, 128])
def test_gravity_cantilever(strategy, n):
    diffs, err64, _ = run_pair([benchmark_rod(n)], None, benchmark_params(), 200, strategy,
                               checkpoints=[1, 10, 50, 200])
    assert_bitwise(diffs)
    assert err64 < 1e-3  # reference: float drift at 200 steps (gpu_cases.cpp:303-304)


@pytest.mark.parametrize("strategy", STRATEGIES)
@pytest.mark.parametrize("n", [16, 48, 128])
def test_one_step_within_1e7_of_fp64(strategy, n):
    _, err64, _ = run_pair([benchmark_rod(n)], None, benchmark_params(), 1, strategy)
    assert err64 < 1e-7


@pytest.mark.parametrize("strategy", STRATEGIES)
def test_loaded_driven_rod(strategy):
    """Tip force, mid torque, a pinned root moving kinematically and the root's
    ghost frame twisted a little every step through set_loads
    (gpu_cases.cpp:205-228)."""
    n = 48
    rod = benchmark_rod(n)
    rod.extForce[-1] = (0.0, 0.02, 0.05)
    rod.extTorque[n // 2] = (2e-3, 0.0, -1e-3)
    rod.kinVel[0] = (0.0, 0.05, 0.1)
    ghost = rod.num_frames - 1
    rest = tuple(rod.q[ghost])
    params = benchmark_params()
    b = rp.Batch([rod], strategy=strategy)
    o32 = oracle_twin(rod, 32)
    op = oracle_params(params)
    for s in range(60):
        qd = rp.rod._qnormalize(rp.rod._qmul(rp.quat_axis_angle((1, 0, 0), 0.5e-3 * (s + 1)), rest))
        rod.q[ghost] = qd
        b.set_loads(0, rod)
        a = o32.get()
        a["q"][ghost] = np.float32(np.array(qd))
        o32.set({"q": a["q"]})
        b.step(params, 1)
        o32.step(op, None, 1)
    b.synchronize()
    assert_bitwise(compare_bitwise(b.download(0), oracle_state(o32)))
    # The loads matter: the rod is far from the unloaded one.
    plain = oracle_twin(benchmark_rod(n), 32)
    plain.step(op, None, 60)
    assert max_pos_err(b.download(0)["x"], plain.get()["x"]) > 1e-3


@pytest.mark.parametrize("strategy", STRATEGIES)
def test_contact_floor_and_sphere(strategy):
    diffs, _, _ = run_pair([benchmark_rod(48)], CONTACT_WORLD, benchmark_params(), 200, strategy,
                           checkpoints=[1, 50, 200])
    assert_bitwise(diffs)
    o = oracle_twin(benchmark_rod(48), 32)
    o.step(oracle_params(benchmark_params()), oracle_world(CONTACT_WORLD), 50)
    assert o.contact_counts()[0] >= 5  # the scene really is in contact


@pytest.mark.parametrize("strategy", STRATEGIES)
def test_all_primitive_types(strategy):
    rods = [benchmark_rod(40), benchmark_rod(64)]
    diffs, _, _ = run_pair(rods, ALL_PRIMS_WORLD, benchmark_params(iterations=2), 150, strategy,
                           checkpoints=[20, 150])
    assert_bitwise(diffs)


@pytest.mark.parametrize("strategy", STRATEGIES)
@pytest.mark.parametrize("iterations", [1, 2])
def test_contacts_kept_between_regenerations(strategy, iterations):
    """Contact interval 3 (src/core/solver.cpp:282-284: contacts regenerated at every
    third substep, their multipliers reset at the others): the world-contact
    slots live across substeps, with one and with two iterations. With one
    iteration and an interval of 1 the TILED path keeps no slots at all; this
    case runs the one that keeps them."""
    params = rp.SolverParams(dt=1e-3, substeps=8, iterations=iterations, gravity=(0.0, 0.0, -9.81),
                             contact_interval=3)
    diffs, _, _ = run_pair([benchmark_rod(40), benchmark_rod(64)], ALL_PRIMS_WORLD, params, 120, strategy,
                           checkpoints=[10, 120])
    assert_bitwise(diffs)


@pytest.mark.parametrize("strategy", STRATEGIES)
def test_per_rod_stiffness_and_friction_scales(strategy):
    """setMaterialScales: stiffness on a swinging cantilever, friction on an
    incline that holds at mu = 0.5 and slides at half (gpu_cases.cpp:316-419)."""
    rods = [benchmark_rod(16), benchmark_rod(16)]
    diffs, _, b = run_pair(rods, None, benchmark_params(), 20, strategy, scales=([1.0, 3.0], None))
    assert_bitwise(diffs)
    x = b.get_state()["x"]
    assert max_pos_err(x[:17], x[17:]) > 1e-4
    # Next to the clamp, where bending stiffness holds the rod up, the stiffer
    # rod (Young's scale 3) sags less; a scale applied inverted fails this.
    assert x[17 + 1, 2] > x[1, 2] and x[17 + 2, 2] > x[2, 2]
    tilt = np.deg2rad(20.0)
    normal = (np.sin(tilt), 0.0, np.cos(tilt))
    incline = rp.CollisionWorld(primitives=[rp.Primitive.plane((0, 0, 0), normal, 0.5)])
    start = tuple(np.array(normal) * (REFERENCE_MATERIAL.radius + 1e-4) + np.array([0, -0.5, 0]))
    rod = rp.make_straight_rod(16, 1.0, REFERENCE_MATERIAL, start, (0, 1, 0))
    diffs, _, b = run_pair([rod, rod], incline, benchmark_params(), 150, strategy, scales=([1.0, 1.0], [1.0, 0.5]))
    assert_bitwise(diffs)
    x = b.get_state()["x"]
    assert max_pos_err(x[:17], x[17:]) > 1e-3


def helix_rope():
    n_turn, turns = 20, 3
    t = np.arange(n_turn * turns + 1) / n_turn * 2 * np.pi
    pts = np.stack([0.03 * np.cos(t), 0.03 * np.sin(t), 0.02 + 0.0095 * t / (2 * np.pi)], axis=-1)
    return rp.make_rod_from_centerline(pts, rp.RodMaterial(2e6, 0.35, 1000.0, 5e-3), (0, 0, 1))


# The coil holds more contacts than the reference's default pool
# (max(16, segments / 2)); a larger pool keeps every contact, as the CPU does.
SELF_WORLD = rp.CollisionWorld(primitives=[rp.Primitive.plane((0, 0, 0), (0, 0, 1), 0.6)], self_collision=True,
                               self_friction=0.3, hash_table_size=256, self_capacity=512)
SELF_PARAMS = rp.SolverParams(dt=1e-3, substeps=16, iterations=4, gravity=(0, 0, -9.81), linear_damping=1.0,
                              angular_damping=1.0)


@pytest.mark.parametrize("strategy", STRATEGIES)
def test_self_collision_coil(strategy):
    rope = helix_rope()
    diffs, _, b = run_pair([rope], SELF_WORLD, SELF_PARAMS, 40, strategy, checkpoints=[1, 5, 40])
    assert_bitwise(diffs)
    assert b.self_contact_overflow() == 0
    no_self = rp.CollisionWorld(primitives=SELF_WORLD.primitives)
    o = oracle_twin(rope, 32)
    o.step(oracle_params(SELF_PARAMS), oracle_world(no_self), 40)
    assert max_pos_err(b.download(0)["x"], o.get()["x"]) > 1e-3  # self-collision mattered


def hairpin(gap=0.03, length=0.3, spacing=0.01):
    """Two parallel strands `gap` apart joined by a half circle, the first
    strand moving towards the second at 0.6 m/s and the second towards the
    first: pairs far outside any candidate list at the start come into
    contact within about 20 steps."""
    n = int(round(length / spacing))
    a = [(i * spacing, 0.0, 0.05) for i in range(n + 1)]
    arc = [(length + 0.5 * gap * np.sin(th), 0.5 * gap - 0.5 * gap * np.cos(th), 0.05)
           for th in np.linspace(0.0, np.pi, 6)[1:-1]]
    b = [(length - i * spacing, gap, 0.05) for i in range(n + 1)]
    pts = np.array(a + arc + b)
    rod = rp.make_rod_from_centerline(pts, rp.RodMaterial(2e6, 0.35, 1000.0, 5e-3), (0, 0, 1))
    rod.v[:] = 0.0
    rod.v[: n + 1, 1] = 0.6
    rod.v[-(n + 1):, 1] = -0.6
    return rod


@pytest.mark.parametrize("strategy", STRATEGIES)
def test_self_collision_candidate_lists_slow_fast_and_closing(strategy):
    """The TILED broadphase keeps each segment's candidates (its later
    segments within the cull radius plus one rod radius) while the rod's
    summed displacement bound stays below an eighth of that skin, and scans
    again once it does not. Three rods against the oracle, bit for bit: a
    helix thrown at 4 m/s (a rescan every few substeps), the same helix at rest
    (lists kept throughout) and a hairpin closing on itself (its contacts
    come from segments that were nobody's candidates at the start)."""
    fast, slow, pin = helix_rope(), helix_rope(), hairpin()
    fast.v[:, 0] = 4.0
    world = rp.CollisionWorld(primitives=[rp.Primitive.plane((0, 0, -0.5), (0, 0, 1), 0.6)], self_collision=True,
                              self_friction=0.3, hash_table_size=256, self_capacity=600)
    rods = [fast, slow, pin]
    diffs, _, b = run_pair(rods, world, SELF_PARAMS, 60, strategy, checkpoints=[1, 10, 20, 30, 60])
    assert_bitwise(diffs)
    assert b.self_contact_overflow() == 0
    # The hairpin's strands met: without self-collision they pass through.
    o = oracle_twin(pin, 32)
    o.step(oracle_params(SELF_PARAMS), oracle_world(rp.CollisionWorld(primitives=world.primitives)), 60)
    assert max_pos_err(b.download(2)["x"], o.get()["x"]) > 1e-3


def test_self_collision_overflow_is_counted():
    """Contacts beyond the per-rod pool are counted, never silently dropped
    (gpu_solver.cu:311-320)."""
    rope = helix_rope()
    b = rp.Batch([rope], strategy="plain", world=SELF_WORLD)
    b.step(SELF_PARAMS, 2)
    assert b.self_contact_overflow() == 0
    tiny = rp.CollisionWorld(primitives=SELF_WORLD.primitives, self_collision=True, self_friction=0.3,
                             hash_table_size=256, self_capacity=1)
    for strategy in STRATEGIES:
        tb = rp.Batch([rope], strategy=strategy, world=tiny)
        tb.step(SELF_PARAMS, 2)
        assert tb.self_contact_overflow() > 0


def test_self_capacity_per_segment():
    """A pool of ceil(k * segments) slots (at least 16) is the pool of that
    many fixed slots: same bits, same overflow count; out-of-range values are
    refused."""
    from ropephysics_amd import _abi

    rope = helix_rope()  # 60 segments
    base = dict(primitives=SELF_WORLD.primitives, self_collision=True, self_friction=0.3, hash_table_size=256)
    for fixed, per_seg in ((510, 8.5), (16, 0.2)):
        runs = []
        for world in (rp.CollisionWorld(**base, self_capacity=fixed),
                      rp.CollisionWorld(**base, self_capacity_per_segment=per_seg)):
            b = rp.Batch([rope], strategy="tiled", world=world)
            b.step(SELF_PARAMS, 20)
            runs.append((b.get_state(), b.counters()))
        (sa, ca), (sb, cb) = runs
        for k in sa:
            assert np.array_equal(bits(sa[k]), bits(sb[k])), (fixed, k)
        assert ca["self_contact_overflow"] == cb["self_contact_overflow"]
        assert ca["lds_bytes_per_workgroup"] == cb["lds_bytes_per_workgroup"]
        assert (ca["self_contact_overflow"] > 0) == (fixed == 16)
    with pytest.raises(_abi.RopePhysicsError) as e:
        rp.Batch([rope], world=rp.CollisionWorld(**base, self_capacity_per_segment=-1.0))
    assert e.value.code == _abi.RP_ERR_INVALID


def pinned_coil(turns=5, per_turn=16, pitch=0.008, radius=0.025):
    """A helix whose pitch (8 mm) is below the rope's diameter (10 mm), both
    ends pinned: the turns keep pressing on each other."""
    t = np.arange(per_turn * turns + 1) / per_turn * 2 * np.pi
    pts = np.stack([radius * np.cos(t), radius * np.sin(t), 0.03 + pitch * t / (2 * np.pi)], axis=-1)
    rod = rp.make_rod_from_centerline(pts, rp.RodMaterial(2e6, 0.35, 1000.0, 5e-3), (0, 0, 1))
    rod.pin_particle(0)
    rod.pin_particle(rod.num_segments)
    return rod


@pytest.mark.parametrize("strategy", STRATEGIES)
@pytest.mark.parametrize("table", [16, 64, 256])
def test_self_collision_pinned_coil_hash_tables(strategy, table):
    """Dense, changing self-contact sets. Small tables make neighbour cells
    share buckets, so a bucket is visited more than once and its contacts
    appear more than once in the list, as on the CPU. The pool is sized so that nothing overflows (the CPU list is
    unbounded). The batch also holds a free helix and a rod without contacts
    of its own."""
    world = rp.CollisionWorld(primitives=[rp.Primitive.plane((0, 0, 0), (0, 0, 1), 0.6)], self_collision=True,
                              self_friction=0.3, hash_table_size=table, self_capacity=700)
    rods = [pinned_coil(), helix_rope(), benchmark_rod(24)]
    diffs, _, b = run_pair(rods, world, SELF_PARAMS, 30, strategy, checkpoints=[1, 3, 30])
    assert b.self_contact_overflow() == 0
    assert_bitwise(diffs)
    o = oracle_twin(rods[0], 32)
    op, ow = oracle_params(SELF_PARAMS), oracle_world(world)
    seen = 0
    for _ in range(3):
        o.step(op, ow, 1)
        seen = max(seen, o.contact_counts()[1])
    assert seen >= 10  # the coil really is in contact


@pytest.mark.parametrize("tile_config", [9, 11, 12, 13])
def test_self_collision_lane_groups_across_wavefronts(tile_config):
    """Self-colliding rods of 61-81 particles at one or two particles per lane,
    packed side by side: lane groups start mid-wavefront and span wavefront
    boundaries, which the group-parallel scans and broadphase must handle.
    Bitwise equal to the oracle, every shape."""
    world = rp.CollisionWorld(primitives=[rp.Primitive.plane((0, 0, 0), (0, 0, 1), 0.6)], self_collision=True,
                              self_friction=0.3, hash_table_size=64, self_capacity=700)
    rods = [benchmark_rod(24), pinned_coil(), helix_rope(), pinned_coil(), benchmark_rod(40)]
    diffs, _, b = run_pair(rods, world, SELF_PARAMS, 12, "tiled", checkpoints=[2, 12], tile_config=tile_config)
    assert b.counters()["tile_config"] == tile_config
    assert b.self_contact_overflow() == 0
    assert_bitwise(diffs)


def test_self_collision_longest_segment_in_every_row():
    """The broadphase's cell is the rod's longest segment plus 2r, found by a
    segmented max over each lane group's lanes within a wavefront (row moves
    of 1-8 lanes, then across 16-lane rows). Coils with one segment three
    times longer than the rest, at positions that land in different rows of a
    wavefront, packed one particle per lane so lane groups start mid-wavefront;
    small hash tables make the contact list depend on the exact cell. Bitwise
    equal to the oracle."""
    world = rp.CollisionWorld(primitives=[rp.Primitive.plane((0, 0, 0), (0, 0, 1), 0.6)], self_collision=True,
                              self_friction=0.3, hash_table_size=64, self_capacity=700)
    base = np.array(pinned_coil().x, dtype=np.float64)
    rods = [benchmark_rod(24)]
    for m in (2, 13, 21, 34, 45, 58, 71):
        pts = base.copy()
        t = pts[m + 1] - pts[m]
        pts[m + 1:] += 2.0 * t  # segment m three times its neighbours' length
        rod = rp.make_rod_from_centerline(pts, rp.RodMaterial(2e6, 0.35, 1000.0, 5e-3), (0, 0, 1))
        rod.pin_particle(0)
        rod.pin_particle(rod.num_segments)
        rods.append(rod)
    diffs, _, b = run_pair(rods, world, SELF_PARAMS, 8, "tiled", checkpoints=[2, 8], tile_config=12)
    assert b.counters()["tile_config"] == 12
    assert b.self_contact_overflow() == 0
    assert_bitwise(diffs)
    o = oracle_twin(rods[4], 32)
    o.step(oracle_params(SELF_PARAMS), oracle_world(world), 2)
    assert o.contact_counts()[1] >= 10  # the coils are in self contact


def _reverse_particles(rod):
    """The same rod with its particles numbered from the other end: stretch
    constraint k joins particles n-1-k and n-2-k."""
    r = rod.copy()
    n = r.num_particles
    for name in ("x", "xPrev", "v", "invMass", "mass", "extForce", "kinVel"):
        setattr(r, name, getattr(r, name)[::-1].copy())
    r.sP0 = (n - 1 - rod.sP0).astype(rod.sP0.dtype)
    r.sP1 = (n - 1 - rod.sP1).astype(rod.sP1.dtype)
    return r


@pytest.mark.parametrize("tile_config", [0, 12])
def test_self_collision_rod_with_relabelled_particles(tile_config):
    """A coil whose segments do not join particles k and k + 1 (particles
    numbered from the other end; the broadphase takes segment ends from the
    constraint table): bitwise equal to the oracle, and, since relabelling
    changes no arithmetic, to the same coil numbered the usual way, particle
    for particle."""
    world = rp.CollisionWorld(primitives=[rp.Primitive.plane((0, 0, 0), (0, 0, 1), 0.6)], self_collision=True,
                              self_friction=0.3, hash_table_size=64, self_capacity=700)
    coil = pinned_coil()
    rods = [_reverse_particles(coil), benchmark_rod(24), coil]
    diffs, _, b = run_pair(rods, world, SELF_PARAMS, 8, "tiled", checkpoints=[2, 8], tile_config=tile_config)
    assert b.self_contact_overflow() == 0
    assert_bitwise(diffs)
    rel, plain = b.download(0), b.download(2)
    assert np.array_equal(bits(rel["x"][::-1]), bits(plain["x"]))
    assert np.array_equal(bits(rel["q"]), bits(plain["q"]))


@pytest.mark.parametrize("strategy", STRATEGIES)
def test_ragged_batch_and_alone_equals_mixed(strategy):
    sizes = [7, 20, 33, 64, 97, 130, 160, 48]
    rods = [benchmark_rod(n) for n in sizes]
    diffs, _, b = run_pair(rods, CONTACT_WORLD, benchmark_params(), 30, strategy)
    assert_bitwise(diffs)
    mixed = b.get_state()
    for i in (0, 3, 6):
        alone = rp.Batch([rods[i]], strategy=strategy, world=CONTACT_WORLD)
        alone.step(benchmark_params(), 30)
        s = alone.download(0)
        for k in ("x", "v", "q", "omega"):
            part = b.rod_slice(mixed[k], i, "particle" if k in ("x", "v") else "frame")
            assert np.array_equal(bits(part), bits(s[k]))


def test_every_tile_configuration_gives_the_same_bits():
    rods = [benchmark_rod(n) for n in (12, 40, 64, 90)]
    ref = None
    for cfg in rp.tile_configs():
        try:
            b = rp.Batch(rods, strategy="tiled", world=CONTACT_WORLD, tile_config=cfg)
            b.step(benchmark_params(), 12)
        except rp.RopePhysicsError as e:
            assert e.code == -4  # this shape cannot hold these rods
            continue
        s = b.get_state()
        if ref is None:
            ref = s
        else:
            for k in s:
                assert np.array_equal(bits(s[k]), bits(ref[k])), (cfg, k)
    assert ref is not None


def test_plain_equals_tiled_and_runs_repeat_bitwise():
    rods = [benchmark_rod(n) for n in (32, 128)] * 8
    outs = []
    for strategy in ("plain", "tiled", "tiled", "tiled"):
        b = rp.Batch(rods, strategy=strategy)
        b.step(benchmark_params(), 100)
        outs.append(b.get_state())
    for o in outs[1:]:
        for k in o:
            assert np.array_equal(bits(o[k]), bits(outs[0][k]))


def test_launch_length_does_not_change_the_bits():
    """TILED runs a call's steps in one launch (up to its steps-per-launch
    bound): 60 steps in one launch, in calls of 7 and in calls of one step give
    the same bits, with world contacts and self-collision on."""
    rods = [helix_rope(), benchmark_rod(24), benchmark_rod(101)]
    outs = []
    for calls in ([60], [7] * 8 + [4], [1] * 60):
        b = rp.Batch(rods, strategy="tiled", world=SELF_WORLD)
        for k in calls:
            b.step(SELF_PARAMS, k)
        assert b.counters()["last_step_launches"] == 1
        outs.append(b.get_state())
    for o in outs[1:]:
        for k in o:
            assert np.array_equal(bits(o[k]), bits(outs[0][k])), k


def test_calls_longer_than_one_launch():
    """A call of 8200 steps runs as two launches (the bound is 8192 steps per
    launch) and equals 82 calls of 100 steps, bit for bit."""
    rods = [benchmark_rod(16), benchmark_rod(9)]
    params = benchmark_params(substeps=2)
    one = rp.Batch(rods, strategy="tiled", world=CONTACT_WORLD)
    one.step(params, 8200)
    assert one.counters()["last_step_launches"] == 2
    many = rp.Batch(rods, strategy="tiled", world=CONTACT_WORLD)
    for _ in range(82):
        many.step(params, 100)
    a, b = one.get_state(), many.get_state()
    for k in a:
        assert np.array_equal(bits(a[k]), bits(b[k])), k
    assert np.all(np.isfinite(a["x"]))


def test_paired_workgroups_give_the_same_bits():
    """A TILED launch with between one and two workgroups per compute unit,
    where workgroups b and b + units trade issue priority by step count at
    each of the first 64 steps and every 64 after, and the tiles are placed by
    estimated cost (the costliest alone on a unit, the rest paired costliest
    with cheapest, so the launch order differs from the packing order): routing
    cables of 130-160 segments (self-collision, world contacts, gripper paths),
    one per 256-lane workgroup whatever the LDS budget, 1.5 per compute unit,
    300 steps in one launch, equal PLAIN bit for bit."""
    from ropephysics_amd import _abi, scenes

    units = _abi.device_info()["compute_units"]
    w = scenes.routing_rollouts(num_rods=8 * units, steps=300)
    pick = [i for i, r in enumerate(w.rods) if r.num_segments >= 130][: 3 * units // 2]
    assert len(pick) == 3 * units // 2
    rods = [w.rods[i] for i in pick]
    outs = []
    for strategy in ("tiled", "plain"):
        b = rp.Batch(rods, strategy=strategy, world=w.world, tile_config=12 if strategy == "tiled" else 0)
        b.set_material_scales(w.youngs_scale[pick], w.friction_scale[pick])
        b.set_paths(0, [w.info["paths"][i] for i in pick])
        b.step(w.params, w.steps)
        b.synchronize()
        if strategy == "tiled":
            c = b.counters()
            assert c["tile_config_name"] == "tp1-tc1-w4" and c["last_step_launches"] == 1
            assert c["workgroups"] == len(rods) and units < c["workgroups"] <= 2 * units
        outs.append(b.get_state())
        b.destroy()
    for k in outs[0]:
        assert np.array_equal(bits(outs[0][k]), bits(outs[1][k])), k
    assert np.all(np.isfinite(outs[0]["x"]))


def test_full_size_uniform_batch_sampled_rods_bitwise():
    """The 4096 x 64 floor workload at full size for 30 steps: a sample of rods
    against the FP32 oracle, bit for bit; every rod finite and above the floor."""
    from ropephysics_amd import scenes

    w = scenes.uniform_floor_batch(steps=30)
    b = rp.Batch(w.prototype, w.num_rods, strategy="tiled", world=w.world)
    b.set_material_scales(w.youngs_scale, w.friction_scale)
    R = w.num_rods
    b.set_state(0, R, x=w.x0, v=np.zeros_like(w.x0), q=np.tile(w.prototype.q, (R, 1)),
                omega=np.tile(w.prototype.omega, (R, 1)))
    b.step(w.params, 30)
    s = b.get_state()
    assert np.all(np.isfinite(s["x"]))
    assert s["x"][:, 2].min() > -w.prototype.material.radius
    nP = w.prototype.num_particles
    ow = oracle_world(w.world)
    for r in (0, 1, 777, 2048, R - 1):
        rod = w.prototype.copy()
        rod.x = w.x0[r * nP:(r + 1) * nP].copy()
        rod.xPrev = rod.x.copy()
        o = oracle_twin(rod, 32)
        o.step(oracle_params(w.params, compliance_scale=compliance_scale_f32(w.youngs_scale[r]),
                             friction_scale=float(np.float32(w.friction_scale[r]))), ow, 30)
        got = b.download(r)
        assert_bitwise(compare_bitwise(got, oracle_state(o)))


def test_extreme_rod_sizes():
    """The smallest rods (1, 2 and 3 segments) and the longest clamped chain
    the TILED path takes (rp_device_query's max_tiled_segments, as the
    reference reports maxFusedSegments), packed together: bitwise equal to the
    oracle under both strategies. One segment more does not fit any tile shape:
    AUTO runs it PLAIN (still bitwise equal) and TILED refuses it at creation."""
    from ropephysics_amd import _abi

    nmax = rp.device_info()["max_tiled_segments"]
    assert nmax >= 255
    rods = [benchmark_rod(1), benchmark_rod(2), benchmark_rod(3), benchmark_rod(nmax)]
    for strategy in STRATEGIES:
        diffs, _, b = run_pair(rods, None, benchmark_params(), 20, strategy, checkpoints=[1, 20])
        assert_bitwise(diffs)
    big = [benchmark_rod(nmax + 1)]
    diffs, _, b = run_pair(big, None, benchmark_params(), 5, "auto")
    assert b.counters()["strategy"] == "plain"
    assert_bitwise(diffs)
    with pytest.raises(rp.RopePhysicsError) as e:
        rp.Batch(big, strategy="tiled")
    assert e.value.code == _abi.RP_ERR_CAPACITY


def test_invalid_inputs_are_rejected():
    rod = benchmark_rod(8)
    b = rp.Batch(rod, 2)
    with pytest.raises(ValueError):
        b.set_state(0, 1, x=np.zeros((3, 3)))
    with pytest.raises(ValueError):
        b.get_state(1, 5)
    from ropephysics_amd import _abi
    out = np.zeros(1000, np.float32)
    assert _abi.lib().rp_batch_get_state(b._h, 1, 5, _abi.fptr(out), None, None, None) == _abi.RP_ERR_INVALID
    assert b"outside" in _abi.lib().rp_last_error()
    with pytest.raises(rp.RopePhysicsError):
        b.set_material_scales([1.0, -1.0], None)
    with pytest.raises(ValueError):
        b.set_material_scales([1.0], None)
    bad = rod.copy()
    bad.sP1 = bad.sP1.copy()
    bad.sP1[3] = 999
    with pytest.raises(rp.RopePhysicsError) as e:
        rp.Batch([bad])
    assert e.value.code == -1
    with pytest.raises(rp.RopePhysicsError):
        b.step(rp.SolverParams(substeps=0), 1)
