"""GPU tests of the device-side drivers and observers, and of the two
application tasks' verdicts (routing; hold/slide)."""
import numpy as np
import pytest

import oracle
import ropephysics_amd as rp
from ropephysics_amd.apps import HangingCable, RoutingTask
from helpers import bits, compare_bitwise, compliance_scale_f32, oracle_params, oracle_state, oracle_twin, oracle_world

pytestmark = pytest.mark.gpu


def _routing_small(num, seed, segs=(48, 61, 80)):
    task = RoutingTask()
    rng = np.random.default_rng(seed)
    rods, paths = [], []
    for k in range(num):
        n = segs[k % len(segs)]
        rods.append(task.build(n))
        paths.append(task.path(n, task.base_waypoints() + rng.normal(0, 0.015, (8, 2))))
    return task, rods, paths


@pytest.mark.parametrize("strategy", ["plain", "tiled"])
def test_gripper_path_driver_bitwise(strategy):
    """The device evaluates each rod's gripper schedule in FP64 every step;
    the oracle is driven with the same schedule evaluated on the host
    (attemptOnCpu, application_cases.cpp:275-289). Bitwise equal throughout,
    with the routing world (floor, pegs, box clip) and self-collision on."""
    task, rods, paths = _routing_small(3, 7)
    world = task.world()
    params = task.params()
    ys = np.array([0.5, 1.0, 3.0])
    fs = np.array([0.7, 1.0, 1.3])
    b = rp.Batch(rods, strategy=strategy, world=world)
    b.set_material_scales(ys, fs)
    b.set_paths(0, paths)
    twins = [(oracle_twin(r, 32), oracle_params(params, compliance_scale=compliance_scale_f32(ys[i]),
                                                friction_scale=float(np.float32(fs[i]))))
             for i, r in enumerate(rods)]
    ow = oracle_world(world)
    steps = 400
    for s in range(steps):
        for i, (o, op) in enumerate(twins):
            v = np.float32(np.array(paths[i].velocity(s)))
            a = o.get()
            a["kinVel"][paths[i].particle] = v
            o.set({"kinVel": a["kinVel"]})
            o.step(op, ow, 1)
    b.step(params, steps)
    b.synchronize()
    for i, (o, _) in enumerate(twins):
        d = compare_bitwise(b.download(i), oracle_state(o))
        assert all(v == 0 for v in d.values()), (i, d)
    assert list(b.get_clock()) == [steps] * 3
    # The gripper really moved the cable end.
    x = b.download(0)["x"]
    assert np.linalg.norm(x[-1, :2] - np.array(paths[0].start)) > 0.05


def test_clock_and_paths_resume_across_calls():
    task, rods, paths = _routing_small(2, 3)
    params = task.params()
    a = rp.Batch(rods, strategy="tiled", world=task.world())
    a.set_paths(0, paths)
    a.step(params, 250)
    b = rp.Batch(rods, strategy="tiled", world=task.world())
    b.set_paths(0, paths)
    for chunk in (1, 63, 64, 100, 22):
        b.step(params, chunk)
    sa, sb = a.get_state(), b.get_state()
    for k in sa:
        assert np.array_equal(bits(sa[k]), bits(sb[k]))
    assert list(b.get_clock()) == [250, 250]


@pytest.mark.parametrize("strategy", ["plain", "tiled"])
def test_hang_observer_matches_host_sampling(strategy):
    """The device observer's max rise equals the host-side HangOutcome fed the
    FP32 oracle's short-end heights every 50 steps (bitwise, FP64)."""
    task = HangingCable(segments=60, duration=0.3, dt=1e-3, substeps=8, iterations=2)
    rods = [task.build(r) for r in (1.2, 3.5)]
    mus = [0.5, 0.05]
    world = task.world(1.0)
    b = rp.Batch(rods, strategy=strategy, world=world)
    b.set_material_scales(None, mus)
    z0 = [float(r.x[0, 2]) for r in rods]
    b.set_observers(0, [(0, 50, z0[i]) for i in range(2)])
    steps = task.steps()
    b.step(task.params(), steps)
    got = b.get_observations()
    ow = oracle_world(world)
    for i, r in enumerate(rods):
        o = oracle_twin(r, 32)
        op = oracle_params(task.params(), friction_scale=float(np.float32(mus[i])))
        rise = 0.0
        for s in range(1, steps + 1):
            o.step(op, ow, 1)
            if s % 50 == 0:
                rise = max(rise, float(o.get()["x"][0, 2]) - z0[i])
        assert got[i] == rise
    assert got[1] > got[0]  # the slippery, lopsided drape rises more


def test_routing_verdicts_agree_with_fp64_oracle():
    """Full 6 s routing rollouts of ten cables of the routing workload (its
    seeded sizes, paths, stiffness and friction): the routed / not-routed
    verdict from the GPU's FP32 positions equals the verdict from the FP64
    oracle's, and the sample holds both outcomes. One more cable is sent
    straight to the end point (not routed)."""
    from ropephysics_amd import scenes

    w = scenes.routing_rollouts()
    task, paths = w.info["task"], w.info["paths"]
    sel = list(range(3, w.num_rods, 103))
    rods = [w.rods[r] for r in sel] + [w.rods[sel[0]]]
    cable_paths = [paths[r] for r in sel] + [task.path(w.rods[sel[0]].num_segments,
                                                       np.tile(task.base_waypoints()[-1], (8, 1)))]
    ys = np.array([w.youngs_scale[r] for r in sel] + [1.0])
    fs = np.array([w.friction_scale[r] for r in sel] + [1.0])
    b = rp.Batch(rods, strategy="tiled", world=w.world)
    b.set_material_scales(ys, fs)
    b.set_paths(0, cable_paths)
    b.step(w.params, w.steps)
    b.synchronize()
    twins = [oracle_twin(r, 64) for r in rods]
    worlds = []
    for i in range(len(rods)):
        ow = oracle_world(w.world)
        for pr in ow.prims:
            pr.friction = pr.friction * fs[i]
        ow.self_friction = ow.self_friction * fs[i]
        worlds.append(ow)
        a = twins[i].get()
        twins[i].set({"sComp": a["sComp"] / ys[i], "bComp": a["bComp"] / ys[i]})
    kv = np.array([[p.velocity(s) for s in range(w.steps)] for p in cable_paths])
    oracle.step_many_driven(twins, oracle_params(w.params, precision=64), worlds, w.steps, 8,
                            [p.particle for p in cable_paths], kv)
    gpu = [task.evaluate(b.download(i)["x"].astype(np.float64)) for i in range(len(rods))]
    cpu = [task.evaluate(o.get()["x"]) for o in twins]
    assert [g["routed"] for g in gpu] == [c["routed"] for c in cpu], (gpu, cpu)
    verdicts = [g["routed"] for g in gpu]
    assert any(verdicts[:-1]) and not all(verdicts[:-1])
    assert not verdicts[-1]


def test_harness_learning_cross_entropy():
    """The cross-entropy loop of runHarnessLearning (application_cases.cpp:
    353-486) on a small population: every iteration's rollouts are one ragged
    batch reused with new motions, stiffness and friction; the run is
    reproducible to the bit from its seed; starting from the hand-written
    motion, which leaves the cable slack and routes few cables, the learned
    motion routes most fresh cables and beats the hand-written one by at least
    0.3, the margin the reference's own case checks."""
    from ropephysics_amd.learning import HarnessLearner

    runs = [HarnessLearner(population=128, seed=11).run(iterations=4) for _ in range(2)]
    a, b = runs
    assert [(i.success_rate, i.mean_score) for i in a.iterations] == \
        [(i.success_rate, i.mean_score) for i in b.iterations]
    assert np.array_equal(a.mean, b.mean)
    assert a.learned_rate == b.learned_rate and a.hand_rate == b.hand_rate
    assert len(a.iterations) == 4 and all(0.0 <= i.success_rate <= 1.0 for i in a.iterations)
    print("routed per iteration", [i.success_rate for i in a.iterations], "learned", a.learned_rate,
          "hand-written", a.hand_rate, flush=True)
    assert a.hand_rate <= 0.2
    assert a.learned_rate >= 0.8
    assert a.learned_rate - a.hand_rate >= 0.3
    assert a.mean.shape == (8, 2)


def test_harness_learning_verdicts_agree_with_cpu():
    """The reference's learning case reruns learned and hand-written attempts
    on the CPU and requires the same routed verdicts as on the GPU
    (application_cases.cpp:453-480). Here: after 4 iterations of 128
    rollouts, the softest and the stiffest fresh cable under each motion
    (four attempts) are rerun by the FP64 oracle in colour order with the
    same stiffness, friction and gripper schedule; the GPU's verdicts agree
    on all four, and the four hold both outcomes."""
    from concurrent.futures import ThreadPoolExecutor

    from ropephysics_amd.learning import HarnessLearner

    learner = HarnessLearner(population=128, seed=11)
    res = learner.run(iterations=4)
    task = learner.task
    ys, fr = res.fresh["youngs"], res.fresh["friction"]
    picks = [int(np.argmin(ys)), int(np.argmax(ys))]
    attempts = [(r, name, wp) for r in picks for name, wp in (("learned", res.mean),
                                                              ("hand", task.hand_written_waypoints()))]

    def cpu(attempt):
        r, name, wp = attempt
        n = int(res.fresh["segments"][r])
        rod = task.build(n)
        o = oracle_twin(rod, 64)
        a = o.get()
        o.set({"sComp": a["sComp"] / float(ys[r]), "bComp": a["bComp"] / float(ys[r])})
        ow = oracle_world(task.world())
        for pr in ow.prims:
            pr.friction = pr.friction * float(fr[r])
        ow.self_friction = ow.self_friction * float(fr[r])
        path = task.path(n, wp)
        kv = np.array([[path.velocity(s) for s in range(learner.steps)]])
        oracle.step_many_driven([o], oracle_params(task.params(), precision=64), [ow], learner.steps, 1,
                                [path.particle], kv)
        return task.evaluate(o.get()["x"])["routed"], res.fresh[name][r]

    with ThreadPoolExecutor(4) as ex:
        pairs = list(ex.map(cpu, attempts))
    print("CPU vs GPU routed verdicts", pairs, flush=True)
    assert all(c == g for c, g in pairs)
    assert any(c for c, _ in pairs) and not all(c for c, _ in pairs)
