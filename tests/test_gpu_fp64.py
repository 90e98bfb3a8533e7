"""GPU parity of FP64 batches (rp_batch_create64) against the FP64 oracle.

An FP64 batch computes the reference CPU solver's arithmetic in its default
precision (crs::Real = double, src/core/math3.h:14-22). The bar:
  * every FP64 word of x, v, q and omega equal to the FP64 oracle run in the
    GPU's colour order, at every checkpoint, for both strategies and every
    tile shape built for FP64;
  * routed / not-routed verdicts of full routing rollouts equal to the FP64
    oracle's, cable for cable (the positions themselves bitwise equal);
  * the FP32 batch within the float envelope of the FP64 batch;
  * precision mismatches refused with RP_ERR_STATE.
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle
import ropephysics_amd as rp
from ropephysics_amd import _abi
from helpers import benchmark_params, benchmark_rod, max_pos_err, oracle_params, oracle_state, oracle_twin, \
    oracle_world
from test_gpu_parity import ALL_PRIMS_WORLD, CONTACT_WORLD, SELF_PARAMS, SELF_WORLD, helix_rope, pinned_coil

pytestmark = pytest.mark.gpu

STRATEGIES = ["plain", "tiled"]
FP64_TILE_CONFIGS = [1, 2, 5, 6, 9, 10, 11, 12]  # one constraint trip per colour, at most 256 lanes


def bits64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def diff_words(gpu: dict, ora: dict) -> dict:
    out = {}
    for k in ("x", "v", "q", "omega"):
        assert np.asarray(gpu[k]).dtype == np.float64
        out[k] = int(np.count_nonzero(bits64(gpu[k]) != bits64(ora[k])))
    return out


def run64(rods, world, params, steps, strategy, checkpoints=None, scales=None, tile_config=0):
    """Step rods in an FP64 batch and each rod on the FP64 oracle in colour
    order; returns the differing FP64 words per field over rods and
    checkpoints, and the batch."""
    b = rp.Batch(rods, strategy=strategy, world=world, tile_config=tile_config, precision=64)
    assert b.precision == 64 and _abi.lib().rp_batch_precision(b._h) == 64
    ys, fs = (None, None) if scales is None else scales
    if scales is not None:
        b.set_material_scales(ys, fs)
    twins = []
    for i, r in enumerate(rods):
        cs = 1.0 / float(ys[i]) if ys is not None else 1.0
        f = float(fs[i]) if fs is not None else 1.0
        twins.append((oracle_twin(r, 64), oracle_params(params, compliance_scale=cs, friction_scale=f)))
    ow = oracle_world(world)
    diffs = {"x": 0, "v": 0, "q": 0, "omega": 0}
    done = 0
    for cp in checkpoints or [steps]:
        b.step(params, cp - done)
        for o, op in twins:
            o.step(op, ow, cp - done)
        done = cp
        b.synchronize()
        for i, (o, _) in enumerate(twins):
            d = diff_words(b.download(i), oracle_state(o))
            for k in diffs:
                diffs[k] += d[k]
    return diffs, b


def assert_bitwise(diffs):
    assert all(v == 0 for v in diffs.values()), diffs


@pytest.mark.parametrize("strategy", STRATEGIES)
def test_fp64_cantilever_bitwise_and_fp32_within_float_envelope(strategy):
    n = 48
    diffs, b64 = run64([benchmark_rod(n)], None, benchmark_params(), 200, strategy, checkpoints=[1, 10, 200])
    assert_bitwise(diffs)
    b32 = rp.Batch([benchmark_rod(n)], strategy=strategy)
    b32.step(benchmark_params(), 200)
    x32, x64 = b32.download(0)["x"], b64.download(0)["x"]
    assert x32.dtype == np.float32
    err = max_pos_err(x32, x64)
    assert 0.0 < err < 1e-3  # the reference's float-drift bar at 200 steps (gpu_cases.cpp:303-304)


@pytest.mark.parametrize("strategy", STRATEGIES)
def test_fp64_contacts_all_primitives_and_scales(strategy):
    diffs, _ = run64([benchmark_rod(48)], CONTACT_WORLD, benchmark_params(), 120, strategy, checkpoints=[1, 120])
    assert_bitwise(diffs)
    rods = [benchmark_rod(40), benchmark_rod(64)]
    diffs, _ = run64(rods, ALL_PRIMS_WORLD, benchmark_params(iterations=2), 100, strategy,
                     scales=([0.5, 3.0], [0.7, 1.3]))
    assert_bitwise(diffs)


@pytest.mark.parametrize("strategy", STRATEGIES)
def test_fp64_self_collision(strategy):
    world = rp.CollisionWorld(primitives=[rp.Primitive.plane((0, 0, 0), (0, 0, 1), 0.6)], self_collision=True,
                              self_friction=0.3, hash_table_size=64, self_capacity=700)
    rods = [pinned_coil(), helix_rope(), benchmark_rod(24)]
    diffs, b = run64(rods, world, SELF_PARAMS, 30, strategy, checkpoints=[1, 30])
    assert b.self_contact_overflow() == 0
    assert_bitwise(diffs)
    diffs, b = run64([helix_rope()], SELF_WORLD, SELF_PARAMS, 40, strategy, checkpoints=[5, 40])
    assert_bitwise(diffs)


def test_fp64_every_tile_shape_same_bits_and_fp32_only_shapes_refused():
    rods = [benchmark_rod(n) for n in (12, 40, 64, 90)]
    ref = None
    for cfg in FP64_TILE_CONFIGS:
        b = rp.Batch(rods, strategy="tiled", world=CONTACT_WORLD, tile_config=cfg, precision=64)
        b.step(benchmark_params(), 40)
        s = b.get_state()
        assert b.counters()["tile_config"] == cfg
        if ref is None:
            ref = s
        for k in s:
            assert np.array_equal(bits64(s[k]), bits64(ref[k])), (cfg, k)
    for cfg in sorted(set(rp.tile_configs()) - set(FP64_TILE_CONFIGS)):
        with pytest.raises(_abi.RopePhysicsError) as e:
            rp.Batch(rods, strategy="tiled", tile_config=cfg, precision=64)
        assert e.value.code == _abi.RP_ERR_INVALID


def test_precision_mismatch_is_refused():
    import ctypes as C

    rod = benchmark_rod(16)
    b64 = rp.Batch([rod], strategy="tiled", precision=64)
    b32 = rp.Batch([rod], strategy="tiled")
    L = _abi.lib()
    x32 = np.zeros((17, 3), np.float32)
    x64 = np.zeros((17, 3), np.float64)
    assert L.rp_batch_get_state(b64._h, 0, 1, _abi.fptr(x32), None, None, None) == _abi.RP_ERR_STATE
    assert L.rp_batch_get_state64(b32._h, 0, 1, _abi.dptr(x64), None, None, None) == _abi.RP_ERR_STATE
    sp32 = rp.batch.step_params(benchmark_params(), 32)
    sp64 = rp.batch.step_params(benchmark_params(), 64)
    assert L.rp_batch_step(b64._h, C.byref(sp32), 1) == _abi.RP_ERR_STATE
    assert L.rp_batch_step64(b32._h, C.byref(sp64), 1) == _abi.RP_ERR_STATE
    assert L.rp_batch_precision(b32._h) == 32
    # Shared calls take either kind.
    assert L.rp_batch_set_clock(b64._h, 0, 1, None) == 0
    c = b64.counters()
    assert c["steps_taken"] == 0
    # Nothing ran: both states are still the uploaded rod.
    assert np.array_equal(b64.download(0)["x"], rod.x)


def test_fp64_routing_rollouts_bitwise_and_verdicts_equal_oracle():
    """Full 6 s rollouts of eleven cables of the routing workload (seeded
    sizes, paths, stiffness and friction; one cable sent straight to the end
    point): the FP64 batch's final positions equal the FP64 oracle's driven by
    the same gripper schedule, word for word, so the routed / not-routed
    verdicts are the oracle's, and the sample holds both outcomes."""
    from ropephysics_amd import scenes

    w = scenes.routing_rollouts()
    task, paths = w.info["task"], w.info["paths"]
    sel = list(range(3, w.num_rods, 103))
    rods = [w.rods[r] for r in sel] + [w.rods[sel[0]]]
    cable_paths = [paths[r] for r in sel] + [task.path(w.rods[sel[0]].num_segments,
                                                       np.tile(task.base_waypoints()[-1], (8, 1)))]
    ys = np.array([w.youngs_scale[r] for r in sel] + [1.0], np.float64)
    fs = np.array([w.friction_scale[r] for r in sel] + [1.0], np.float64)
    b = rp.Batch(rods, strategy="tiled", world=w.world, precision=64)
    b.set_material_scales(ys, fs)
    b.set_paths(0, cable_paths)
    b.step(w.params, w.steps)
    b.synchronize()
    ow = oracle_world(w.world)
    twins = [oracle_twin(r, 64) for r in rods]

    def roll(i):
        kv = np.array([[cable_paths[i].velocity(s) for s in range(w.steps)]])
        op = oracle_params(w.params, compliance_scale=1.0 / ys[i], friction_scale=fs[i])
        oracle.step_many_driven([twins[i]], op, [ow], w.steps, 1, [cable_paths[i].particle], kv)

    with ThreadPoolExecutor(8) as ex:
        list(ex.map(roll, range(len(rods))))
    for i, o in enumerate(twins):
        d = diff_words(b.download(i), oracle_state(o))
        assert all(v == 0 for v in d.values()), (i, d)
    gpu = [task.evaluate(b.download(i)["x"]) for i in range(len(rods))]
    cpu = [task.evaluate(o.get()["x"]) for o in twins]
    assert [g["routed"] for g in gpu] == [c["routed"] for c in cpu]
    verdicts = [g["routed"] for g in gpu]
    assert any(verdicts[:-1]) and not all(verdicts[:-1])
    assert not verdicts[-1]


def test_config1_tip_loaded_cantilever_fp64_meets_1e4_of_length():
    """The single tip-loaded cantilever (1 m, 64 segments, E = 1e7, Poisson
    0.5, no gravity, 0.01 N ramped over 0.5 s, 20 substeps, 2 s), the load set
    before every step. The FP64 batch equals the FP64 oracle in colour order
    word for word and stays within 1e-4 of the rod's length of the reference's
    sequential FP64 CPU path at 1 s and 2 s. The FP32 batch equals the FP32
    oracle word for word -- the reference GPU path's arithmetic -- whose own
    error against FP64 (about 4e-3 m at 1 s: the large, undamped deflection
    amplifies FP32 rounding) is what FP32 can do here."""
    from ropephysics_amd import scenes

    rod, params, steps, tip = scenes.cantilever_tip_load()
    b64 = rp.Batch([rod], strategy="tiled", precision=64)
    b32 = rp.Batch([rod], strategy="tiled")
    o64, seq, o32 = oracle_twin(rod, 64), oracle_twin(rod, 64), oracle_twin(rod, 32)
    op, op_seq = oracle_params(params), oracle_params(params, coloured=False)
    force = np.array(rod.extForce, np.float64)
    checked = 0
    for s in range(steps):
        force[-1] = (0.0, 0.0, tip(s))
        b64.set_loads_range(0, 1, force=force)
        b32.set_loads_range(0, 1, force=force)
        for o in (o64, seq, o32):
            o.set({"extForce": force})
        b64.step(params, 1)
        b32.step(params, 1)
        o64.step(op, None, 1)
        seq.step(op_seq, None, 1)
        o32.step(op, None, 1)
        if s + 1 in (1000, 2000):
            g64, g32 = b64.download(0), b32.download(0)
            assert all(v == 0 for v in diff_words(g64, oracle_state(o64)).values())
            ref32 = oracle_state(o32)
            for k in ("x", "v", "q", "omega"):
                assert np.array_equal(np.asarray(g32[k], np.float32).view(np.uint32),
                                      np.asarray(ref32[k], np.float32).view(np.uint32)), k
            assert max_pos_err(g64["x"], seq.get()["x"]) <= 1e-4 * 1.0
            assert max_pos_err(g32["x"], seq.get()["x"]) == max_pos_err(ref32["x"], seq.get()["x"])
            checked += 1
    assert checked == 2
    assert g64["x"][-1, 2] < -0.3  # large deflection
