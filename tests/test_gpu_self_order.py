"""GPU parity of the coloured self-contact order (CollisionWorld.self_order = 1).

This order is an opt-in of this build: the reference projects self contacts
one at a time in list order (src/core/solver.cpp:163-216). The coloured order
colours the list greedily and projects colour by colour. In TILED batches the
lanes of a rod's group project one colour's contacts at once. The bar:
  * every FP32 (FP64) word of x, v, q and omega equals the FP32 (FP64) oracle
    run in the same coloured order, for both strategies, at every checkpoint;
  * every tile shape gives the same bits;
  * rods in list order and in coloured order side by side in one batch each
    equal their own oracle;
  * the coloured order really differs from the list order on a dense coil
    (several colours in use), so these tests exercise the parallel rounds.
"""
import dataclasses

import numpy as np
import pytest

import ropephysics_amd as rp
from helpers import bits, compare_bitwise, oracle_params, oracle_state, oracle_twin, oracle_world
from test_gpu_parity import SELF_PARAMS, assert_bitwise, benchmark_rod, helix_rope, pinned_coil, run_pair
from test_gpu_fp64 import run64

pytestmark = pytest.mark.gpu

COLOURED = rp.CollisionWorld(primitives=[rp.Primitive.plane((0, 0, 0), (0, 0, 1), 0.6)], self_collision=True,
                             self_friction=0.3, hash_table_size=64, self_capacity=700, self_order=1)
LISTED = dataclasses.replace(COLOURED, self_order=0)


def coloured_rods():
    return [pinned_coil(), helix_rope(), benchmark_rod(24), pinned_coil(turns=3, per_turn=20)]


def test_the_coil_needs_several_colours():
    o = oracle_twin(pinned_coil(), 64)
    op, ow = oracle_params(SELF_PARAMS), oracle_world(COLOURED)
    most = 0
    for _ in range(3):
        o.step(op, ow, 1)
        most = max(most, int(o.self_contacts()["colour"].max(initial=0)))
    assert most >= 2


@pytest.mark.parametrize("strategy", ["plain", "tiled"])
def test_coloured_order_bitwise_fp32(strategy):
    print(f"coloured order FP32 {strategy}", flush=True)
    diffs, _, b = run_pair(coloured_rods(), COLOURED, SELF_PARAMS, 30, strategy, checkpoints=[1, 3, 30])
    assert b.self_contact_overflow() == 0
    assert_bitwise(diffs)
    listed, _, lb = run_pair(coloured_rods()[:1], LISTED, SELF_PARAMS, 30, strategy, checkpoints=[30])
    assert_bitwise(listed)
    assert not np.array_equal(bits(lb.download(0)["x"]), bits(b.download(0)["x"]))


@pytest.mark.parametrize("strategy", ["plain", "tiled"])
def test_coloured_order_bitwise_fp64(strategy):
    print(f"coloured order FP64 {strategy}", flush=True)
    diffs, b = run64(coloured_rods(), COLOURED, SELF_PARAMS, 30, strategy, checkpoints=[1, 30])
    assert b.self_contact_overflow() == 0
    assert all(v == 0 for v in diffs.values()), diffs


@pytest.mark.parametrize("tile_config", [1, 3, 4, 9, 11, 12, 13])
def test_coloured_order_every_tile_shape(tile_config):
    print(f"coloured order, tile configuration {tile_config}", flush=True)
    diffs, _, b = run_pair(coloured_rods(), COLOURED, SELF_PARAMS, 12, "tiled", checkpoints=[2, 12],
                           tile_config=tile_config)
    assert b.counters()["tile_config"] == tile_config
    assert_bitwise(diffs)


@pytest.mark.parametrize("strategy", ["plain", "tiled"])
def test_list_and_coloured_rods_in_one_batch(strategy):
    rods = coloured_rods() + coloured_rods()
    wor = [0, 1, 0, 1, 1, 0, 1, 0]
    worlds = [LISTED, COLOURED]
    b = rp.Batch(rods, strategy=strategy, world=worlds, world_of_rod=wor)
    twins = [(oracle_twin(r, 32), oracle_world(worlds[w])) for r, w in zip(rods, wor)]
    op = oracle_params(SELF_PARAMS)
    done = 0
    for cp in (2, 20):
        b.step(SELF_PARAMS, cp - done)
        for o, ow in twins:
            o.step(op, ow, cp - done)
        done = cp
        b.synchronize()
        for i, (o, _) in enumerate(twins):
            d = compare_bitwise(b.download(i), oracle_state(o))
            assert all(v == 0 for v in d.values()), (i, cp, d)


def test_server_runs_coloured_requests_as_alone():
    """Requests in the coloured order, in list order and without self-collision,
    served together: each result is bitwise what its rods give alone in a batch
    (the server launches the coloured-order kernel variant when a running rod
    asks for it)."""
    from test_gpu_server import PARAMS, _alone_in_batch, _assert_same, _serve, _split

    plain_world = rp.CollisionWorld(primitives=COLOURED.primitives)
    requests = [([pinned_coil(), helix_rope()], COLOURED, 40, np.array([1.0, 1.5]), np.array([1.0, 0.8])),
                ([pinned_coil()], LISTED, 30, np.array([0.7]), np.array([1.2])),
                ([benchmark_rod(30), benchmark_rod(12)], plain_world, 25, np.array([1.0, 2.0]), np.array([1.0, 1.0])),
                ([pinned_coil(turns=3, per_turn=20)], COLOURED, 20, np.array([1.0]), np.array([1.0]))]
    server = rp.Server(PARAMS, hbm_budget_bytes=64 << 20, steps_per_launch=32)
    results = _serve(server, requests, arrivals=[0, 0, 1, 2])
    for k, (rods, world, steps, ys, fs) in enumerate(requests):
        _assert_same(_split(results[k], rods), _alone_in_batch(rods, world, steps, ys, fs))
