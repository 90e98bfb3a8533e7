"""GPU tests of the continuous-batching server: each request's result is
bitwise the one the same rods get alone (in a Batch, and alone in a server),
whatever else is resident and whenever it was admitted; horizons are honoured
to the step; admission follows the HBM budget and freed ranges are reused."""
import numpy as np
import pytest

import ropephysics_amd as rp
from helpers import benchmark_rod, bits, compare_bitwise, oracle_params, oracle_state, oracle_twin, oracle_world

pytestmark = pytest.mark.gpu

PARAMS = rp.SolverParams(dt=1e-3, substeps=10, iterations=1, gravity=(0.0, 0.0, -9.81))


def _world(seed: int, self_collision: bool) -> rp.CollisionWorld:
    rng = np.random.default_rng(seed)
    prims = [rp.Primitive.plane((0, 0, -0.02), (0, 0, 1), float(rng.uniform(0.3, 0.9)))]
    prims.append(rp.Primitive.sphere((float(rng.uniform(0.2, 0.6)), 0.0, -0.09), 0.08, 0.4))
    prims.append(rp.Primitive.capsule((0.3, -0.1, -0.03), (0.3, 0.1, -0.03), 0.012, 0.5))
    prims.append(rp.Primitive.box((0.7, 0.0, -0.035), (0.05, 0.05, 0.02), (1.0, 0.0, 0.0, 0.0), 0.6))
    return rp.CollisionWorld(primitives=prims, self_collision=self_collision, self_friction=0.3,
                             hash_table_size=64)


def _request(seed: int):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 6))
    rods = [benchmark_rod(int(rng.integers(8, 80))) for _ in range(n)]
    ys = rng.uniform(0.5, 2.0, n)
    fs = rng.uniform(0.5, 1.5, n)
    world = _world(seed, bool(rng.integers(0, 2)))
    steps = int(rng.integers(10, 150))
    return rods, world, steps, ys, fs


def _alone_in_batch(rods, world, steps, ys, fs):
    b = rp.Batch(rods, strategy="tiled", world=world, tile_config=rp.server.DEFAULT_TILE_CONFIG)
    b.set_material_scales(ys, fs)
    b.step(PARAMS, steps)
    b.synchronize()
    return [b.download(i) for i in range(len(rods))]


def _split(res, rods):
    out, p0, f0 = [], 0, 0
    for r in rods:
        P, F = r.num_particles, r.num_frames
        out.append({"x": res["x"][p0:p0 + P], "v": res["v"][p0:p0 + P], "q": res["q"][f0:f0 + F],
                    "omega": res["omega"][f0:f0 + F]})
        p0 += P
        f0 += F
    return out


def _assert_same(got, want):
    for g, w in zip(got, want):
        d = compare_bitwise(g, w)
        assert all(v == 0 for v in d.values()), d


def _serve(server, requests, arrivals):
    """Submit request k once `arrivals[k]` chunks have run (admission retried
    when the budget is full), step until all are done; results by index."""
    pending = list(range(len(requests)))
    ids, results, chunk = {}, {}, 0
    while pending or len(results) < len(requests):
        for k in list(pending):
            if arrivals[k] > chunk:
                continue
            rods, world, steps, ys, fs = requests[k]
            try:
                ids[server.submit(rods, world, steps, ys, fs)] = k
                pending.remove(k)
            except rp.ServerFull:
                break
        server.step()
        chunk += 1
        server.synchronize()
        for rid in server.finished():
            results[ids[rid]] = server.result(rid)
            server.release(rid)
        assert chunk < 1000
    return results


def test_request_results_equal_alone():
    requests = [_request(s) for s in range(8)]
    server = rp.Server(PARAMS, hbm_budget_bytes=64 << 20, steps_per_launch=32)
    results = _serve(server, requests, arrivals=[0, 0, 1, 1, 2, 3, 5, 5])
    for k, (rods, world, steps, ys, fs) in enumerate(requests):
        _assert_same(_split(results[k], rods), _alone_in_batch(rods, world, steps, ys, fs))
    # And alone in a fresh server.
    rods, world, steps, ys, fs = requests[3]
    solo = rp.Server(PARAMS, hbm_budget_bytes=8 << 20, steps_per_launch=32)
    r = _serve(solo, [requests[3]], arrivals=[0])[0]
    for k in r:
        assert np.array_equal(bits(r[k]), bits(results[3][k]))


def test_horizon_is_exact_and_matches_fp32_oracle():
    """A horizon that is not a multiple of the chunk: the rods stop exactly
    there (checked against the FP32 oracle stepping that many steps)."""
    rod = benchmark_rod(33)
    world = _world(7, True)
    server = rp.Server(PARAMS, hbm_budget_bytes=8 << 20, steps_per_launch=16)
    rid = server.submit([rod], world, 37)
    for _ in range(3):
        server.step()
    server.synchronize()
    assert server.finished() == [rid]
    assert server.info(rid)["steps_done"] == 37
    got = server.result(rid)
    o = oracle_twin(rod, 32)
    o.step(oracle_params(PARAMS), oracle_world(world), 37)
    d = compare_bitwise(got, oracle_state(o))
    assert all(v == 0 for v in d.values()), d


def test_admission_by_budget_and_reuse():
    """A budget that holds one of these requests at a time: the second waits
    for the first to be released, then runs in the freed ranges and still
    gives its alone result."""
    big = [benchmark_rod(120) for _ in range(6)]
    world = _world(3, False)
    budget = 250_000
    server = rp.Server(PARAMS, hbm_budget_bytes=budget, steps_per_launch=64)
    s = server.stats()
    assert s["device_bytes"] <= budget * 1.01
    a = server.submit(big, world, 64)
    with pytest.raises(rp.ServerFull):
        server.submit(big, world, 64)
    server.step()
    server.synchronize()
    assert server.finished() == [a]
    first = server.result(a)
    server.release(a)
    assert server.stats()["used_particles"] == 0
    b = server.submit(big, world, 64)
    server.step()
    server.synchronize()
    second = server.result(b)
    for k in first:
        assert np.array_equal(bits(first[k]), bits(second[k]))
    with pytest.raises(rp.ServerFull):
        server.submit([benchmark_rod(120)] * 40, world, 8)  # larger than the whole budget


def test_time_sliced_steps_give_the_same_results():
    """Time-sliced server steps (each workgroup runs the steps it can start in
    a 300 us slice, up to its rods' horizons): every request's result equals
    the one it gets alone in a Batch, bit for bit, and horizons are exact."""
    requests = [_request(s) for s in range(8)]
    server = rp.Server(PARAMS, hbm_budget_bytes=64 << 20, time_slice_us=300)
    results = _serve(server, requests, arrivals=[0, 0, 1, 1, 2, 3, 5, 5])
    for k, (rods, world, steps, ys, fs) in enumerate(requests):
        _assert_same(_split(results[k], rods), _alone_in_batch(rods, world, steps, ys, fs))
    rod = benchmark_rod(33)
    world = _world(7, True)
    solo = rp.Server(PARAMS, hbm_budget_bytes=8 << 20, time_slice_us=300)
    rid = solo.submit([rod], world, 37)
    rounds = 0
    while not solo.finished():
        solo.step()
        solo.synchronize()
        rounds += 1
        assert rounds < 1000
    assert solo.info(rid)["steps_done"] == 37
    o = oracle_twin(rod, 32)
    o.step(oracle_params(PARAMS), oracle_world(world), 37)
    d = compare_bitwise(solo.result(rid), oracle_state(o))
    assert all(v == 0 for v in d.values()), d
    with pytest.raises(rp.RopePhysicsError):
        rp.Server(PARAMS, hbm_budget_bytes=8 << 20, time_slice_us=-1)


def test_admission_and_batch_share_the_packer_lds_bound():
    """A rod whose LDS need lies just above what the packer gives one
    workgroup is refused at submission (and by a TILED batch), never admitted
    and left unpackable; at the largest pool the bound allows, the request is
    admitted and steps. The pool size is found by bisection on TILED batch
    creation."""
    from ropephysics_amd import _abi

    def world(cap):
        return rp.CollisionWorld(primitives=[rp.Primitive.plane((0, 0, -0.02), (0, 0, 1), 0.5)],
                                 self_collision=True, self_friction=0.3, hash_table_size=256, self_capacity=cap)

    rod = benchmark_rod(100)

    def tiled_ok(cap):
        try:
            rp.Batch([rod], strategy="tiled", world=world(cap), tile_config=rp.server.DEFAULT_TILE_CONFIG).destroy()
            return True
        except rp.RopePhysicsError as e:
            assert e.code == _abi.RP_ERR_CAPACITY
            return False

    lo, hi = 16, 8000
    assert tiled_ok(lo) and not tiled_ok(hi)
    while hi - lo > 1:
        mid = (lo + hi) // 2
        lo, hi = (mid, hi) if tiled_ok(mid) else (lo, mid)
    for cap in range(hi, hi + 6):
        assert not tiled_ok(cap)
        server = rp.Server(PARAMS, hbm_budget_bytes=16 << 20)
        with pytest.raises(rp.ServerFull):
            server.submit([rod], world(cap), 8)
    server = rp.Server(PARAMS, hbm_budget_bytes=16 << 20)
    rid = server.submit([rod], world(lo), 8)
    server.step()
    server.synchronize()
    assert server.finished() == [rid]
    assert np.all(np.isfinite(server.result(rid)["x"]))


def test_serving_autotune_picks_a_pair_and_results_do_not_depend_on_it():
    """serving.autotune_shapes serves a sample once per shape pair and returns
    the fastest; a request's result is the same bit for bit under the chosen
    shapes and the defaults, in both server modes."""
    from ropephysics_amd import serving
    from dataclasses import replace

    reqs = [replace(q, steps=min(q.steps, 24)) for q in serving.make_requests(5, seed=11)]
    pairs = serving.SHAPE_PAIRS[:3]
    for slice_us in (0, 300):
        best, trials = serving.autotune_shapes(reqs[:4], 1 << 30, time_slice_us=slice_us, steps=24, pairs=pairs)
        assert best in pairs
        assert [p for p, _ in trials] == list(pairs)
        assert all(t is not None and t > 0 for _, t in trials)
        got = {}
        for key, kw in (("default", {}), ("tuned", {"tile_config": best[0], "self_tile_config": best[1]})):
            server = rp.Server(serving.PARAMS, hbm_budget_bytes=1 << 30, time_slice_us=slice_us, **kw)
            q = reqs[4]
            rid = server.submit(q.rods, q.world, q.steps, q.youngs, q.friction, x=q.x)
            rounds = 0
            while not server.finished():
                server.step()
                server.synchronize()
                rounds += 1
                assert rounds < 1000
            got[key] = server.result(rid)
            server.destroy()
        for k in got["default"]:
            assert np.array_equal(bits(got["default"][k]), bits(got["tuned"][k])), k
