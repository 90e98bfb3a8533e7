"""GPU test of rp_batch_step_noalloc, the one entry point that allocates
nothing: it steps a GLOBAL FP32 batch with the rods' working arrays in the
library's fixed per-device region instead of the batch's own. Its results
must equal rp_batch_step's bit for bit, on two calls in a row (the second on
a smaller batch, so the region holds the first call's larger contents), and
it must never allocate the batch's own working arrays.
"""
import numpy as np
import pytest

import ropephysics_amd as rp
from ropephysics_amd import _abi
from helpers import REFERENCE_MATERIAL, benchmark_params, benchmark_rod, bits

pytestmark = pytest.mark.gpu

FLOOR = rp.CollisionWorld(primitives=[rp.Primitive.plane((0, 0, -0.0051), (0, 0, 1), 0.5),
                                      rp.Primitive.sphere((1.2, 0, -0.15), 0.12, 0.4)])
SELF = rp.CollisionWorld(primitives=[rp.Primitive.plane((0, 0, 0), (0, 0, 1), 0.6)], self_collision=True,
                         self_friction=0.3, hash_table_size=256, self_capacity=700)


def long_rod(n: int) -> rp.Rod:
    rod = rp.make_straight_rod(n, n * 2e-3, REFERENCE_MATERIAL, (0, 0, 0), (1, 0, 0))
    rp.clamp_root_exact(rod)
    return rod


def coil() -> rp.Rod:
    t = np.arange(5 * 16 + 1) / 16 * 2 * np.pi
    pts = np.stack([0.025 * np.cos(t), 0.025 * np.sin(t), 0.03 + 0.008 * t / (2 * np.pi)], axis=-1)
    rod = rp.make_rod_from_centerline(pts, rp.RodMaterial(2e6, 0.35, 1000.0, 5e-3), (0, 0, 1))
    rod.pin_particle(0)
    rod.pin_particle(rod.num_segments)
    return rod


def _pair(rods, world, params, chunks):
    """The same rods stepped by rp_batch_step and by rp_batch_step_noalloc in
    the same chunks; returns both states and the no-alloc batch's counters."""
    a = rp.Batch(rods, strategy="global", world=world)
    b = rp.Batch(rods, strategy="global", world=world)
    before = b.counters()["device_bytes"]
    for n in chunks:
        a.step(params, n)
        b.step_noalloc(params, n)
    a.synchronize()
    b.synchronize()
    c = b.counters()
    assert c["scratch_bytes"] == 0 and c["device_bytes"] == before  # nothing allocated for it
    assert c["steps_taken"] == sum(chunks)
    return a.get_state(), b.get_state(), c


def _assert_equal(sa, sb):
    for k in sa:
        assert np.array_equal(bits(sa[k]), bits(sb[k])), k


def test_noalloc_equals_step_twice_second_smaller():
    params = benchmark_params(iterations=2)
    # First call: 2000- and 1200-segment rods and a 64-segment one.
    sa, sb, _ = _pair([long_rod(2000), long_rod(1200), benchmark_rod(64)], FLOOR, params, [7, 30])
    _assert_equal(sa, sb)
    # Second call, smaller: the region still holds the first call's arrays.
    sa, sb, _ = _pair([long_rod(700)], FLOOR, params, [1, 12])
    _assert_equal(sa, sb)


def test_noalloc_self_collision_equals_step():
    params = rp.SolverParams(dt=1e-3, substeps=16, iterations=4, gravity=(0, 0, -9.81), linear_damping=1.0,
                             angular_damping=1.0)
    sa, sb, c = _pair([coil(), benchmark_rod(40)], SELF, params, [3, 9])
    _assert_equal(sa, sb)
    assert c["self_contact_overflow"] == 0


def test_noalloc_refusals():
    params = benchmark_params()
    with pytest.raises(rp.RopePhysicsError) as e:
        rp.Batch([benchmark_rod(16)], strategy="tiled").step_noalloc(params, 1)
    assert e.value.code == _abi.RP_ERR_STATE
    with pytest.raises(rp.RopePhysicsError) as e:
        rp.Batch([long_rod(900)], strategy="global", precision=64).step_noalloc(params, 1)
    assert e.value.code == _abi.RP_ERR_STATE
    # Working arrays beyond the fixed region: 100 rods of 2000 segments.
    big = rp.Batch(long_rod(2000), 100, strategy="global")
    with pytest.raises(rp.RopePhysicsError) as e:
        big.step_noalloc(params, 1)
    assert e.value.code == _abi.RP_ERR_CAPACITY and "RP_NOALLOC_SCRATCH_BYTES" in str(e.value)
    assert big.counters()["steps_taken"] == 0
