"""Trajectory recording on the device (rp_batch_record / rp_batch_get_frames).

The reference records a trajectory by downloading every rod's positions after
each frame's steps (Batch::downloadPositions, src/gpu/gpu_batch.cpp:445-455)
and writing them as FP32 (TrajectoryWriter::frame, src/validation/scenes.cpp:
45-68). Here the stepping kernels store each frame themselves. The bar:
  * frame k bitwise equal to the positions a second batch downloads after
    stepping (k + 1) * every steps, rounded to FP32, for both strategies and
    both precisions, across launch boundaries;
  * frame k bitwise equal to the oracle's positions at the same step (FP32
    oracle for FP32 batches, FP64 oracle rounded to FP32 for FP64 batches);
  * recording does not change the state: a recording batch ends bitwise where
    a batch that never recorded does;
  * the frame cap, restarts with a larger buffer, stopping, and the .rodtraj
    file of a uniform batch read back unchanged.
"""
import numpy as np
import pytest

import ropephysics_amd as rp
from ropephysics_amd import _abi, trajectory
from helpers import benchmark_params, benchmark_rod, bits, oracle_params, oracle_state, oracle_twin, oracle_world
from test_gpu_parity import CONTACT_WORLD, SELF_PARAMS, SELF_WORLD, helix_rope

pytestmark = pytest.mark.gpu


def mixed_rods():
    return [benchmark_rod(7), helix_rope(), benchmark_rod(33), benchmark_rod(97)]


def stepwise_frames(rods, world, params, strategy, precision, every, frames):
    """Positions after each `every` steps, downloaded from a batch that
    records nothing, rounded to FP32."""
    b = rp.Batch(rods, strategy=strategy, world=world, precision=precision)
    out = []
    for _ in range(frames):
        b.step(params, every)
        out.append(b.download_positions().astype(np.float32))
    return np.stack(out), b


@pytest.mark.parametrize("precision", [32, 64])
@pytest.mark.parametrize("strategy", ["plain", "tiled"])
def test_frames_equal_downloaded_state_and_oracle(strategy, precision):
    rods, every, F = mixed_rods(), 7, 12
    print(f"recording {strategy} FP{precision}: every {every}, {F} frames, 100 steps", flush=True)
    a = rp.Batch(rods, strategy=strategy, world=SELF_WORLD, precision=precision)
    a.record(every, F)
    a.step(SELF_PARAMS, 100)  # frames at steps 7..84; steps 91 and 98 are past the cap
    assert a.recorded_frames() == F
    got = a.frames()
    assert got.shape == (F, int(a.particle_start[-1]), 3) and got.dtype == np.float32

    want, b = stepwise_frames(rods, SELF_WORLD, SELF_PARAMS, strategy, precision, every, F)
    assert np.array_equal(bits(got), bits(want))

    # Recording leaves the physics alone.
    b.step(SELF_PARAMS, 100 - every * F)
    sa, sb = a.get_state(), b.get_state()
    for k in ("x", "v", "q", "omega"):
        assert np.array_equal(np.asarray(sa[k]).view(np.uint8), np.asarray(sb[k]).view(np.uint8)), k

    # The oracle at the same steps, rod by rod.
    ow = oracle_world(SELF_WORLD)
    op = oracle_params(SELF_PARAMS, precision=precision)
    ps = a.particle_start
    for r, rod in enumerate(rods):
        o = oracle_twin(rod, precision)
        for k in range(F):
            o.step(op, ow, every)
            x = np.asarray(oracle_state(o)["x"]).reshape(-1, 3).astype(np.float32)
            assert np.array_equal(bits(got[k, ps[r]:ps[r + 1]]), bits(x)), (r, k)


def test_cap_restart_and_stop():
    rods = mixed_rods()
    b = rp.Batch(rods, strategy="tiled", world=SELF_WORLD)
    b.record(5, 2)
    b.step(SELF_PARAMS, 30)
    first = b.frames()
    assert first.shape[0] == 2
    b.step(SELF_PARAMS, 10)  # past the cap: the buffer is not touched
    assert np.array_equal(bits(b.frames()), bits(first))

    # A restart with a larger buffer: frames count from the current clocks,
    # through a re-captured graph that holds the new buffer.
    b.record(9, 20)
    assert b.recorded_frames() == 0
    b.step(SELF_PARAMS, 9 * 20 + 3)
    assert b.recorded_frames() == 20
    frames = b.frames()
    ref = rp.Batch(rods, strategy="tiled", world=SELF_WORLD)
    ref.step(SELF_PARAMS, 40)
    for k in range(20):
        ref.step(SELF_PARAMS, 9)
        assert np.array_equal(bits(frames[k]), bits(ref.download_positions())), k

    # Stop: no frames to read, stepping goes on unchanged.
    b.record(0, 0)
    assert b.recorded_frames() == 0
    with pytest.raises(RuntimeError):
        b.frames(0, 1)
    b.step(SELF_PARAMS, 17)
    ref.step(SELF_PARAMS, 3 + 17)
    assert np.array_equal(bits(b.download_positions()), bits(ref.download_positions()))


def test_invalid_recording_calls_are_refused():
    b = rp.Batch([benchmark_rod(10)], strategy="tiled", world=CONTACT_WORLD)
    lib = _abi.lib()
    assert lib.rp_batch_record(b._h, -1, 4) == _abi.RP_ERR_INVALID
    assert lib.rp_batch_record(b._h, 4, 0) == _abi.RP_ERR_INVALID
    assert lib.rp_batch_get_frames(b._h, 0, 1, None) == _abi.RP_ERR_INVALID
    out = np.zeros((2, 11, 3), np.float32)
    assert lib.rp_batch_get_frames(b._h, 0, 1, _abi.fptr(out)) == _abi.RP_ERR_STATE
    b.record(3, 2)
    assert lib.rp_batch_get_frames(b._h, 1, 2, _abi.fptr(out)) == _abi.RP_ERR_INVALID
    assert lib.rp_batch_get_frames(b._h, 0, 2, _abi.fptr(out)) == _abi.RP_OK


def test_rodtraj_of_a_uniform_batch_and_split_of_a_ragged_one(tmp_path):
    rods = [benchmark_rod(24) for _ in range(5)]
    p = benchmark_params()
    b = rp.Batch(rods, strategy="tiled", world=CONTACT_WORLD)
    frames, wall = trajectory.record(b, p, frames=6, steps_per_frame=4)
    assert wall.shape == (6,) and np.all(wall > 0)
    path = str(tmp_path / "uniform.rodtraj")
    n = trajectory.write_rodtraj(path, frames.reshape(6, 5, 25, 3), radius=0.005, fps=250.0, wall=wall)
    assert n == 28 + 6 * (4 + 5 * 25 * 12 + 4)
    t = trajectory.read_rodtraj(path)
    assert t.num_frames == 6 and t.positions.shape == (6, 5, 25, 3)
    assert np.array_equal(bits(t.positions.reshape(6, -1, 3)), bits(frames))
    assert all(c.shape == (0, 3) for c in t.contacts)

    ragged = [benchmark_rod(n) for n in (10, 24, 10, 40, 24)]
    rb = rp.Batch(ragged, strategy="tiled", world=CONTACT_WORLD)
    fr, _ = trajectory.record(rb, p, frames=3, steps_per_frame=5)
    groups = trajectory.split_by_size(fr, rb.particle_start)
    assert sorted(groups) == [11, 25, 41]
    rods11, pos11 = groups[11]
    assert list(rods11) == [0, 2] and pos11.shape == (3, 2, 11, 3)
    ps = rb.particle_start
    assert np.array_equal(bits(pos11[:, 1]), bits(fr[:, ps[2]:ps[3]]))
