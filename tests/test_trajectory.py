"""The .rodtraj trajectory format and scene sidecar (ropephysics_amd.trajectory),
on the CPU.

The layout is the reference's (src/validation/scenes.cpp:13-21, written by
TrajectoryWriter, :45-90, read by tools/render_scene.py:56-75). The bytes are
checked against the layout written out field by field with `struct`, not
against the writer's own code path."""
import json
import struct

import numpy as np
import pytest

import ropephysics_amd as rp
from ropephysics_amd import trajectory


def test_layout_byte_for_byte(tmp_path):
    rng = np.random.default_rng(3)
    pos = rng.standard_normal((2, 3, 4, 3)).astype(np.float32)
    wall = np.array([0.25, 0.5], np.float32)
    contacts = [np.zeros((0, 3), np.float32), np.array([[1, 2, 3], [4, 5, 6]], np.float32)]
    path = tmp_path / "t.rodtraj"
    n = trajectory.write_rodtraj(str(path), pos, radius=0.004, fps=60.0, wall=wall, contacts=contacts)
    raw = path.read_bytes()
    want = b"RODTRAJ1" + struct.pack("<iii", 2, 3, 4) + struct.pack("<ff", 0.004, 60.0)
    for k in range(2):
        want += struct.pack("<f", float(wall[k]))
        want += struct.pack("<%df" % (3 * 4 * 3), *pos[k].reshape(-1).tolist())
        want += struct.pack("<i", len(contacts[k]))
        want += struct.pack("<%df" % (3 * len(contacts[k])), *contacts[k].reshape(-1).tolist())
    assert raw == want and n == len(want)


def test_round_trip_and_rejections(tmp_path):
    rng = np.random.default_rng(4)
    pos = rng.standard_normal((5, 2, 7, 3)).astype(np.float32)
    path = tmp_path / "r.rodtraj"
    trajectory.write_rodtraj(str(path), pos, radius=0.005, fps=30.0)
    t = trajectory.read_rodtraj(str(path))
    assert t.num_frames == 5 and t.radius == np.float32(0.005) and t.fps == 30.0
    assert np.array_equal(t.positions, pos) and np.all(t.wall == 0)
    assert all(c.shape == (0, 3) for c in t.contacts)

    raw = path.read_bytes()
    (tmp_path / "bad.rodtraj").write_bytes(b"RODTRAJ2" + raw[8:])
    with pytest.raises(ValueError):
        trajectory.read_rodtraj(str(tmp_path / "bad.rodtraj"))
    (tmp_path / "long.rodtraj").write_bytes(raw + b"\0\0\0\0")
    with pytest.raises(ValueError):
        trajectory.read_rodtraj(str(tmp_path / "long.rodtraj"))
    with pytest.raises(ValueError):
        trajectory.write_rodtraj(str(tmp_path / "x.rodtraj"), pos[0], radius=0.005, fps=30.0)
    with pytest.raises(ValueError):
        trajectory.write_rodtraj(str(tmp_path / "x.rodtraj"), pos, radius=0.005, fps=30.0, wall=np.zeros(4))


def test_split_by_size():
    sizes = [3, 5, 3, 2, 5, 5]
    ps = np.concatenate([[0], np.cumsum(sizes)])
    F = 4
    frames = np.arange(F * ps[-1] * 3, dtype=np.float32).reshape(F, ps[-1], 3)
    groups = trajectory.split_by_size(frames, ps)
    assert sorted(groups) == [2, 3, 5]
    for n, (rods, pos) in groups.items():
        assert list(rods) == [r for r, s in enumerate(sizes) if s == n]
        assert pos.shape == (F, len(rods), n, 3)
        for j, r in enumerate(rods):
            assert np.array_equal(pos[:, j], frames[:, ps[r]:ps[r + 1]])


def test_scene_json(tmp_path):
    world = rp.CollisionWorld(primitives=[rp.Primitive.plane((0, 0, 0), (0, 0, 1), 0.5),
                                          rp.Primitive.capsule((0, 0, 0), (0, 0, 0.05), 0.015, 0.5)])
    params = rp.SolverParams(dt=1e-3, substeps=10, iterations=1, gravity=(0, 0, -9.81))
    path = tmp_path / "s.json"
    trajectory.write_scene_json(str(path), "routing", "cables routed past two pegs", world, 0.004, 62.5, 375, 12, 101,
                                params, 16, {"simulated_seconds": 6.0, "wall_seconds": 1.5, "segment_substeps": 6e9,
                                             "device": "MI355X"}, labels=[("peg A", (0.2, -0.1, 0.0))])
    doc = json.loads(path.read_text())
    assert doc["frames"] == 375 and doc["rods"] == 12 and doc["particles"] == 101
    assert doc["solver"] == {"dt": 1e-3, "substeps": 10, "iterations": 1, "steps_per_frame": 16}
    assert doc["timing"]["realtime_factor"] == 4.0 and doc["timing"]["segment_substeps_per_sec"] == 4e9
    assert [p["type"] for p in doc["primitives"]] == ["plane", "capsule"]
    assert doc["primitives"][1]["radius"] == 0.015 and doc["labels"][0]["text"] == "peg A"
