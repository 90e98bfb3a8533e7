"""Trajectories recorded on the device, in the reference's `.rodtraj` format.

Reference: the trajectory writer and scene sidecar of the demo scenes
(src/validation/scenes.cpp:13-21 for the layout, :45-90 TrajectoryWriter,
:104-160 writePrimitiveJson and writeJson). The renderer reads the same layout
(tools/render_scene.py:56-75).

`.rodtraj`, little-endian:
    char[8]  "RODTRAJ1"
    int32    numFrames, numRods, numParticles
    float32  radius, fps
    per frame:
        float32  seconds spent producing the frame
        float32  numRods * numParticles * 3 positions, rod-major
        int32    numContacts
        float32  numContacts * 3 contact points

The format has one particle count for every rod. A ragged batch is written
as one file per rod size (`split_by_size`).

Frames come from `Batch.record` / `Batch.frames`: the stepping kernels store
them into a device buffer, with no download or host step between frames.
The frame time is therefore the device time of the recorded steps divided
evenly among the frames, not a per-frame wall clock. Contacts are written as
none, as the reference writes for frames downloaded from a GPU batch
(TrajectoryWriter::frame on positions, scenes.cpp:61-68).
"""
from __future__ import annotations

import json
import struct
from dataclasses import dataclass, field

import numpy as np

MAGIC = b"RODTRAJ1"
_PRIM_NAMES = {0: "plane", 1: "sphere", 2: "capsule", 3: "box"}


@dataclass
class Trajectory:
    radius: float
    fps: float
    wall: np.ndarray  # (F,) float32 seconds per frame
    positions: np.ndarray  # (F, R, N, 3) float32
    contacts: list = field(default_factory=list)  # F arrays (C, 3) float32

    @property
    def num_frames(self) -> int:
        return int(self.positions.shape[0])


def write_rodtraj(path: str, positions, radius: float, fps: float, wall=None, contacts=None) -> int:
    """Write frames (F, R, N, 3) as a `.rodtraj` file. `wall` is the seconds
    per frame (default 0), `contacts` F arrays of points or None. Returns the
    bytes written."""
    pos = np.asarray(positions, np.float32)
    if pos.ndim != 4 or pos.shape[3] != 3:
        raise ValueError("positions must be (frames, rods, particles, 3)")
    F, R, N, _ = pos.shape
    wall = np.zeros(F, np.float32) if wall is None else np.asarray(wall, np.float32)
    if wall.shape != (F,):
        raise ValueError("one frame time per frame")
    if contacts is not None and len(contacts) != F:
        raise ValueError("one contact array per frame")
    with open(path, "wb") as f:
        f.write(MAGIC)
        f.write(struct.pack("<3i2f", F, R, N, float(radius), float(fps)))
        for k in range(F):
            f.write(struct.pack("<f", float(wall[k])))
            f.write(np.ascontiguousarray(pos[k], "<f4").tobytes())
            c = np.zeros((0, 3), np.float32) if contacts is None else np.asarray(contacts[k], "<f4").reshape(-1, 3)
            f.write(struct.pack("<i", len(c)))
            f.write(np.ascontiguousarray(c).tobytes())
        return f.tell()


def read_rodtraj(path: str) -> Trajectory:
    """Read a `.rodtraj` file."""
    with open(path, "rb") as f:
        raw = f.read()
    if raw[:8] != MAGIC:
        raise ValueError(f"{path}: not a RODTRAJ1 file")
    F, R, N = struct.unpack_from("<3i", raw, 8)
    radius, fps = struct.unpack_from("<2f", raw, 20)
    off = 28
    per = R * N * 3
    wall = np.empty(F, np.float32)
    pos = np.empty((F, R, N, 3), np.float32)
    contacts = []
    for k in range(F):
        wall[k] = struct.unpack_from("<f", raw, off)[0]
        off += 4
        pos[k] = np.frombuffer(raw, "<f4", per, off).reshape(R, N, 3)
        off += per * 4
        (c,) = struct.unpack_from("<i", raw, off)
        off += 4
        contacts.append(np.frombuffer(raw, "<f4", c * 3, off).reshape(c, 3).copy())
        off += c * 12
    if off != len(raw):
        raise ValueError(f"{path}: {len(raw) - off} trailing bytes")
    return Trajectory(float(radius), float(fps), wall, pos, contacts)


def split_by_size(frames: np.ndarray, particle_start) -> dict:
    """Frames (F, P, 3) of a ragged batch grouped by rod size: {particles per
    rod: (rod indices, (F, R_n, n, 3))}, rods in batch order within a group."""
    ps = np.asarray(particle_start, np.int64)
    sizes = np.diff(ps)
    out = {}
    for n in sorted(set(sizes.tolist())):
        rods = np.nonzero(sizes == n)[0]
        idx = (ps[rods][:, None] + np.arange(n)[None, :]).reshape(-1)
        out[int(n)] = (rods, frames[:, idx, :].reshape(frames.shape[0], len(rods), n, 3))
    return out


def record(batch, params, frames: int, steps_per_frame: int) -> tuple:
    """Step `batch` for frames * steps_per_frame steps with recording on.
    Returns the frames (F, P, 3) and the device seconds per frame (the
    recorded steps' HIP-event time divided evenly)."""
    batch.record(steps_per_frame, frames)
    batch.set_timing(True)
    batch.step(params, frames * steps_per_frame)
    batch.synchronize()
    ms = batch.counters()["last_step_gpu_ms"]
    batch.set_timing(False)
    got = batch.recorded_frames()
    if got != frames:
        raise RuntimeError(f"{got} of {frames} frames recorded")
    return batch.frames(0, frames), np.full(frames, ms / 1e3 / frames, np.float32)


def write_scene_json(path: str, name: str, description: str, world, radius: float, fps: float, frames: int,
                     rods: int, particles: int, params, steps_per_frame: int, timing: dict, labels=()) -> None:
    """The scene sidecar of writeJson (scenes.cpp:126-160): frame rate and
    counts, solver settings, timing, labels and the collision primitives."""
    prims = []
    for p in world.primitives:
        prims.append({"type": _PRIM_NAMES.get(int(p.type), str(p.type)), "a": list(map(float, p.a)),
                      "b": list(map(float, p.b)), "normal": list(map(float, p.normal)), "radius": float(p.radius),
                      "half_extents": list(map(float, p.half_extents))})
    sim = float(timing.get("simulated_seconds", 0.0))
    wall = float(timing.get("wall_seconds", 0.0))
    seg = float(timing.get("segment_substeps", 0.0))
    doc = {
        "name": name, "description": description, "fps": float(fps), "frames": int(frames), "rods": int(rods),
        "particles": int(particles), "radius": float(radius),
        "solver": {"dt": float(params.dt), "substeps": int(params.substeps), "iterations": int(params.iterations),
                   "steps_per_frame": int(steps_per_frame)},
        "timing": {"simulated_seconds": sim, "wall_seconds": wall,
                   "realtime_factor": sim / wall if wall > 0 else 0.0, "segment_substeps": seg,
                   "segment_substeps_per_sec": seg / wall if wall > 0 else 0.0,
                   "threads": int(timing.get("threads", 1)), "device": str(timing.get("device", "MI355X"))},
        "labels": [{"text": t, "at": list(map(float, at))} for t, at in labels],
        "primitives": prims,
    }
    with open(path, "w") as f:
        json.dump(doc, f, indent=2)
