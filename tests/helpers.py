"""Shared test plumbing: scenes, oracle twins of product rods, comparisons.

The oracle (``oracle/``) is the checker here; the product path is
``ropephysics_amd`` over ``libropephysics.so``.
"""
from __future__ import annotations

import numpy as np

import oracle
import ropephysics_amd as rp

REFERENCE_MATERIAL = rp.RodMaterial(1e7, 0.35, 1000.0, 5e-3)  # validation/support.h:55-62


def benchmark_rod(n: int, length: float = 1.0) -> rp.Rod:
    """The reference's benchmark rod: straight along +x, root clamped
    (gpu_cases.cpp:27-33)."""
    rod = rp.make_straight_rod(n, length, REFERENCE_MATERIAL, (0, 0, 0), (1, 0, 0))
    rp.clamp_root_exact(rod)
    return rod


def benchmark_params(substeps: int = 8, iterations: int = 1) -> rp.SolverParams:
    return rp.SolverParams(dt=1e-3, substeps=substeps, iterations=iterations, gravity=(0.0, 0.0, -9.81))


def oracle_world(w: rp.CollisionWorld | None):
    if w is None:
        return None
    prims = [oracle.make_prim(p.type, p.a, p.b, p.normal, p.half_extents, p.rotation, p.radius, p.friction)
             for p in w.primitives]
    return oracle.World(prims=prims, self_collision=w.self_collision, self_friction=w.self_friction,
                        self_skip=w.self_skip, margin_radii=w.contact_margin_radii, hash_table_size=w.hash_table_size,
                        self_order=w.self_order)


def oracle_params(p: rp.SolverParams, coloured=True, compliance_scale=1.0, friction_scale=1.0, precision=32):
    """Step parameters for the oracle. In FP32 the batch's per-rod scales are
    the FP32 values the device holds (compliance scale = 1 / youngs in FP32)."""
    return oracle.Params(dt=p.dt, substeps=p.substeps, iterations=p.iterations, gravity=p.gravity,
                         lin_damp=p.linear_damping, ang_damp=p.angular_damping,
                         contact_interval=p.contact_interval, coloured=coloured,
                         compliance_scale=compliance_scale, friction_scale=friction_scale)


def compliance_scale_f32(youngs_scale: float) -> float:
    return float(np.float32(1.0) / np.float32(youngs_scale))


def oracle_twin(rod: rp.Rod, precision: int, kind: str = "oracle") -> oracle.Rod:
    """An oracle rod holding exactly `rod` (rounded to FP32 when precision=32)."""
    o = oracle.Rod.straight(2, 1.0, rod.material.as_tuple(), precision=precision, kind=kind)
    o.set(rod.arrays())
    return o


def bits(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def bit_equal(a, b) -> bool:
    return np.array_equal(bits(a), bits(b))


def max_abs(a, b) -> float:
    return float(np.max(np.abs(np.asarray(a, np.float64) - np.asarray(b, np.float64)))) if np.size(a) else 0.0


def max_pos_err(a, b) -> float:
    d = np.asarray(a, np.float64) - np.asarray(b, np.float64)
    return float(np.max(np.sqrt((d * d).sum(-1)))) if d.size else 0.0


def gpu_rod_state(batch: rp.Batch, index: int) -> dict:
    return batch.download(index)


def oracle_state(o: oracle.Rod) -> dict:
    a = o.get()
    return {"x": a["x"], "v": a["v"], "q": a["q"], "omega": a["omega"]}


def compare_bitwise(gpu: dict, ora: dict) -> dict:
    """Per field: number of FP32 words that differ."""
    return {k: int(np.count_nonzero(bits(gpu[k]) != bits(ora[k]))) for k in ("x", "v", "q", "omega")}
