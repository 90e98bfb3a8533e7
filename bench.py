#!/usr/bin/env python3
"""Throughput of the batched XPBD rod stepper on one MI355X.

Workload (default): the uniform batch of 4096 rods x 64 segments over a floor
plane with per-rod stiffness and friction, dt = 1e-3, 10 substeps, 1000 steps
(synthetic inputs from a fixed seed). A "step" is one step of every rod in the
batch. The TILED stepper is timed over --steps steps after --warmup untimed
ones; the PLAIN one-rod-per-workgroup launch (the reference's launch shape) is
timed on the same workload beside it, and the reference CPU path on a bounded
sample of it.

Prints one JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import ropephysics_amd as rp  # noqa: E402
from ropephysics_amd import scenes  # noqa: E402

METRIC = "segment-substeps/s, ragged-batched vs plain on 1 GPU; p50/p99 solve latency"

# Algorithmic FP32 operations per segment-substep (adds, multiplies, divides
# and square roots, each counted once) of the contact-free step with one sweep,
# counted by running the oracle with an operation-counting scalar over a
# clamped 64-segment rod for 10 steps of 10 substeps
# (tests/test_oracle_pins.py::test_flop_count_frozen checks this value).
FLOPS_PER_SEGMENT_SUBSTEP = 888.39

# Bytes the TILED kernel moves through LDS per segment-substep with one sweep
# and no contacts: predict 112 (x, force, kinematic velocity, orientation,
# torque read; x, q written), stretch 96, bend 64, finish 48.
LDS_BYTES_PER_SEGMENT_SUBSTEP = 320.0

PEAK_FP32_TFLOPS = 157.3   # MI355X published FP32 vector/matrix peak
PEAK_HBM_GBS = 8000.0
PEAK_LDS_GBS = 128.0 * 256 * 2.4  # 128 B/clk/CU x 256 CUs x 2.4 GHz


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_batch(w: scenes.Workload, strategy: str) -> rp.Batch:
    b = rp.Batch(w.prototype, w.num_rods, strategy=strategy, world=w.world)
    b.set_material_scales(w.youngs_scale, w.friction_scale)
    return b


def reset(b: rp.Batch, w: scenes.Workload):
    R = w.num_rods
    p = w.prototype
    b.set_state(0, R, x=w.x0, v=np.zeros_like(w.x0), q=np.tile(p.q, (R, 1)), omega=np.tile(p.omega, (R, 1)))


def time_steps(b: rp.Batch, params, steps: int):
    b.synchronize()
    t0 = time.perf_counter()
    b.step(params, steps)
    b.synchronize()
    t1 = time.perf_counter()
    c = b.counters()
    return t1 - t0, c


def make_ragged_batch(w: scenes.Workload, strategy: str) -> rp.Batch:
    b = rp.Batch(w.rods, strategy=strategy, world=w.world)
    b.set_material_scales(w.youngs_scale, w.friction_scale)
    if "paths" in w.info:
        b.set_paths(0, w.info["paths"])
    return b


def reset_ragged(b: rp.Batch, w: scenes.Workload):
    b.upload_all(w.rods)
    b.set_clock()


def ragged_comparison(steps: int, plain_steps: int, warmup: int) -> dict:
    """The ragged routing workload (1024 cables of 48-160 segments with
    self-collision, gripper paths on the device), TILED and PLAIN, each from
    the workload's initial state over its first `steps` / `plain_steps`
    steps."""
    w = scenes.routing_rollouts()
    log(f"ragged workload {w.name}: {w.num_rods} rods, {w.segments_total} segments")
    out = {"workload": w.name, "rods": w.num_rods, "segments_total": w.segments_total,
           "segments": "48-160 uniform", "self_collision": True, "substeps": w.params.substeps}
    for strategy, k in (("tiled", steps), ("plain", plain_steps)):
        b = make_ragged_batch(w, strategy)
        reset_ragged(b, w)
        b.step(w.params, min(warmup, 4))
        b.synchronize()
        reset_ragged(b, w)
        b.set_timing(True)
        secs, c = time_steps(b, w.params, k)
        work = float(w.segments_total) * w.params.substeps * k
        ent = {"value": work / secs, "steps": k, "seconds": secs,
               "kernel_ms_avg": c["last_step_gpu_ms"] / max(1, c["last_step_launches"]),
               "self_contact_overflow": c["self_contact_overflow"]}
        if strategy == "tiled":
            ent.update({"tile_config": c.get("tile_config_name"), "autotune_trials_ms": b.tune_report(),
                        "workgroups": c["workgroups"],
                        "lanes_busy": c["lanes_busy"], "lanes_total": c["lanes_total"],
                        "lds_bytes_per_workgroup": c["lds_bytes_per_workgroup"]})
            x = b.download_positions()
            if not np.all(np.isfinite(x)):
                raise RuntimeError("non-finite state in the ragged run")
        log(f"ragged {strategy}: {secs:.3f} s for {k} steps, {ent['value'] / 1e9:.3f} G seg-substeps/s")
        out[strategy] = ent
        b.destroy()
    out["speedup_vs_plain"] = out["tiled"]["value"] / out["plain"]["value"]
    return out


def cpu_baseline(w: scenes.Workload, stride: int, threads: int):
    """The reference CPU path (its own sources, built into oracle/_ref) on every
    `stride`-th rod of the workload for all its steps, FP64, on `threads` host
    threads with a static partition over rods."""
    import oracle  # the CPU checker/baseline; never part of the product path

    kind = "ref64" if oracle.available("ref64") else "oracle"
    sel = np.arange(0, w.num_rods, stride)
    nP = w.prototype.num_particles
    arrays = w.prototype.arrays()
    rods = []
    for r in sel:
        o = oracle.Rod.straight(2, 1.0, w.prototype.material.as_tuple(), precision=64, kind=kind)
        a = dict(arrays)
        a["x"] = w.x0[r * nP:(r + 1) * nP]
        a["xPrev"] = a["x"]
        a["sComp"] = arrays["sComp"] / float(w.youngs_scale[r])
        a["bComp"] = arrays["bComp"] / float(w.youngs_scale[r])
        o.set(a)
        rods.append(o)
    # Per-rod stiffness and friction, as the reference CPU applies them: the
    # rod's material and the world's friction. The scales are folded into
    # per-rod parameters here; the work per step is the same.
    worlds = []
    for r in sel:
        pr = w.world.primitives[0]
        worlds.append(oracle.World(prims=[oracle.make_prim(pr.type, pr.a, pr.b, pr.normal, pr.half_extents,
                                                           pr.rotation, pr.radius,
                                                           pr.friction * float(w.friction_scale[r]))]))
    p = w.params
    params = oracle.Params(dt=p.dt, substeps=p.substeps, iterations=p.iterations, gravity=p.gravity,
                           coloured=False)
    t0 = time.perf_counter()
    oracle.step_many(rods, params, worlds, w.steps, threads)
    dt = time.perf_counter() - t0
    work = len(sel) * w.prototype.num_segments * p.substeps * w.steps
    return {"value": work / dt, "unit": "segment-substeps/s", "cores": threads,
            "kind": "reference" if kind == "ref64" else "port",
            "sample": f"{len(sel)} of {w.num_rods} rods (every {stride}th), all {w.steps} steps, FP64, "
                      f"sequential sweep order, {threads} threads static partition; {dt:.1f} s",
            "seconds": dt}


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default: the workload's 1000)")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--plain-steps", type=int, default=None, help="PLAIN timed steps (default: --steps)")
    ap.add_argument("--rods", type=int, default=4096)
    ap.add_argument("--cpu-stride", type=int, default=16)
    ap.add_argument("--cpu-threads", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--tile-config", type=int, default=0)
    ap.add_argument("--ragged-steps", type=int, default=1000, help="steps of the ragged comparison (0: skip)")
    ap.add_argument("--ragged-plain-steps", type=int, default=200)
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="first run one device-to-device copy of this many MiB (a known-traffic kernel for "
                         "scaling hardware byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: every box here has one, and this path has no cross-GPU "
              f"exchange (--gpus {args.gpus} requested).", file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration_mib:
        import torch

        n = args.copy_calibration_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy calibration: {args.copy_calibration_mib} MiB read + {args.copy_calibration_mib} MiB written")
        del src, dst

    w = scenes.uniform_floor_batch(num_rods=args.rods)
    K = args.steps if args.steps is not None else w.steps
    W = args.warmup
    log(f"workload {w.name}: {w.num_rods} rods, {w.segments_total} segments, K={K}, W={W}")
    info = rp.device_info()
    log("device", info)

    # ---- TILED
    b = make_batch(w, "tiled")
    if args.tile_config:
        b.destroy()
        b = rp.Batch(w.prototype, w.num_rods, strategy="tiled", world=w.world, tile_config=args.tile_config)
        b.set_material_scales(w.youngs_scale, w.friction_scale)
    reset(b, w)
    b.step(w.params, W)
    b.synchronize()
    reset(b, w)
    b.set_timing(True)
    secs, c = time_steps(b, w.params, K)
    x = b.download_positions()
    if not np.all(np.isfinite(x)):
        raise RuntimeError("non-finite state after the timed run")
    work = float(w.segments_total) * w.params.substeps * K
    rate = work / secs
    gpu_ms = c["last_step_gpu_ms"]
    launches = max(1, c["last_step_launches"])
    kernel_ms = gpu_ms / launches
    work_per_launch = work / launches
    kernel_rate = work_per_launch / (kernel_ms * 1e-3)
    log(f"TILED {c.get('tile_config_name')}: {secs:.3f} s wall, {gpu_ms:.2f} ms device over {launches} launches, "
        f"{rate / 1e9:.3f} G seg-substeps/s; lanes {c['lanes_busy']}/{c['lanes_total']}, "
        f"{c['workgroups']} workgroups, LDS {c['lds_bytes_per_workgroup']} B")
    tiled_info = {"tile_config": c.get("tile_config_name"), "workgroups": c["workgroups"],
                  "lanes_busy": c["lanes_busy"], "lanes_total": c["lanes_total"],
                  "lds_bytes_per_workgroup": c["lds_bytes_per_workgroup"], "launches": launches,
                  "kernel_ms_avg": kernel_ms, "autotune_trials_ms": b.tune_report()}
    b.destroy()

    # ---- PLAIN, same workload
    plain = None
    if not args.no_plain:
        Kp = args.plain_steps if args.plain_steps is not None else K
        pb = make_batch(w, "plain")
        reset(pb, w)
        pb.step(w.params, min(W, 4))
        pb.synchronize()
        reset(pb, w)
        pb.set_timing(True)
        psecs, pc = time_steps(pb, w.params, Kp)
        pwork = float(w.segments_total) * w.params.substeps * Kp
        prate = pwork / psecs
        plain = {"value": prate, "steps": Kp, "seconds": psecs, "ms_per_step": psecs * 1e3 / Kp,
                 "kernel_ms_avg": pc["last_step_gpu_ms"] / max(1, pc["last_step_launches"])}
        log(f"PLAIN: {psecs:.3f} s for {Kp} steps, {prate / 1e9:.3f} G seg-substeps/s")
        pb.destroy()

    # ---- roofline of the dominant kernel (k_tiled_step)
    achieved_tflops = FLOPS_PER_SEGMENT_SUBSTEP * kernel_rate / 1e12
    roofline = {
        "bound": "mfma",
        "bound_detail": "FP32 vector-ALU bound; the path has no dense contraction and issues no MFMA, so "
                        "the FP32 peak is the compute roofline (unpacked FP32 tops out at half of it)",
        "achieved": achieved_tflops, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
        "frac": achieved_tflops / PEAK_FP32_TFLOPS, "traffic": None,
        "kernel": "k_tiled_step", "flops_per_segment_substep": FLOPS_PER_SEGMENT_SUBSTEP,
        "segment_substeps_per_launch": work_per_launch, "kernel_ms_avg": kernel_ms,
        "lds": {"bytes_per_segment_substep": LDS_BYTES_PER_SEGMENT_SUBSTEP,
                "achieved_GBs": LDS_BYTES_PER_SEGMENT_SUBSTEP * kernel_rate / 1e9, "peak_GBs": PEAK_LDS_GBS,
                "frac": LDS_BYTES_PER_SEGMENT_SUBSTEP * kernel_rate / 1e9 / PEAK_LDS_GBS},
    }
    traffic_path = os.path.join(HERE, "profiles", "k_tiled_step_uniform-4096x64-floor_traffic.json")
    if os.path.exists(traffic_path) and not args.tile_config:
        with open(traffic_path) as f:
            tr = json.load(f)
        roofline["traffic"] = tr.get("hbm_bytes_per_launch")
        roofline["traffic_source"] = os.path.relpath(traffic_path, HERE)

    ragged = None
    if args.ragged_steps > 0:
        ragged = ragged_comparison(args.ragged_steps, args.ragged_plain_steps, W)

    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(w, args.cpu_stride, threads)
        log(f"CPU baseline: {cpu['value'] / 1e6:.2f} M seg-substeps/s ({cpu['sample']})")

    out = {
        "metric": METRIC, "value": rate, "unit": "segment-substeps/s", "n_gpus": 1, "steps": K, "warmup": W,
        "ms_per_step": secs * 1e3 / K, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (seeded)",
        "config": {"workload": w.name, **w.info, "strategy": "tiled", "parallelism": "single GPU"},
        "tiled": tiled_info,
        "plain": plain,
        "speedup_vs_plain": (rate / plain["value"]) if plain else None,
        "ragged": ragged,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "device": {"arch": info["arch"], "compute_units": info["compute_units"]},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
