#!/usr/bin/env python3
"""Throughput of the batched XPBD rod stepper on one MI355X.

Headline workload: the ragged batch of 1024 cable-routing rollouts (cables of
48-160 segments, 1.2 m, a gripper path per cable evaluated on the device,
floor, two capsule pegs and a box clip, self-collision on, dt = 1e-3, 10
substeps, the whole 6 s = 6000 steps; synthetic inputs from a fixed seed). A
"step" is one step of every rod in the batch. The ragged-packed TILED stepper
is timed over --steps steps after --warmup untimed ones (the autotuner runs
during the warm-up), starting again from step 0; `config.timed_steps` states
the window. The PLAIN one-rod-per-workgroup launch (the reference's launch
shape) is timed over the same steps when there are at most 400 of them.

Whatever --steps is, the full-rollout leg runs the whole 6000 steps: TILED in
FP32 and in FP64 (one launch each; routed verdicts; `roofline` comes from the
FP32 launch), and PLAIN over four 100-step windows spread over the rollout,
each started from TILED's state at that step (bitwise PLAIN's own) and checked
bit for bit at its end. The routed verdicts are compared with the reference
CPU path's (its own sources, FP64) on a 64-cable sample, in sequential and in
colour order; the same sample, sequential, is the all-core CPU baseline, and a
smaller one the 1-thread baseline.

Secondary: the routing rollout with the coloured self-contact order and with
trajectory recording; 512 rods of 2000 segments (beyond LDS, the GLOBAL
strategy); the uniform batch of 4096 rods x 64 segments over a floor plane
with per-rod stiffness and friction, 1000 steps, TILED and PLAIN; the
hold/slide sweep of 3600 draped cables (8 s) in FP32 and FP64 with its
verdicts spot-checked against the reference CPU path; the cross-entropy
harness-learning loop (10 iterations of 1024 rollouts); and the serving
stream (requests of 1-256 rods of 32-256 segments, 0.5-8 s horizons, random
contact shapes, self-collision in half) through the continuous-batching
server: the capacity of each mode (64-step chunks in lockstep; 50 ms time
slices in which each workgroup runs the steps it can) on 40 requests arriving
at once, then 200 requests arriving as a Poisson stream at 80% of the
time-sliced mode's capacity through the time-sliced server (the headline
p50/p99) and the first 40 of them, at the same rate, through the chunked one.

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

# Algorithmic FP32 operations (adds, multiplies, divides and square roots, each
# counted once) per segment-substep, counted by running the oracle with an
# operation-counting scalar (tests/test_oracle_pins.py::test_flop_counts_frozen
# re-derives both):
#  - contact-free step, one sweep: a clamped 64-segment rod, 10 steps of 10
#    substeps;
#  - cable routing: every 64th cable of the routing workload from its initial
#    pose, 10 steps of 10 substeps, with the routing world (floor, two pegs,
#    clip: contact generation and projection) and self-collision (spatial hash,
#    closest points of every candidate pair).
FLOPS_PER_SEGMENT_SUBSTEP = 888.39
FLOPS_PER_SEGMENT_SUBSTEP_ROUTING = 1483.08

# Bytes the TILED kernel moves through LDS per segment-substep with one sweep
# and no contacts: predict 112 (x, force, kinematic velocity, orientation,
# torque read; x, q written), stretch 96, bend 64, finish 48.
LDS_BYTES_PER_SEGMENT_SUBSTEP = 320.0

PEAK_FP32_TFLOPS = 157.3   # MI355X published FP32 vector/matrix peak
PEAK_HBM_GBS = 8000.0
PEAK_LDS_GBS = 128.0 * 256 * 2.4  # 128 B/clk/CU x 256 CUs x 2.4 GHz

TRAFFIC = {
    "routing": os.path.join(HERE, "profiles", "k_tiled_step_routing-1024-ragged_traffic.json"),
    "uniform": os.path.join(HERE, "profiles", "k_tiled_step_uniform-4096x64-floor_traffic.json"),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def time_steps(b: rp.Batch, params, steps: int):
    b.synchronize()
    t0 = time.perf_counter()
    b.step(params, steps)
    b.synchronize()
    t1 = time.perf_counter()
    return t1 - t0, b.counters()


def roofline(flops_per_unit: float, kernel_rate: float, kernel_ms: float, units_per_launch: float, traffic_key):
    achieved = flops_per_unit * kernel_rate / 1e12
    out = {
        "bound": "valu",
        "bound_detail": "FP32 vector-ALU bound: the path has no dense contraction and issues no MFMA; the peak is "
                        "MI355X's FP32 vector figure. Bitwise parity with the CPU rules out FMA contraction and "
                        "packed FP32, and division and square root are IEEE sequences, so unpacked non-FMA issue "
                        "(16 lanes x 4 SIMDs x 256 CUs x 2.4 GHz = 39.3e12 lane instructions/s) caps frac near 0.25",
        "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS,
        "traffic": None, "kernel": "k_tiled_step", "flops_per_segment_substep": flops_per_unit,
        "segment_substeps_per_launch": units_per_launch, "kernel_ms_avg": kernel_ms,
        "lds": {"bytes_per_segment_substep_contact_free": LDS_BYTES_PER_SEGMENT_SUBSTEP,
                "achieved_GBs_contact_free_share": LDS_BYTES_PER_SEGMENT_SUBSTEP * kernel_rate / 1e9,
                "peak_GBs": PEAK_LDS_GBS},
    }
    path = TRAFFIC.get(traffic_key)
    if path and os.path.exists(path):
        with open(path) as f:
            tr = json.load(f)
        out["traffic"] = tr.get("hbm_bytes_per_launch")
        out["traffic_source"] = os.path.relpath(path, HERE)
        out["traffic_launch_steps"] = tr.get("launch", {}).get("steps_per_launch")
    return out


# ------------------------------------------------------------------ ragged routing (headline)

def make_ragged_batch(w: scenes.Workload, strategy: str, tile_config: int = 0, precision: int = 32) -> rp.Batch:
    b = rp.Batch(w.rods, strategy=strategy, world=w.world, tile_config=tile_config, precision=precision)
    b.set_material_scales(w.youngs_scale, w.friction_scale)
    b.set_paths(0, w.info["paths"])
    return b


def reset_ragged(b: rp.Batch, w: scenes.Workload):
    b.upload_all(w.rods)
    b.set_clock()


def routing_leg(w: scenes.Workload, steps: int, plain_steps: int, warmup: int, tile_config: int,
                no_plain: bool) -> dict:
    out = {}
    for strategy, k in (("tiled", steps), ("plain", plain_steps)):
        if strategy == "plain" and (no_plain or k <= 0):
            continue
        b = make_ragged_batch(w, strategy, tile_config if strategy == "tiled" else 0)
        reset_ragged(b, w)
        b.step(w.params, warmup if strategy == "tiled" else min(warmup, 4))
        b.synchronize()
        reset_ragged(b, w)
        b.set_timing(True)
        secs, c = time_steps(b, w.params, k)
        work = float(w.segments_total) * w.params.substeps * k
        launches = max(1, c["last_step_launches"])
        ent = {"value": work / secs, "steps": k, "seconds": secs, "ms_per_step": secs * 1e3 / k,
               "device_ms": c["last_step_gpu_ms"], "launches": launches,
               "kernel_ms_avg": c["last_step_gpu_ms"] / launches,
               "self_contact_overflow": c["self_contact_overflow"]}
        if strategy == "tiled":
            ent.update({"tile_config": c.get("tile_config_name"), "autotune_trials_ms": b.tune_report(),
                        "workgroups": c["workgroups"], "lanes_busy": c["lanes_busy"],
                        "lanes_total": c["lanes_total"], "lds_bytes_per_workgroup": c["lds_bytes_per_workgroup"]})
            x = b.download_positions()
            if not np.all(np.isfinite(x)):
                raise RuntimeError("non-finite state after the timed routing run")
        log(f"routing {strategy} {ent.get('tile_config', '')}: {secs:.3f} s for {k} steps, "
            f"{ent['value'] / 1e9:.3f} G seg-substeps/s, device {c['last_step_gpu_ms']:.1f} ms over "
            f"{launches} launches")
        out[strategy] = ent
        b.destroy()
    return out


PLAIN_WINDOW_STARTS = (0, 1500, 3000, 4500)
PLAIN_WINDOW_STEPS = 100


def full_rollout_leg(w: scenes.Workload, precision: int, warmup: int, tile_config: int = 0,
                     plain_windows: bool = True) -> tuple:
    """The whole 6000-step routing rollout, whatever --steps is. TILED runs it
    as one call (one launch), timed by wall clock and by HIP events on the
    batch's stream; routed verdicts come from its final positions. PLAIN (one
    launch per step) runs four stated windows of 100 steps spread over the
    rollout, each started from the state a second TILED batch reaches at that
    step (the two strategies are bitwise equal, so this is PLAIN's own state
    there; each window's end state is checked against TILED's bit for bit).
    PLAIN's rollout rate is taken as its rate over the windows."""
    steps = w.steps
    task = w.info["task"]
    b = make_ragged_batch(w, "tiled", tile_config, precision)
    reset_ragged(b, w)
    b.step(w.params, warmup)
    b.synchronize()
    reset_ragged(b, w)
    b.set_timing(True)
    secs, c = time_steps(b, w.params, steps)
    work = float(w.segments_total) * w.params.substeps * steps
    launches = max(1, c["last_step_launches"])
    kernel_ms = c["last_step_gpu_ms"] / launches
    x = b.get_state()["x"]
    ps = b.particle_start
    if not np.all(np.isfinite(x)):
        raise RuntimeError("non-finite state after the routing rollout")
    verdicts = [task.evaluate(x[ps[i]:ps[i + 1]].astype(np.float64))["routed"] for i in range(w.num_rods)]
    cfg = c["tile_config"]
    tiled = {"value": work / secs, "unit": "segment-substeps/s", "steps": steps, "seconds": secs,
             "device_ms": c["last_step_gpu_ms"], "launches": launches, "kernel_ms_avg": kernel_ms,
             "segment_substeps_per_launch": work / launches, "tile_config": c.get("tile_config_name"),
             "workgroups": c["workgroups"], "self_contact_overflow": c["self_contact_overflow"],
             "motion_margin": c["motion_margin"], "low_margin_steps": c["low_margin_steps"],
             "routed": int(sum(verdicts))}
    log(f"routing rollout FP{precision} TILED {tiled['tile_config']}: {secs:.3f} s for {steps} steps in {launches} "
        f"launch(es), {tiled['value'] / 1e9:.3f} G seg-substeps/s, routed {tiled['routed']} of {w.num_rods}")
    b.destroy()
    out = {"dtype": f"f{precision}", "tiled": tiled}
    if plain_windows:
        snap = make_ragged_batch(w, "tiled", cfg, precision)
        pl = make_ragged_batch(w, "plain", 0, precision)
        reset_ragged(pl, w)
        pl.step(w.params, PLAIN_WINDOW_STEPS)  # graph capture and code objects
        pl.synchronize()
        pl.set_timing(True)
        reset_ragged(snap, w)
        done, plain_secs, plain_dev, equal = 0, 0.0, 0.0, True
        for s0 in PLAIN_WINDOW_STARTS:
            snap.step(w.params, s0 - done)
            st = snap.get_state()
            pl.set_state(0, w.num_rods, **st)
            pl.set_clock(clock=snap.get_clock())
            wsecs, wc = time_steps(pl, w.params, PLAIN_WINDOW_STEPS)
            plain_secs += wsecs
            plain_dev += wc["last_step_gpu_ms"]
            snap.step(w.params, PLAIN_WINDOW_STEPS)
            done = s0 + PLAIN_WINDOW_STEPS
            a, e = pl.get_state(), snap.get_state()
            equal = equal and all(np.array_equal(np.ascontiguousarray(a[k]).view(np.uint8),
                                                 np.ascontiguousarray(e[k]).view(np.uint8)) for k in a)
        wwork = float(w.segments_total) * w.params.substeps * PLAIN_WINDOW_STEPS * len(PLAIN_WINDOW_STARTS)
        out["plain_windows"] = {
            "value": wwork / plain_secs, "unit": "segment-substeps/s",
            "windows": [f"{s0}-{s0 + PLAIN_WINDOW_STEPS - 1}" for s0 in PLAIN_WINDOW_STARTS],
            "steps": PLAIN_WINDOW_STEPS * len(PLAIN_WINDOW_STARTS), "seconds": plain_secs,
            "device_ms": plain_dev, "bitwise_equal_tiled_at_window_ends": bool(equal)}
        out["speedup_vs_plain"] = tiled["value"] / out["plain_windows"]["value"]
        log(f"routing FP{precision} PLAIN over windows {out['plain_windows']['windows']}: "
            f"{out['plain_windows']['value'] / 1e9:.4f} G seg-substeps/s; TILED rollout / PLAIN "
            f"{out['speedup_vs_plain']:.1f}x; window ends bitwise equal to TILED: {equal}")
        snap.destroy()
        pl.destroy()
    return out, verdicts


def trajectory_leg(w: scenes.Workload, every: int, warmup: int) -> dict:
    """The routing workload recorded on the device every `every` steps (the
    reference's trajectory frames, src/validation/scenes.cpp:13-21, taken by the
    stepping kernel instead of a download per frame): the same full rollout
    timed without and with recording, the frames' size and download time, and
    the last frame checked bitwise against the final positions."""
    from ropephysics_amd import trajectory

    b = make_ragged_batch(w, "tiled")
    reset_ragged(b, w)
    b.step(w.params, warmup)
    b.synchronize()
    steps = w.steps
    frames = steps // every
    work = float(w.segments_total) * w.params.substeps * steps
    reset_ragged(b, w)
    plain_secs, _ = time_steps(b, w.params, steps)
    reset_ragged(b, w)
    b.record(every, frames)
    rec_secs, c = time_steps(b, w.params, steps)
    assert b.recorded_frames() == frames
    t0 = time.perf_counter()
    fr = b.frames(0, frames)
    dl = time.perf_counter() - t0
    last_ok = bool(np.array_equal(fr[-1].view(np.uint32), b.download_positions().astype(np.float32).view(np.uint32)))
    groups = trajectory.split_by_size(fr[:1], b.particle_start)
    ent = {"every_steps": every, "fps": 1.0 / (every * w.params.dt), "frames": frames,
           "frame_bytes": int(fr[0].nbytes), "buffer_bytes": int(fr.nbytes),
           "value_without": work / plain_secs, "value_with": work / rec_secs,
           "overhead": rec_secs / plain_secs - 1.0, "download_seconds": dl,
           "download_GBs": fr.nbytes / dl / 1e9, "last_frame_equals_final_positions": last_ok,
           "rodtraj_files": len(groups), "unit": "segment-substeps/s"}
    log(f"trajectory: {frames} frames every {every} steps ({fr.nbytes / 1e6:.0f} MB on the device): "
        f"{ent['value_with'] / 1e9:.3f} G seg-substeps/s recording vs {ent['value_without'] / 1e9:.3f} G without "
        f"({100 * ent['overhead']:+.1f}%); download {dl * 1e3:.0f} ms; last frame equals final positions: {last_ok}")
    b.record(0, 0)
    b.destroy()
    return ent


def routing_coloured_leg(w: scenes.Workload, steps: int, warmup: int, gv) -> tuple:
    """The routing workload with every cable's self contacts in the coloured
    order (an opt-in of this build; the reference projects them in list
    order): throughput, the longest launches, and routed verdicts against the
    list-order FP32 rollout's."""
    import dataclasses

    world = dataclasses.replace(w.world, self_order=1)
    b = rp.Batch(w.rods, strategy="tiled", world=world)
    b.set_material_scales(w.youngs_scale, w.friction_scale)
    b.set_paths(0, w.info["paths"])
    reset_ragged(b, w)
    b.step(w.params, warmup)
    b.synchronize()
    reset_ragged(b, w)
    b.set_timing(True)
    secs, c = time_steps(b, w.params, steps)
    work = float(w.segments_total) * w.params.substeps * steps
    launches = max(1, c["last_step_launches"])
    ent = {"value": work / secs, "unit": "segment-substeps/s", "dtype": "f32", "steps": steps, "seconds": secs,
           "kernel_ms_avg": c["last_step_gpu_ms"] / launches, "launches": launches,
           "tile_config": c.get("tile_config_name"), "self_contact_overflow": c["self_contact_overflow"]}
    verdicts = None
    if steps == w.steps:
        task = w.info["task"]
        verdicts = [task.evaluate(b.download(i)["x"])["routed"] for i in range(w.num_rods)]
        ent["routed"] = int(sum(verdicts))
        if gv is not None:
            ent["verdicts_vs_list_order_gpu"] = {"cables": w.num_rods,
                                                 "agree": int(sum(1 for a, g in zip(verdicts, gv) if a == g))}
    log(f"routing, coloured self order {ent['tile_config']}: {secs:.3f} s for {steps} steps, "
        f"{ent['value'] / 1e9:.3f} G seg-substeps/s" +
        (f"; verdicts equal to the list order's on {ent['verdicts_vs_list_order_gpu']['agree']} of {w.num_rods}"
         if "verdicts_vs_list_order_gpu" in ent else ""))
    b.destroy()
    return ent, verdicts


def cpu_baseline_routing(w: scenes.Workload, stride: int, steps: int, threads: int, coloured: bool = False) -> dict:
    """The reference CPU path (its own sources, built into oracle/_ref) on every
    `stride`-th cable of the routing workload over its first `steps` steps, FP64,
    the gripper schedule applied before each step, on `threads` host threads
    with a static partition over rods; in the reference's default sequential
    sweep order, or in colour order (the order the GPU sweeps in)."""
    import oracle  # the CPU checker/baseline; never part of the product path

    kind = "ref64" if oracle.available("ref64") else "oracle"
    sel = np.arange(0, w.num_rods, stride)
    paths = w.info["paths"]
    rods, worlds = [], []
    p = w.params
    for r in sel:
        rod = w.rods[r]
        o = oracle.Rod.straight(2, 1.0, rod.material.as_tuple(), precision=64, kind=kind)
        a = rod.arrays()
        a["sComp"] = a["sComp"] / float(w.youngs_scale[r])
        a["bComp"] = a["bComp"] / float(w.youngs_scale[r])
        o.set(a)
        rods.append(o)
        prims = [oracle.make_prim(pr.type, pr.a, pr.b, pr.normal, pr.half_extents, pr.rotation, pr.radius,
                                  pr.friction * float(w.friction_scale[r])) for pr in w.world.primitives]
        worlds.append(oracle.World(prims=prims, self_collision=True,
                                   self_friction=w.world.self_friction * float(w.friction_scale[r]),
                                   self_skip=w.world.self_skip, hash_table_size=w.world.hash_table_size))
    kv = np.array([[paths[r].velocity(s) for s in range(steps)] for r in sel])
    params = oracle.Params(dt=p.dt, substeps=p.substeps, iterations=p.iterations, gravity=p.gravity,
                           coloured=coloured)
    t0 = time.perf_counter()
    oracle.step_many_driven(rods, params, worlds, steps, threads, [paths[r].particle for r in sel], kv)
    dt = time.perf_counter() - t0
    segs = sum(w.rods[r].num_segments for r in sel)
    work = segs * p.substeps * steps
    verdicts = None
    if steps == w.steps:
        task = w.info["task"]
        verdicts = {int(r): task.evaluate(o.get()["x"])["routed"] for r, o in zip(sel, rods)}
    return {"value": work / dt, "unit": "segment-substeps/s", "cores": threads, "_verdicts": verdicts,
            "kind": "reference" if kind == "ref64" else "port",
            "sample": f"{len(sel)} of {w.num_rods} cables (every {stride}th, {segs} segments), first {steps} of "
                      f"{w.steps} steps, gripper schedule applied per step, FP64, "
                      f"{'colour' if coloured else 'sequential'} sweep order, "
                      f"{threads} threads static partition over rods; {dt:.1f} s",
            "seconds": dt}


def cpu_model() -> str:
    """The host CPU's model name and logical core count."""
    name = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    name = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return f"{name} ({os.cpu_count()} logical CPUs visible)"


# ------------------------------------------------------------------ long rods (GLOBAL strategy)

def global_leg(num_rods: int, warmup: int, segments: int = 2000, steps: int = 200) -> dict:
    """Rods too long for a workgroup's LDS: `num_rods` clamped rods of 2000
    segments (4 m) over a floor with friction, gravity, dt = 1e-3, 10
    substeps, stepped by the GLOBAL strategy (working arrays in HBM, one rod
    per workgroup, many steps per launch). The reference runs such rods with
    its multi-kernel strategy only."""
    mat = scenes.REFERENCE_MATERIAL
    rod = rp.make_straight_rod(segments, 4.0, mat, (0.0, 0.0, 0.1), (1.0, 0.0, 0.0))
    rp.clamp_root_exact(rod)
    world = rp.CollisionWorld(primitives=[rp.Primitive.plane((0.0, 0.0, 0.0), (0.0, 0.0, 1.0), 0.6)])
    params = rp.SolverParams(dt=1e-3, substeps=10, iterations=1, gravity=(0.0, 0.0, -9.81))
    b = rp.Batch(rod, num_rods, strategy="auto", world=world)
    b.step(params, warmup)
    b.synchronize()
    b.upload(rod)
    b.set_clock()
    b.set_timing(True)
    secs, c = time_steps(b, params, steps)
    work = float(num_rods) * segments * params.substeps * steps
    launches = max(1, c["last_step_launches"])
    out = {"config": {"rods": num_rods, "segments": segments, "length_m": 4.0, "dt": 1e-3, "substeps": 10,
                      "steps": steps, "contacts": "floor plane, Coulomb friction"},
           "strategy": c["strategy"], "value": work / secs, "unit": "segment-substeps/s", "seconds": secs,
           "kernel_ms_avg": c["last_step_gpu_ms"] / launches, "launches": launches,
           "scratch_bytes": c["scratch_bytes"], "lanes_per_workgroup": c["lanes_total"] // max(1, c["workgroups"]),
           "lds_bytes_per_workgroup": c["lds_bytes_per_workgroup"]}
    log(f"long rods ({c['strategy']}): {num_rods} x {segments} segments, {steps} steps in {launches} launches, "
        f"{out['value'] / 1e9:.3f} G seg-substeps/s, {c['scratch_bytes'] / 1e6:.0f} MB of HBM working arrays, "
        f"{c['lds_bytes_per_workgroup']} B of LDS per workgroup (positions and orientations)")
    b.destroy()
    return out


# ------------------------------------------------------------------ serving stream (latency)

def serving_leg(num: int, load: float, budget_gib: float, slice_us: int, capacity_num: int = 40,
                tune_num: int = 16, tune_steps: int = 512) -> dict:
    """The serving workload through the server in its two modes: chunked
    steps (64 steps per server step, every resident request in lockstep) and
    time-sliced steps (`slice_us` per server step, each workgroup running the
    steps it starts in the slice). First each mode's capacity with the first
    `capacity_num` requests arriving at once; then `num` requests arriving as
    a Poisson stream at `load` times the time-sliced mode's capacity through
    the time-sliced server (the headline p50/p99), and the first
    `capacity_num` of the same stream at the same rate through the chunked one.
    Latency is from arrival to result on the host. The time-sliced server's
    tile shapes come from `serving.autotune_shapes` on the first `tune_num`
    requests cut to `tune_steps` steps (0 requests: the server's defaults).
    The chunked server keeps its defaults: its throughput is set by the
    full-horizon heavy requests a short sample leaves out (tuned on one, it
    picked a pair 4% slower over the whole capacity run)."""
    from ropephysics_amd import serving

    reqs = serving.make_requests(num)
    arrivals = serving.poisson_arrivals(num, 1.0)  # scaled to the rate below
    horizon = np.array([q.steps * serving.PARAMS.dt for q in reqs])
    head = reqs[:capacity_num]
    work_sat = sum(q.work for q in head)
    out = {"requests": num, "share_of_stream": num / 10000.0,
           "segments_mean": float(np.mean([q.segments for q in reqs])),
           "rods_mean": float(np.mean([q.info["rods"] for q in reqs])),
           "horizon_s_mean": float(horizon.mean()), "self_collision_share": float(np.mean(
               [q.info["self_collision"] for q in reqs])), "load": load, "capacity_requests": capacity_num,
           "rate_basis": "load x the time-sliced mode's saturated capacity"}
    modes = (("chunked", {}), ("time_sliced", {"time_slice_us": slice_us}))
    servers, capacity = {}, {}
    budget = int(budget_gib * (1 << 30))
    for mode, kw in modes:
        tune = None
        if tune_num > 0 and "time_slice_us" in kw:
            best, trials = serving.autotune_shapes(head[:tune_num], budget, time_slice_us=kw.get("time_slice_us", 0),
                                                   steps=tune_steps)
            name = rp.tile_configs()
            tune = {"sample_requests": min(tune_num, len(head)), "steps": tune_steps,
                    "trials_s": [[name[a], name[b], t] for (a, b), t in trials],
                    "chosen": [name[best[0]], name[best[1]]]}
            kw = dict(kw, tile_config=best[0], self_tile_config=best[1])
            log(f"serving {mode}: shapes {name[best[0]]} / {name[best[1]]} (autotuned on {tune['sample_requests']} "
                f"requests of {tune_steps} steps)")
        server = rp.Server(serving.PARAMS, hbm_budget_bytes=budget, **kw)
        serving.run_stream(server, serving.make_requests(2, seed=5), np.zeros(2))  # code objects, first launches
        sat = serving.run_stream(server, head, np.zeros(len(head)))
        capacity[mode] = len(head) / sat["wall"]
        servers[mode] = server
        out["hbm_budget_bytes"] = server.stats()["device_bytes"]
        out[mode] = {"saturated": {"requests": len(head), "wall_s": sat["wall"], "value": work_sat / sat["wall"],
                                   "unit": "segment-substeps/s", "requests_per_s": capacity[mode],
                                   "launches": sat["launches"]}}
        if "time_slice_us" in kw:
            out[mode]["time_slice_us"] = slice_us
        if tune is not None:
            out[mode]["autotune"] = tune
        log(f"serving {mode}: saturated on {len(head)} requests {work_sat / sat['wall'] / 1e9:.3f} G "
            f"seg-substeps/s ({capacity[mode]:.2f} req/s)")
    rate = load * capacity["time_sliced"]
    for mode, _ in modes:
        n = num if mode == "time_sliced" else len(head)
        run = serving.run_stream(servers[mode], reqs[:n], arrivals[:n] / rate)
        lat = run["latency"]
        work = sum(q.work for q in reqs[:n])
        out[mode]["poisson"] = {
            "requests": n, "rate_requests_per_s": rate, "load_of_own_capacity": rate / capacity[mode],
            "wall_s": run["wall"], "value": work / run["wall"], "unit": "segment-substeps/s",
            "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
            "max_latency_s": float(lat.max()), "p50_latency_over_horizon": float(np.percentile(lat / horizon[:n], 50)),
            "max_resident_requests": run["max_resident"], "launches": run["launches"]}
        log(f"serving {mode}: {n} requests Poisson at {rate:.2f} req/s ({100 * rate / capacity[mode]:.0f}% of its "
            f"capacity): p50 {out[mode]['poisson']['p50_latency_s']:.3f} s, p99 "
            f"{out[mode]['poisson']['p99_latency_s']:.3f} s, {work / run['wall'] / 1e9:.3f} G seg-substeps/s")
        servers[mode].destroy()
    return out


# ------------------------------------------------------------------ uniform batch (secondary)

def uniform_leg(steps: int, warmup: int, no_plain: bool) -> dict:
    w = scenes.uniform_floor_batch(steps=steps)
    R, p = w.num_rods, w.prototype

    def make(strategy):
        b = rp.Batch(p, R, strategy=strategy, world=w.world)
        b.set_material_scales(w.youngs_scale, w.friction_scale)
        return b

    def reset(b):
        b.set_state(0, R, x=w.x0, v=np.zeros_like(w.x0), q=np.tile(p.q, (R, 1)), omega=np.tile(p.omega, (R, 1)))

    out = {"config": {"workload": w.name, **w.info}}
    for strategy in ("tiled", "plain"):
        if strategy == "plain" and no_plain:
            continue
        b = make(strategy)
        reset(b)
        b.step(w.params, warmup if strategy == "tiled" else min(warmup, 4))
        b.synchronize()
        reset(b)
        b.set_timing(True)
        secs, c = time_steps(b, w.params, steps)
        work = float(w.segments_total) * w.params.substeps * steps
        launches = max(1, c["last_step_launches"])
        kernel_ms = c["last_step_gpu_ms"] / launches
        ent = {"value": work / secs, "steps": steps, "seconds": secs, "kernel_ms_avg": kernel_ms,
               "launches": launches}
        if strategy == "tiled":
            ent.update({"tile_config": c.get("tile_config_name"), "workgroups": c["workgroups"],
                        "lds_bytes_per_workgroup": c["lds_bytes_per_workgroup"]})
            ent["roofline"] = roofline(FLOPS_PER_SEGMENT_SUBSTEP, (work / launches) / (kernel_ms * 1e-3),
                                       kernel_ms, work / launches, "uniform")
        log(f"uniform {strategy}: {secs:.3f} s for {steps} steps, {ent['value'] / 1e9:.3f} G seg-substeps/s")
        out[strategy] = ent
        b.destroy()
    if "plain" in out:
        out["speedup_vs_plain"] = out["tiled"]["value"] / out["plain"]["value"]
    return out


# ------------------------------------------------------------------ harness learning (application)

def cem_leg(iterations: int, population: int, cpu: bool) -> dict:
    """The cross-entropy harness-learning loop (runHarnessLearning,
    application_cases.cpp:353-486) on the routing task: `iterations` batches of
    `population` rollouts of 6 s, then the learned and the hand-written motion
    on the same fresh cables; with `cpu`, the softest and stiffest fresh cable
    under each motion rerun by the reference CPU path (FP64, colour order) and
    the verdicts compared."""
    from ropephysics_amd.learning import HarnessLearner

    t0 = time.perf_counter()
    learner = HarnessLearner(population=population)
    res = learner.run(iterations)
    wall = time.perf_counter() - t0
    out = {"population": population, "iterations": [
        {"success_rate": i.success_rate, "mean_score": i.mean_score, "seconds": i.seconds}
        for i in res.iterations], "learn_seconds": res.learn_seconds, "wall_seconds": wall,
        "learned_rate": res.learned_rate, "hand_rate": res.hand_rate,
        "learned_waypoints": res.mean.tolist()}
    if cpu:
        import oracle

        kind = "ref64" if oracle.available("ref64") else "oracle"
        task = learner.task
        ys = res.fresh["youngs"]
        picks = [int(np.argmin(ys)), int(np.argmax(ys))]
        agree = compared = 0
        for r in picks:
            n = int(res.fresh["segments"][r])
            for name, wp in (("learned", res.mean), ("hand", task.hand_written_waypoints())):
                rod = task.build(n)
                o = oracle.Rod.straight(2, 1.0, rod.material.as_tuple(), precision=64, kind=kind)
                arr = rod.arrays()
                arr["sComp"] = arr["sComp"] / float(ys[r])
                arr["bComp"] = arr["bComp"] / float(ys[r])
                o.set(arr)
                fr = float(res.fresh["friction"][r])
                prims = [oracle.make_prim(pr.type, pr.a, pr.b, pr.normal, pr.half_extents, pr.rotation, pr.radius,
                                          pr.friction * fr) for pr in task.world().primitives]
                w = task.world()
                world = oracle.World(prims=prims, self_collision=True, self_friction=w.self_friction * fr,
                                     self_skip=w.self_skip, hash_table_size=w.hash_table_size)
                path = task.path(n, wp)
                p = task.params()
                kv = np.array([[path.velocity(s_) for s_ in range(learner.steps)]])
                oracle.step_many_driven([o], oracle.Params(dt=p.dt, substeps=p.substeps, iterations=p.iterations,
                                                           gravity=p.gravity, coloured=True), [world],
                                        learner.steps, 1, [path.particle], kv)
                agree += int(task.evaluate(o.get()["x"])["routed"] == res.fresh[name][r])
                compared += 1
        out["cpu_cross_check"] = {"attempts": compared, "agree": agree, "kind": "reference" if kind == "ref64"
                                  else "port"}
    log(f"harness learning: {iterations} x {population} rollouts, {res.learn_seconds:.1f} s on the device; "
        f"routed per iteration {[round(i.success_rate, 3) for i in res.iterations]}; fresh cables: learned "
        f"{res.learned_rate:.3f}, hand-written {res.hand_rate:.3f}")
    return out


# ------------------------------------------------------------------ hold/slide sweep (verdicts)

def drape_cpu_verdicts(w: scenes.Workload, sel, steps: int, threads: int) -> dict:
    """Hold/slide verdicts of the reference CPU path (its own sources, FP64,
    colour order) for rods `sel` of the drape sweep, each with its stiffness
    applied to the compliances and its friction to the bar and floor."""
    import oracle  # the CPU checker; never part of the product path

    kind = "ref64" if oracle.available("ref64") else "oracle"
    task = w.info["task"]
    rods, worlds = [], []
    for r in sel:
        rod = w.rods[r]
        o = oracle.Rod.straight(2, 1.0, rod.material.as_tuple(), precision=64, kind=kind)
        a = rod.arrays()
        a["sComp"] = a["sComp"] / float(w.youngs_scale[r])
        a["bComp"] = a["bComp"] / float(w.youngs_scale[r])
        o.set(a)
        rods.append(o)
        prims = [oracle.make_prim(pr.type, pr.a, pr.b, pr.normal, pr.half_extents, pr.rotation, pr.radius,
                                  pr.friction * float(w.friction_scale[r])) for pr in w.world.primitives]
        worlds.append(oracle.World(prims=prims, margin_radii=w.world.contact_margin_radii))
    p = w.params
    params = oracle.Params(dt=p.dt, substeps=p.substeps, iterations=p.iterations, gravity=p.gravity, coloured=True)
    t0 = time.perf_counter()
    oracle.step_many(rods, params, worlds, steps, threads)
    dt = time.perf_counter() - t0
    return {"kind": "reference" if kind == "ref64" else "port", "seconds": dt, "threads": threads,
            "held": {int(r): task.held_from_final(o.get()["x"]) for r, o in zip(sel, rods)}}


def drape_leg(steps: int | None, warmup: int, cpu_rods: int, threads: int) -> dict:
    """The hold/slide sweep: 3600 draped cables stepped as one FP32 batch and
    as one FP64 batch; held/slid from the final height (the cable's highest
    particle within 5 cm of the bar's top). FP32 cannot start a resting cable
    sliding here (gravity's per-substep motion is 0.4 of the FP32 spacing at
    the bar's height), FP64 can; both verdict sets are compared with the
    reference CPU path's on a spot sample: a quarter the cables FP64 holds, a
    quarter those FP32 holds and FP64 lets slide, the rest spread over the
    grid (each group drawn evenly from its list)."""
    w = scenes.drape_sweep(steps=steps)
    steps = w.steps
    task = w.info["task"]
    out = {"config": {"workload": w.name, **{k: v for k, v in w.info.items() if k not in ("task", "meta")}},
           "float_motion_margin": task.float_motion_margin()}
    held = {}
    for prec in (32, 64):
        b = rp.Batch(w.rods, strategy="tiled", world=w.world, precision=prec)
        b.set_material_scales(w.youngs_scale, w.friction_scale)
        b.step(w.params, warmup)
        b.synchronize()
        b.upload_all(w.rods)
        b.set_timing(True)
        secs, c = time_steps(b, w.params, steps)
        work = float(w.segments_total) * w.params.substeps * steps
        launches = max(1, c["last_step_launches"])
        x = b.download_positions()
        ps = b.particle_start
        held[prec] = [task.held_from_final(x[ps[i]:ps[i + 1]]) for i in range(w.num_rods)] \
            if steps == task.steps() else None
        out[f"fp{prec}"] = {"value": work / secs, "unit": "segment-substeps/s", "steps": steps, "seconds": secs,
                            "kernel_ms_avg": c["last_step_gpu_ms"] / launches, "launches": launches,
                            "tile_config": c.get("tile_config_name"),
                            "held": None if held[prec] is None else int(sum(held[prec]))}
        log(f"drape FP{prec} {c.get('tile_config_name')}: {secs:.3f} s for {steps} steps, "
            f"{work / secs / 1e9:.3f} G seg-substeps/s, held {out[f'fp{prec}']['held']} of {w.num_rods}")
        b.destroy()
    if held[32] is None:
        return out
    out["fp32_vs_fp64_agree"] = int(sum(1 for a, c in zip(held[32], held[64]) if a == c))
    if cpu_rods > 0:
        def even(ids, k):
            return [ids[int(j)] for j in np.linspace(0, len(ids) - 1, min(k, len(ids)))] if ids and k > 0 else []

        held64 = [i for i in range(w.num_rods) if held[64][i]]
        only32 = [i for i in range(w.num_rods) if held[32][i] and not held[64][i]]
        sel = even(held64, cpu_rods // 4)
        sel += [i for i in even(only32, cpu_rods // 4) if i not in sel]
        rest = [i for i in range(w.num_rods) if i not in sel]
        sel += even(rest, cpu_rods - len(sel))
        cpu = drape_cpu_verdicts(w, sel, steps, threads)
        out["cpu_spot_check"] = {
            "rods": len(sel), "kind": cpu["kind"], "seconds": cpu["seconds"], "threads": threads,
            "order": "colour", "held_cpu": int(sum(cpu["held"].values())),
            "agree_fp64": int(sum(1 for r, v in cpu["held"].items() if held[64][r] == v)),
            "agree_fp32": int(sum(1 for r, v in cpu["held"].items() if held[32][r] == v))}
        log(f"drape verdicts vs reference CPU FP64 on {len(sel)} cables: FP64 GPU "
            f"{out['cpu_spot_check']['agree_fp64']} agree, FP32 GPU {out['cpu_spot_check']['agree_fp32']} agree")
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed routing steps (default: the whole 6000)")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--plain-steps", type=int, default=None,
                    help="PLAIN timed routing steps of the headline window (default: --steps when at most 400; "
                         "longer windows take PLAIN's rate from the full-rollout leg's windows)")
    ap.add_argument("--rods", type=int, default=1024, help="routing cables")
    ap.add_argument("--tile-config", type=int, default=0)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--uniform-steps", type=int, default=1000, help="steps of the uniform leg (0: skip)")
    ap.add_argument("--cpu-stride", type=int, default=16)
    ap.add_argument("--cpu-steps", type=int, default=6000, help="routing steps of the CPU sample")
    ap.add_argument("--cpu-threads", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--fp64-routing", type=int, default=1, help="also run the routing rollout in FP64 (0: skip)")
    ap.add_argument("--cpu-1thread-stride", type=int, default=128,
                    help="1-thread CPU baseline: every this-many-th routing cable over the whole rollout")
    ap.add_argument("--global-rods", type=int, default=512, help="long-rod (GLOBAL) leg: rods of 2000 segments "
                                                                 "(0: skip)")
    ap.add_argument("--coloured-routing", type=int, default=1,
                    help="also run the routing workload with the coloured self-contact order (0: skip)")
    ap.add_argument("--drape-steps", type=int, default=None, help="steps of the hold/slide sweep (default: the "
                                                                   "whole 8 s = 16000; 0: skip)")
    ap.add_argument("--drape-cpu-rods", type=int, default=32, help="cables of the CPU verdict spot check")
    ap.add_argument("--record-every", type=int, default=16, help="trajectory leg: steps per recorded frame "
                    "(0: skip)")
    ap.add_argument("--cem-iterations", type=int, default=10, help="harness-learning iterations (0: skip)")
    ap.add_argument("--cem-population", type=int, default=1024)
    ap.add_argument("--serve-requests", type=int, default=200,
                    help="requests of the time-sliced Poisson stream (0: skip the serving leg)")
    ap.add_argument("--serve-capacity-requests", type=int, default=40,
                    help="requests of the capacity runs and of the chunked Poisson stream")
    ap.add_argument("--serve-load", type=float, default=0.8,
                    help="Poisson rate as a fraction of the time-sliced server's capacity")
    ap.add_argument("--serve-slice-us", type=int, default=50000, help="time slice of the time-sliced server mode")
    ap.add_argument("--serve-budget-gib", type=float, default=4.0, help="server HBM budget")
    ap.add_argument("--serve-tune-requests", type=int, default=16,
                    help="requests of each serving mode's tile-shape autotuning sample (0: the server's defaults)")
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

    info = rp.device_info()
    log("device", info)
    w = scenes.routing_rollouts(num_rods=args.rods)
    K = args.steps if args.steps is not None else w.steps
    Kp = args.plain_steps if args.plain_steps is not None else (K if K <= 400 else 0)
    W = args.warmup
    log(f"workload {w.name}: {w.num_rods} cables, {w.segments_total} segments, K={K} (PLAIN {Kp}), W={W}")

    leg = routing_leg(w, K, Kp, W, args.tile_config, args.no_plain)
    t = leg["tiled"]
    # The whole rollout in FP32 and FP64 (verdicts, PLAIN windows), whatever K is.
    full32, gv = full_rollout_leg(w, 32, W, args.tile_config, not args.no_plain)
    full64, v64 = full_rollout_leg(w, 64, W, 0, not args.no_plain) if args.fp64_routing else (None, None)
    ft = full32["tiled"]
    kernel_rate = ft["segment_substeps_per_launch"] / (ft["kernel_ms_avg"] * 1e-3)
    roof = roofline(FLOPS_PER_SEGMENT_SUBSTEP_ROUTING, kernel_rate, ft["kernel_ms_avg"],
                    ft["segment_substeps_per_launch"], "routing" if not args.tile_config else None)
    roof["launch"] = f"the full {w.steps}-step routing rollout (full_rollout.fp32.tiled)"
    rcol, vcol = routing_coloured_leg(w, w.steps, W, gv) if args.coloured_routing else (None, None)
    uniform = uniform_leg(args.uniform_steps, W, args.no_plain) if args.uniform_steps > 0 else None
    threads = args.cpu_threads or min(16, os.cpu_count() or 1)
    drape = None
    if args.drape_steps != 0:
        drape = drape_leg(args.drape_steps, W, 0 if args.no_cpu else args.drape_cpu_rods, threads)
    traj = trajectory_leg(w, args.record_every, W) if args.record_every > 0 else None
    glob = global_leg(args.global_rods, W) if args.global_rods > 0 else None
    cem = cem_leg(args.cem_iterations, args.cem_population, not args.no_cpu) if args.cem_iterations > 0 else None
    serving_out = serving_leg(args.serve_requests, args.serve_load, args.serve_budget_gib, args.serve_slice_us,
                              min(args.serve_capacity_requests, args.serve_requests), args.serve_tune_requests) \
        if args.serve_requests > 0 else None

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline_routing(w, args.cpu_stride, min(args.cpu_steps, w.steps), threads)
        cpu["cpu_model"] = cpu_model()
        log(f"CPU baseline: {cpu['value'] / 1e6:.2f} M seg-substeps/s ({cpu['sample']})")
        one = cpu_baseline_routing(w, args.cpu_1thread_stride, w.steps, 1)
        one.pop("_verdicts", None)
        one["cpu_model"] = cpu["cpu_model"]
        cpu["one_thread"] = one
        log(f"CPU baseline, 1 thread: {one['value'] / 1e6:.2f} M seg-substeps/s ({one['sample']})")
        cv = cpu.pop("_verdicts")
        if cv is not None and gv is not None:
            agree = sum(1 for r, v in cv.items() if gv[r] == v)
            ft["verdicts_vs_reference_cpu"] = {"sampled_cables": len(cv), "agree": agree,
                                              "routed_gpu": sum(1 for r in cv if gv[r]),
                                              "routed_cpu": sum(1 for v in cv.values() if v)}
            log(f"routed verdicts, GPU FP32 vs reference CPU FP64 on {len(cv)} cables: {agree} agree")
            if vcol is not None:
                rcol["verdicts_vs_reference_cpu"] = {"sampled_cables": len(cv),
                                                     "agree": sum(1 for r, v in cv.items() if vcol[r] == v)}
                log(f"routed verdicts, coloured self order vs reference CPU FP64 (list order) on {len(cv)} "
                    f"cables: {rcol['verdicts_vs_reference_cpu']['agree']} agree")
        if v64 is not None and args.cpu_steps >= w.steps:
            col = cpu_baseline_routing(w, args.cpu_stride, w.steps, threads, coloured=True)
            cc = col.pop("_verdicts")
            full64["verdicts_vs_reference_cpu_colour_order"] = {
                "sampled_cables": len(cc), "agree_fp64": sum(1 for r, v in cc.items() if v64[r] == v),
                "agree_fp32": sum(1 for r, v in cc.items() if gv is not None and gv[r] == v),
                "routed_cpu": sum(1 for v in cc.values() if v), "cpu_seconds": col["seconds"]}
            log(f"routed verdicts vs reference CPU FP64 in colour order on {len(cc)} cables: FP64 GPU "
                f"{full64['verdicts_vs_reference_cpu_colour_order']['agree_fp64']} agree, FP32 GPU "
                f"{full64['verdicts_vs_reference_cpu_colour_order']['agree_fp32']} agree")

    plain = leg.get("plain")
    if plain is None and "plain_windows" in full32 and K == w.steps:
        plain = full32["plain_windows"]
    out = {
        "metric": METRIC, "value": t["value"], "unit": "segment-substeps/s", "n_gpus": 1, "steps": K, "warmup": W,
        "ms_per_step": t["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (seeded)",
        "config": {"workload": w.name, **{k: v for k, v in w.info.items() if k not in ("paths", "task")},
                   "timed_steps": f"steps 0-{K - 1} of the {w.steps}-step rollout (after {W} untimed warm-up "
                                  f"steps and a reset to step 0)",
                   "strategy": "tiled (ragged-packed persistent launches, hipGraph-captured)",
                   "parallelism": "single GPU"},
        "tiled": t,
        "plain": plain,
        "speedup_vs_plain": (t["value"] / plain["value"]) if plain else None,
        "roofline": roof,
        "cpu_baseline": cpu,
        "full_rollout": {"fp32": full32, "fp64": full64},
        "routing_coloured_self_order": rcol,
        "uniform": uniform,
        "drape": drape,
        "harness_learning": cem,
        "trajectory": traj,
        "global_long_rods": glob,
        "serving": serving_out,
        "p50_latency_s": serving_out["time_sliced"]["poisson"]["p50_latency_s"] if serving_out else None,
        "p99_latency_s": serving_out["time_sliced"]["poisson"]["p99_latency_s"] if serving_out else None,
        "device": {"arch": info["arch"], "compute_units": info["compute_units"]},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
