"""The serving workload and its driver.

Requests (seeded): 1-256 cantilevers of 32-256 segments (0.5 m, reference
material, clamped at the root, horizontal 0.3 m above a floor, 5 cm apart),
a horizon of 0.5-8 s at dt = 1e-3 and 10 substeps, a floor with friction
U[0.3, 0.9] and, each with probability 1/2, a sphere, a capsule and a box
under the rods; self-collision on in half of the requests. Per-rod stiffness
0.4-4x (log-uniform) and friction 0.6-1.4x.

``autotune_shapes`` picks a server's tile shapes by serving a sample of
requests once per candidate pair.

``run_stream`` plays requests against a ``Server`` at given arrival times:
arrived requests are admitted in arrival order while the budget has room,
every running request advances in each launch, and finished results are read
back as soon as their launch completes. Admission runs while a launch is in
flight (submission only queues its uploads behind it), so the host's work
overlaps the device's. Latency is arrival to result on the host, wall
clock.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field, replace

import numpy as np

from .rod import CollisionWorld, Primitive, SolverParams, clamp_root_exact, make_straight_rod
from .scenes import REFERENCE_MATERIAL
from .server import Server, ServerFull

PARAMS = SolverParams(dt=1e-3, substeps=10, iterations=1, gravity=(0.0, 0.0, -9.81))
ROD_LENGTH = 0.5
HEIGHT = 0.3
SPACING = 0.05

_PROTOS: dict = {}


def prototype(n: int):
    r = _PROTOS.get(n)
    if r is None:
        r = make_straight_rod(n, ROD_LENGTH, REFERENCE_MATERIAL, (0.0, 0.0, HEIGHT), (1.0, 0.0, 0.0))
        clamp_root_exact(r)
        _PROTOS[n] = r
    return r


@dataclass
class Request:
    index: int
    rods: list
    x: np.ndarray            # packed initial positions
    world: CollisionWorld
    steps: int
    youngs: np.ndarray
    friction: np.ndarray
    segments: int = 0
    info: dict = field(default_factory=dict)

    @property
    def work(self) -> float:
        return float(self.segments) * PARAMS.substeps * self.steps


def make_request(index: int, rng: np.random.Generator) -> Request:
    nr = int(rng.integers(1, 257))
    ns = rng.integers(32, 257, nr)
    rods = [prototype(int(n)) for n in ns]
    x = np.concatenate([r.x + np.array([0.0, SPACING * i, 0.0]) for i, r in enumerate(rods)])
    width = SPACING * nr
    prims = [Primitive.plane((0.0, 0.0, 0.0), (0.0, 0.0, 1.0), float(rng.uniform(0.3, 0.9)))]
    if rng.uniform() < 0.5:
        prims.append(Primitive.sphere((float(rng.uniform(0.2, 0.45)), float(rng.uniform(0.0, width)), 0.05),
                                      float(rng.uniform(0.03, 0.08)), float(rng.uniform(0.2, 0.8))))
    if rng.uniform() < 0.5:
        xc = float(rng.uniform(0.15, 0.45))
        prims.append(Primitive.capsule((xc, -0.05, 0.08), (xc, width + 0.05, 0.08), float(rng.uniform(0.01, 0.03)),
                                       float(rng.uniform(0.2, 0.8))))
    if rng.uniform() < 0.5:
        prims.append(Primitive.box((float(rng.uniform(0.25, 0.45)), width / 2, 0.04), (0.05, width / 2 + 0.05, 0.04),
                                   (1.0, 0.0, 0.0, 0.0), float(rng.uniform(0.2, 0.8))))
    self_coll = bool(rng.uniform() < 0.5)
    world = CollisionWorld(primitives=prims, self_collision=self_coll, self_friction=0.3, hash_table_size=256)
    steps = int(round(rng.uniform(0.5, 8.0) / PARAMS.dt))
    youngs = 0.4 * 10.0 ** rng.uniform(0.0, 1.0, nr)
    friction = rng.uniform(0.6, 1.4, nr)
    return Request(index, rods, x, world, steps, youngs, friction, int(ns.sum()),
                   {"rods": nr, "primitives": len(prims), "self_collision": self_coll})


def make_requests(num: int, seed: int = 20260917) -> list:
    rng = np.random.default_rng(seed)
    return [make_request(i, rng) for i in range(num)]


def poisson_arrivals(num: int, rate: float, seed: int = 1) -> np.ndarray:
    """Arrival offsets in seconds of a Poisson process of `rate` requests/s."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, num))


def run_stream(server: Server, requests: list, arrivals: np.ndarray, keep_results: bool = False) -> dict:
    """Serve `requests`, request k arriving `arrivals[k]` s after the start.
    Returns per-request latencies (s), the wall time from the first arrival to
    the last result, and (optionally) the results."""
    n = len(requests)
    order = list(np.argsort(arrivals, kind="stable"))
    queue: list = []
    ids: dict = {}
    latency = np.full(n, np.nan)
    results = {} if keep_results else None
    t0 = time.perf_counter()
    nxt = 0
    done = 0
    launches = 0
    max_resident = 0
    def admit():
        nonlocal nxt
        now = time.perf_counter() - t0
        while nxt < n and arrivals[order[nxt]] <= now:
            queue.append(order[nxt])
            nxt += 1
        while queue:
            k = queue[0]
            q = requests[k]
            try:
                rid = server.submit(q.rods, q.world, q.steps, q.youngs, q.friction, x=q.x)
            except ServerFull:
                break
            ids[rid] = k
            queue.pop(0)

    while done < n:
        admit()
        st = server.stats()
        max_resident = max(max_resident, st["resident_requests"])
        if st["running_requests"] == 0:
            if nxt < n:
                time.sleep(max(0.0, arrivals[order[nxt]] - (time.perf_counter() - t0)))
            continue
        server.step()
        launches += 1
        # Requests arriving during the launch are admitted while it runs;
        # they join the next one.
        admit()
        server.synchronize()
        for rid in server.finished():
            k = ids.pop(rid)
            if keep_results:
                results[k] = server.result(rid)
            else:
                server.result(rid)
            latency[k] = time.perf_counter() - t0 - arrivals[k]
            server.release(rid)
            done += 1
    wall = time.perf_counter() - t0 - float(np.min(arrivals))
    return {"latency": latency, "wall": wall, "launches": launches, "max_resident": max_resident,
            "results": results}


# Tile-shape pairs (rods without self-collision, rods with it) that
# autotune_shapes tries, by configuration id: tp3-tc1-w4 / tp1-tc1-w8 (the
# chunked default), tp2-tc1-w4 / tp1-tc1-w8, tp3-tc1-w2 / tp1-tc1-w8,
# tp3-tc1-w4 / tp2-tc1-w4 (the time-sliced default), tp2-tc1-w4 for both and
# tp3-tc1-w4 for both.
SHAPE_PAIRS = ((10, 13), (9, 13), (6, 13), (10, 9), (9, 9), (10, 10))


def autotune_shapes(requests: list, hbm_budget_bytes: int, time_slice_us: int = 0, steps: int = 64,
                    pairs=SHAPE_PAIRS, params: SolverParams = PARAMS, max_launches: int = 10000):
    """Choose a server's tile shapes on a sample of its requests.

    Each (lean, self-colliding) pair in `pairs` serves `requests`, all
    arriving at once and their horizons cut to `steps` steps, in a fresh
    server of the given budget and mode, after a warm-up on the first two; the
    pair with the shortest wall time wins. A pair under which some request is
    refused by an empty server (a rod that shape cannot take) is skipped.
    Returns (best pair, [(pair, seconds or None), ...]). The shapes change the
    time only: every request's result is the same bit for bit.

    There is no counterpart in the reference, which has no server; the batch
    path's autotuner (run inside rp_batch_step) plays the same role for one
    batch, but a server's resident set changes at every admission, and trials
    on it would stall the running requests, so the server's shapes are chosen
    before it serves."""
    sample = [replace(q, steps=min(q.steps, steps)) for q in requests]
    trials = []
    for lean, selfc in pairs:
        server = Server(params, hbm_budget_bytes, tile_config=lean, self_tile_config=selfc,
                        time_slice_us=time_slice_us)
        try:
            t = _serve_at_once(server, sample[:2], max_launches)  # code objects, first launches
            if t is not None:
                t = _serve_at_once(server, sample, max_launches)
        finally:
            server.destroy()
        trials.append(((lean, selfc), t))
    timed = [(p, t) for p, t in trials if t is not None]
    if not timed:
        raise RuntimeError("no tile-shape pair serves these requests")
    return min(timed, key=lambda e: e[1])[0], trials


def _serve_at_once(server: Server, requests: list, max_launches: int):
    """Wall time to serve `requests` arriving at once (admitted in order as
    the budget frees), or None when an empty server refuses one of them."""
    t0 = time.perf_counter()
    pending = list(requests)
    running = 0
    launches = 0
    while pending or running:
        while pending:
            q = pending[0]
            try:
                server.submit(q.rods, q.world, q.steps, q.youngs, q.friction, x=q.x)
            except ServerFull:
                if running == 0:
                    return None
                break
            pending.pop(0)
            running += 1
        server.step()
        server.synchronize()
        launches += 1
        for rid in server.finished():
            server.result(rid)
            server.release(rid)
            running -= 1
        if launches > max_launches:
            raise RuntimeError("the sample did not finish within max_launches server steps")
    return time.perf_counter() - t0
