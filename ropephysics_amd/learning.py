"""Learning to route a wire harness with the cross-entropy method, each
iteration's rollouts stepped as one ragged GPU batch.

Reference: runHarnessLearning (src/validation/application_cases.cpp:353-486).
A population of gripper motions is sampled from a Gaussian over the waypoint
coordinates, starting at the hand-written motion
(`RoutingTask.hand_written_waypoints`: straight at the targets, leaving the
cable slack). Each motion meets its own random cable: stiffness log-uniform
over 0.4-4x, friction 0.6-1.4x (randomCable, application_cases.cpp:314-322). The distribution is refit each
iteration to the best 10% by score, with a floor on its spread. The learned
motion is then compared with the hand-written one on fresh cables, and a few
of those attempts are checked on the CPU.

Differences from the reference:
- The task is this package's routing task (`apps.RoutingTask`: eight
  waypoints, a box clip).
- Cables are ragged: each population slot holds a cable of 48-160 segments,
  drawn once.
- Gripper schedules are evaluated on the device (`Batch.set_paths`), so an
  iteration is one batched call with no per-tick host traffic.
- Random draws use numpy's PCG64 generator with the stated seed.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from .apps import RoutingTask
from .batch import Batch


@dataclass
class Iteration:
    success_rate: float
    mean_score: float
    seconds: float  # device rollout time (wall clock around the batched steps)


@dataclass
class LearningResult:
    mean: np.ndarray  # learned waypoints, (W, 2)
    iterations: list = field(default_factory=list)
    learned_rate: float = 0.0  # routed share of fresh cables under the learned motion
    hand_rate: float = 0.0  # ... under the hand-written motion
    learn_seconds: float = 0.0
    fresh: dict = field(default_factory=dict)  # the evaluation cables and their verdicts


class HarnessLearner:
    """Cross-entropy search over a routing motion's waypoint coordinates.

    One batch holds the whole population. It is reset and reused every
    iteration, with new waypoints, stiffness and friction per slot."""

    def __init__(self, task: RoutingTask | None = None, population: int = 1024, elite_fraction: float = 0.1,
                 sigma0: float = 0.08, sigma_floor: float = 0.003, segments=(48, 160), seed: int = 20260917,
                 steps: int | None = None, precision: int = 32):
        self.task = task or RoutingTask()
        self.population = int(population)
        self.elites = max(1, int(population * elite_fraction))
        self.sigma0 = float(sigma0)
        self.sigma_floor = float(sigma_floor)
        self.rng = np.random.default_rng(seed)
        self.steps = self.task.steps() if steps is None else int(steps)
        self.segments = self.rng.integers(segments[0], segments[1] + 1, self.population)
        protos: dict = {}
        self.rods = [protos.setdefault(int(n), self.task.build(int(n))) for n in self.segments]
        self.batch = Batch(self.rods, strategy="tiled", world=self.task.world(), precision=precision)

    def random_cables(self):
        """Stiffness 0.4 * 10^U(0, 1) and friction 0.6 + 0.8 U(0, 1) per slot."""
        u = self.rng.uniform(0.0, 1.0, (self.population, 2))
        return 0.4 * 10.0 ** u[:, 0], 0.6 + 0.8 * u[:, 1]

    def rollouts(self, waypoints: np.ndarray, youngs: np.ndarray, friction: np.ndarray):
        """Every slot's cable driven along its own waypoints (P, W, 2) for the
        task's horizon; returns the outcomes (apps.RoutingTask.evaluate) and the
        device time."""
        task, b = self.task, self.batch
        b.upload_all(self.rods)
        b.set_clock()
        b.set_material_scales(youngs, friction)
        b.set_paths(0, [task.path(int(n), w) for n, w in zip(self.segments, waypoints)])
        b.synchronize()
        t0 = time.perf_counter()
        b.step(task.params(), self.steps)
        b.synchronize()
        seconds = time.perf_counter() - t0
        x = b.download_positions()
        ps = b.particle_start
        outcomes = [task.evaluate(np.asarray(x[ps[r]:ps[r + 1]], np.float64)) for r in range(self.population)]
        return outcomes, seconds

    def run(self, iterations: int = 10) -> LearningResult:
        task = self.task
        hand = task.hand_written_waypoints()
        W = len(hand)
        mean = hand.reshape(-1).astype(np.float64)
        sigma = np.full(mean.size, self.sigma0)
        result = LearningResult(mean=hand.copy())
        for _ in range(iterations):
            draws = mean[None, :] + sigma[None, :] * self.rng.normal(0.0, 1.0, (self.population, mean.size))
            samples = np.stack([task.clamp_reach(d) for d in draws])  # (P, W, 2)
            youngs, friction = self.random_cables()
            outcomes, seconds = self.rollouts(samples, youngs, friction)
            scores = np.array([o["score"] for o in outcomes])
            routed = np.array([o["routed"] for o in outcomes])
            result.iterations.append(Iteration(float(routed.mean()), float(scores.mean()), seconds))
            result.learn_seconds += seconds
            order = np.argsort(-scores, kind="stable")[:self.elites]
            elite = samples[order].reshape(self.elites, -1)
            mean = elite.mean(axis=0)
            sigma = np.maximum(np.sqrt(((elite - mean) ** 2).mean(axis=0)), self.sigma_floor)
        learned = task.clamp_reach(mean.reshape(W, 2))
        result.mean = learned
        youngs, friction = self.random_cables()
        got_l, _ = self.rollouts(np.repeat(learned[None], self.population, axis=0), youngs, friction)
        got_h, _ = self.rollouts(np.repeat(hand[None], self.population, axis=0), youngs, friction)
        result.learned_rate = float(np.mean([o["routed"] for o in got_l]))
        result.hand_rate = float(np.mean([o["routed"] for o in got_h]))
        result.fresh = {"youngs": youngs, "friction": friction, "segments": self.segments.copy(),
                        "learned": [o["routed"] for o in got_l], "hand": [o["routed"] for o in got_h]}
        return result
