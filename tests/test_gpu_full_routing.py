"""The whole headline workload against the oracle: all 1024 routing cables,
the full 6 s (6000 steps of 10 substeps), one TILED launch on the GPU, and
every cable stepped by the oracle of the same precision on host threads with
the same gripper schedule (src/core/solver.cpp's step, colour order). Every
cable's final x, v, q and omega must be equal word for word.

The oracle side is about 6.4e9 segment-substeps of host work (two to three
minutes on 16 threads), so the test runs only when RP_FULL_ROUTING=1 is set;
the default GPU suite checks ten FP32 and eleven FP64 cables of the same
workload over the same 6 s (tests/test_gpu_trajectory.py,
tests/test_gpu_fp64.py)."""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle
import ropephysics_amd as rp
from helpers import compare_bitwise, compliance_scale_f32, oracle_params, oracle_state, oracle_twin, oracle_world

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(os.environ.get("RP_FULL_ROUTING") != "1",
                                 reason="host-side oracle of the whole workload takes minutes; set RP_FULL_ROUTING=1")]


def _words(a, precision):
    return np.ascontiguousarray(a, np.float32 if precision == 32 else np.float64).view(
        np.uint32 if precision == 32 else np.uint64)


@pytest.mark.parametrize("precision", [32, 64])
def test_every_routing_cable_bitwise_over_the_whole_rollout(precision):
    from ropephysics_amd import scenes

    w = scenes.routing_rollouts()
    paths = w.info["paths"]
    b = rp.Batch(w.rods, strategy="tiled", world=w.world, precision=precision)
    b.set_material_scales(w.youngs_scale, w.friction_scale)
    b.set_paths(0, paths)
    b.step(w.params, w.steps)
    b.synchronize()
    gpu = [b.download(i) for i in range(w.num_rods)]
    ow = oracle_world(w.world)

    def differing(i):
        o = oracle_twin(w.rods[i], precision)
        p = paths[i]
        kv = np.array([[p.velocity(s) for s in range(w.steps)]])
        if precision == 32:
            op = oracle_params(w.params, compliance_scale=compliance_scale_f32(w.youngs_scale[i]),
                               friction_scale=float(np.float32(w.friction_scale[i])))
        else:
            op = oracle_params(w.params, compliance_scale=1.0 / float(w.youngs_scale[i]),
                               friction_scale=float(w.friction_scale[i]))
        oracle.step_many_driven([o], op, [ow], w.steps, 1, [p.particle], kv)
        st = oracle_state(o)
        return sum(int(np.count_nonzero(_words(gpu[i][k], precision) != _words(st[k], precision)))
                   for k in ("x", "v", "q", "omega"))

    with ThreadPoolExecutor(16) as ex:
        diffs = list(ex.map(differing, range(w.num_rods)))
    bad = [i for i, d in enumerate(diffs) if d]
    assert not bad, bad[:20]
