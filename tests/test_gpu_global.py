"""GPU tests of the GLOBAL strategy: rods whose working arrays exceed a
workgroup's LDS, stepped with those arrays in HBM (the counterpart of the
reference's multi-kernel strategy, src/gpu/gpu_solver.cu:417-529,691-730,
which GpuSim falls back to when the fused kernel does not fit,
src/gpu/gpu_suite.cpp:125-130).

The bar is the same as for the other strategies: bit for bit equal to the
oracle of the batch's precision in colour order, and bit for bit equal to
PLAIN (the same per-rod arithmetic) on rods that fit both.
"""
import os
import subprocess
import textwrap

import numpy as np
import pytest

import ropephysics_amd as rp
from ropephysics_amd import _abi
from helpers import (REFERENCE_MATERIAL, benchmark_params, benchmark_rod, bits, compare_bitwise, oracle_params,
                     oracle_state, oracle_twin, oracle_world)
from test_gpu_parity import SELF_PARAMS, SELF_WORLD, helix_rope, pinned_coil

pytestmark = pytest.mark.gpu

FLOOR = rp.CollisionWorld(primitives=[rp.Primitive.plane((0, 0, -0.0051), (0, 0, 1), 0.5)])


def long_rod(n: int = 2000) -> rp.Rod:
    """The benchmark rod at n segments, 4 m long (2 mm segments), root
    clamped."""
    rod = rp.make_straight_rod(n, 4.0, REFERENCE_MATERIAL, (0, 0, 0), (1, 0, 0))
    rp.clamp_root_exact(rod)
    return rod


def _check_against_oracle(rods, world, params, checkpoints, strategy, precision=32, ys=None, fs=None):
    b = rp.Batch(rods, strategy=strategy, world=world, precision=precision)
    if ys is not None or fs is not None:
        b.set_material_scales(ys, fs)
    twins = []
    for i, r in enumerate(rods):
        if precision == 32:
            cs = 1.0 if ys is None else float(np.float32(1.0) / np.float32(ys[i]))
            f = 1.0 if fs is None else float(np.float32(fs[i]))
        else:
            cs = 1.0 if ys is None else 1.0 / float(ys[i])
            f = 1.0 if fs is None else float(fs[i])
        twins.append((oracle_twin(r, precision), oracle_params(params, compliance_scale=cs, friction_scale=f)))
    ow = oracle_world(world)
    done = 0
    for cp in checkpoints:
        b.step(params, cp - done)
        for o, op in twins:
            o.step(op, ow, cp - done)
        done = cp
        b.synchronize()
        for i, (o, _) in enumerate(twins):
            got, want = b.download(i), oracle_state(o)
            for k in ("x", "v", "q", "omega"):
                g = np.ascontiguousarray(got[k])
                w = np.ascontiguousarray(want[k], g.dtype)
                assert np.array_equal(g.view(np.uint8), w.view(np.uint8)), (i, cp, k)
    return b


def test_long_rods_do_not_fit_lds_and_auto_picks_global():
    rod = long_rod()
    with pytest.raises(rp.RopePhysicsError) as e:
        rp.Batch([rod], strategy="plain")
    assert e.value.code == _abi.RP_ERR_CAPACITY and "GLOBAL" in str(e.value)
    with pytest.raises(rp.RopePhysicsError):
        rp.Batch([rod], strategy="tiled")
    b = rp.Batch([rod], strategy="auto")
    c = b.counters()
    assert c["strategy"] == "global" and c["scratch_bytes"] == 0
    b.step(benchmark_params(), 1)
    c = b.counters()
    assert c["scratch_bytes"] > rp.device_info()["lds_per_workgroup"] and c["lanes_total"] == 512


@pytest.mark.parametrize("precision", [32, 64])
def test_long_rods_bitwise_vs_oracle(precision):
    """A 2000-segment rod free of contacts and one resting its sag on a
    floor, 60 steps (checked after 1, 10 and 60). In FP32 its positions and
    orientations (64 KB) are held in LDS for the launch; in FP64 (128 KB) they
    stay in HBM with the other arrays."""
    for world in (None, FLOOR):
        b = _check_against_oracle([long_rod()], world, benchmark_params(), [1, 10, 60], "global", precision)
        lds = b.counters()["lds_bytes_per_workgroup"]
        assert lds == (2001 + 2001) * 16 if precision == 32 else lds == 0


def test_long_rods_with_scales_contacts_and_ragged_sizes():
    """A ragged GLOBAL batch (2000, 1200 and 64 segments) on a floor and a
    sphere with per-rod stiffness and friction scales."""
    world = rp.CollisionWorld(primitives=[rp.Primitive.plane((0, 0, -0.0051), (0, 0, 1), 0.5),
                                          rp.Primitive.sphere((1.5, 0, -0.2), 0.15, 0.3)])
    rods = [long_rod(2000), long_rod(1200), benchmark_rod(64)]
    _check_against_oracle(rods, world, benchmark_params(iterations=2), [5, 40], "global",
                          ys=np.array([0.5, 1.0, 3.0]), fs=np.array([1.2, 0.6, 1.0]))


def test_long_self_colliding_coil():
    """A 640-segment coil pressed onto itself (both ends pinned) with
    self-collision: beyond LDS once its pool is large, stepped in HBM."""
    rod = pinned_coil(turns=40, per_turn=16)
    world = rp.CollisionWorld(primitives=SELF_WORLD.primitives, self_collision=True, self_friction=0.3,
                              hash_table_size=256, self_capacity=6000)
    b = _check_against_oracle([rod], world, SELF_PARAMS, [1, 4, 12], "global")
    assert b.self_contact_overflow() == 0


def test_global_equals_plain_on_rods_that_fit_both():
    """Self-collision, all primitive kinds, gripper paths and observers: the
    GLOBAL batch equals the PLAIN one bit for bit; splitting the steps into
    calls of 1, 7 and 100 does not change the bits."""
    from ropephysics_amd.apps import RoutingTask

    task = RoutingTask()
    rods = [task.build(48), task.build(97), helix_rope()]
    rng = np.random.default_rng(5)
    paths = [task.path(48, task.base_waypoints() + rng.normal(0, 0.01, (8, 2))),
             task.path(97, task.base_waypoints() + rng.normal(0, 0.01, (8, 2)))]
    out = []
    for strategy, chunks in (("plain", [108]), ("global", [108]), ("global", [1, 7, 100])):
        b = rp.Batch(rods, strategy=strategy, world=task.world())
        b.set_material_scales([0.7, 2.0, 1.0], [1.1, 0.8, 1.0])
        b.set_paths(0, paths)
        b.set_observers(0, [(0, 10, 0.0), (5, 10, 0.0), (3, 10, 0.0)])
        for n in chunks:
            b.step(task.params(), n)
        b.synchronize()
        out.append((b.get_state(), b.get_observations(), list(b.get_clock())))
    (s0, o0, c0), rest = out[0], out[1:]
    for s, o, c in rest:
        for k in s0:
            assert np.array_equal(bits(s0[k]), bits(s[k])), k
        assert np.array_equal(o0, o) and c0 == c == [108, 108, 108]


def test_global_records_trajectories():
    """Frames stored by the GLOBAL kernel every 5 steps: frame k equals the
    positions of a second batch stepped 5 (k + 1) steps, bit for bit."""
    rods = [long_rod(900), benchmark_rod(32)]
    b = rp.Batch(rods, strategy="global")
    b.record(5, 4)
    b.step(benchmark_params(), 20)
    b.synchronize()
    assert b.recorded_frames() == 4
    f = b.frames()
    for k in (0, 3):
        ref = rp.Batch(rods, strategy="global")
        ref.step(benchmark_params(), 5 * (k + 1))
        assert np.array_equal(bits(f[k]), bits(ref.download_positions()))


FACADE_PROGRAM = textwrap.dedent("""
    // A program written the way the reference's call sites use Batch
    // (GpuSim::create, src/gpu/gpu_suite.cpp:125-130): the fused strategy
    // first, the multi-kernel one as the fallback.
    #include <cstdio>
    #include <fstream>
    #include <vector>
    #include <ropephysics/batch.hpp>

    struct V3 { double x = 0, y = 0, z = 0; };
    struct Q4 { double w = 1, x = 0, y = 0, z = 0; };
    struct State {
        std::vector<V3> x, v, extForce, kinematicVelocity, omega, invInertia, inertia, extTorque;
        std::vector<Q4> q;
        std::vector<double> invMass;
    };
    struct Stretch { std::vector<int> p0, p1, seg; std::vector<double> restLength; std::vector<V3> compliance; };
    struct Bend { std::vector<int> segA, segB; std::vector<double> restLength; std::vector<V3> compliance, restDarboux; };
    struct Material { double radius = 0; };
    struct RodLike { State state; Stretch stretch; Bend bend; Material material; };
    struct Prim { int type = 0; V3 a, b, normal, halfExtents; Q4 rotation; double radius = 0, friction = 0; };
    struct World {
        std::vector<Prim> primitives;
        double contactMarginRadii = 0.25;
        bool selfCollision = false;
        double selfFriction = 0;
        int selfSkip = 2;
        int hashTableSize = 4096;
    };

    template <typename T> static void rd(std::ifstream& f, T* p, size_t n) { f.read((char*)p, sizeof(T) * n); }
    static void rd3(std::ifstream& f, std::vector<V3>& v, size_t n) { v.resize(n); for (auto& e : v) rd(f, &e.x, 3); }

    int main(int argc, char** argv) {
        std::ifstream f(argv[1], std::ios::binary);
        int n[4];
        rd(f, n, 4);
        RodLike r;
        State& s = r.state;
        rd3(f, s.x, n[0]); rd3(f, s.v, n[0]);
        s.invMass.resize(n[0]); rd(f, s.invMass.data(), n[0]);
        rd3(f, s.extForce, n[0]); rd3(f, s.kinematicVelocity, n[0]);
        s.q.resize(n[1]); for (auto& q : s.q) rd(f, &q.w, 4);
        rd3(f, s.omega, n[1]); rd3(f, s.invInertia, n[1]); rd3(f, s.inertia, n[1]); rd3(f, s.extTorque, n[1]);
        r.stretch.p0.resize(n[2]); r.stretch.p1.resize(n[2]); r.stretch.seg.resize(n[2]);
        rd(f, r.stretch.p0.data(), n[2]); rd(f, r.stretch.p1.data(), n[2]); rd(f, r.stretch.seg.data(), n[2]);
        r.stretch.restLength.resize(n[2]); rd(f, r.stretch.restLength.data(), n[2]);
        rd3(f, r.stretch.compliance, n[2]);
        r.bend.segA.resize(n[3]); r.bend.segB.resize(n[3]);
        rd(f, r.bend.segA.data(), n[3]); rd(f, r.bend.segB.data(), n[3]);
        r.bend.restLength.resize(n[3]); rd(f, r.bend.restLength.data(), n[3]);
        rd3(f, r.bend.compliance, n[3]); rd3(f, r.bend.restDarboux, n[3]);
        rd(f, &r.material.radius, 1);
        World world;
        Prim floor;
        floor.type = 0;
        floor.a.z = -0.0051;
        floor.normal.z = 1;
        floor.friction = 0.5;
        world.primitives.push_back(floor);

        rp::Batch batch;
        const bool fused = batch.create(r, 2, rp::Strategy::kFused, &world);
        std::printf("fused create: %d\\n", int(fused));
        if (!batch.create(r, 2, rp::Strategy::kMultiKernel, &world)) {
            std::printf("multi-kernel create failed: %s\\n", rp_last_error());
            return 2;
        }
        rp_counters c;
        rp_batch_counters(batch.handle(), &c);
        std::printf("strategy %d\\n", c.strategy);
        rp::BatchParams p;
        p.dt = 1e-3f;
        p.substeps = 8;
        p.iterations = 1;
        for (int i = 0; i < 25; ++i) batch.step(p);
        batch.synchronize();
        RodLike out = r;
        batch.download(1, out);
        std::ofstream o(argv[2], std::ios::binary);
        for (auto& e : out.state.x) { float v[3] = {float(e.x), float(e.y), float(e.z)}; o.write((char*)v, 12); }
        for (auto& e : out.state.q) { float v[4] = {float(e.w), float(e.x), float(e.y), float(e.z)}; o.write((char*)v, 16); }
        std::printf("ok\\n");
        return 0;
    }
""")


def test_facade_multi_kernel_fallback_steps_a_long_rod(tmp_path):
    """The C++ facade's create(prototype, 2, kMultiKernel, &world) succeeds
    for a 2000-segment rod and gives the FP32 oracle's bits after 25 steps
    (the facade rounds the FP64 rod to FP32, as the reference's Batch does)."""
    rod = long_rod()
    src = tmp_path / "facade_global.cpp"
    src.write_text(FACADE_PROGRAM)
    exe = tmp_path / "facade_global"
    libdir = os.path.dirname(_abi.LIB_PATH)
    repo = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    res = subprocess.run(["g++", "-std=c++17", "-O1", "-I" + os.path.join(repo, "include"), str(src), "-o", str(exe),
                          "-L" + libdir, "-lropephysics", "-Wl,-rpath," + libdir], capture_output=True, text=True)
    assert res.returncode == 0, res.stderr[-3000:]
    inp = tmp_path / "rod.bin"
    with open(inp, "wb") as f:
        f.write(np.array([rod.num_particles, rod.num_frames, len(rod.sP0), len(rod.bA)], np.int32).tobytes())
        for a in (rod.x, rod.v, rod.invMass, rod.extForce, rod.kinVel, rod.q, rod.omega, rod.invInertia,
                  rod.inertia, rod.extTorque):
            f.write(np.ascontiguousarray(a, np.float64).tobytes())
        for a in (rod.sP0, rod.sP1, rod.sSeg):
            f.write(np.ascontiguousarray(a, np.int32).tobytes())
        for a in (rod.sRest, rod.sComp):
            f.write(np.ascontiguousarray(a, np.float64).tobytes())
        for a in (rod.bA, rod.bB):
            f.write(np.ascontiguousarray(a, np.int32).tobytes())
        for a in (rod.bRest, rod.bComp, rod.bRestD):
            f.write(np.ascontiguousarray(a, np.float64).tobytes())
        f.write(np.array([rod.material.radius], np.float64).tobytes())
    outp = tmp_path / "out.bin"
    run = subprocess.run([str(exe), str(inp), str(outp)], capture_output=True, text=True, timeout=300)
    assert run.returncode == 0, (run.stdout, run.stderr)
    assert "strategy 3" in run.stdout  # RP_STRATEGY_GLOBAL
    raw = np.fromfile(outp, np.float32)
    P, F = rod.num_particles, rod.num_frames
    x, q = raw[:3 * P].reshape(P, 3), raw[3 * P:].reshape(F, 4)
    o = oracle_twin(rod, 32)
    o.step(oracle_params(benchmark_params()), oracle_world(FLOOR), 25)
    want = o.get()
    assert np.array_equal(bits(x), bits(want["x"])) and np.array_equal(bits(q), bits(want["q"]))
