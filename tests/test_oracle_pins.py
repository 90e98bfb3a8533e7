"""Pins on the CPU oracle (the checker), independent of the GPU.

Each pin is chosen so a plausible mistake in the restatement (a dropped term,
a wrong sign or index, a transposed operand) fails it:
  * the reference's own golden trajectory (docs/data/crosscheck.csv);
  * bitwise equality with the reference's CPU sources compiled as-is
    (oracle/_ref, built in the container that holds the checkout);
  * textbook routines and closed forms for the pieces (Cholesky vs LAPACK,
    analytic distances, brute-force closest points, material constants);
  * invariants (momentum in free flight, the unloaded rest state).
"""
import csv
import math
import os

import numpy as np
import pytest

import oracle
import ropephysics_amd as rp
from helpers import REFERENCE_MATERIAL, benchmark_rod, oracle_params, oracle_twin, oracle_world

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
MAT = REFERENCE_MATERIAL.as_tuple()


def test_crosscheck_golden_trajectory():
    """16 segments, clamped root, dt 1e-3, 4 substeps, 200 steps, sequential
    sweeps (reference src/validation/cases.cpp:760-805): every particle and
    frame (ghost included) within 1e-9 of the reference's published dump."""
    rows = list(csv.DictReader(open(os.path.join(GOLDEN, "crosscheck.csv"))))
    r = oracle.Rod.straight(16, 1.0, MAT)
    r.clamp_end(False)
    r.step(oracle.Params(dt=1e-3, substeps=4, iterations=1, coloured=False), None, 200)
    a = r.get()
    ex = np.array([[float(w[k]) for k in "xyz"] for w in rows])
    eq = np.array([[float(w[k]) for k in ("qw", "qx", "qy", "qz")] for w in rows])
    assert len(rows) == 17 and a["x"].shape == (17, 3) and a["q"].shape == (17, 4)
    assert np.max(np.abs(a["x"] - ex)) < 1e-9
    assert np.max(np.abs(a["q"] - eq)) < 1e-9
    # The tip has visibly moved: the trajectory is not trivially at rest.
    assert ex[-1, 2] < -0.1


def _scenarios():
    """(name, build, world, params, steps): gravity, loaded, contact and
    self-collision scenes in the shapes the reference's GPU cases use."""
    out = []
    out.append(("gravity", lambda kind, prec: _clamped(kind, prec, 24), None,
                oracle.Params(dt=1e-3, substeps=8, iterations=1), 60))

    def loaded(kind, prec):
        r = _clamped(kind, prec, 20)
        a = r.get()
        a["extForce"][-1] = (0, 0.02, 0.05)
        a["extTorque"][10] = (2e-3, 0, -1e-3)
        a["kinVel"][0] = (0, 0.05, 0.1)
        r.set(a)
        return r

    out.append(("loaded", loaded, None, oracle.Params(dt=1e-3, substeps=8, iterations=2, lin_damp=0.5,
                                                      ang_damp=0.3), 40))
    world = oracle.World(prims=[
        oracle.make_prim(oracle.PLANE, a=(0, 0, -0.0051), normal=(0, 0, 1), friction=0.5),
        oracle.make_prim(oracle.SPHERE, a=(0.5, 0, -0.0851), radius=0.08, friction=0.3),
        oracle.make_prim(oracle.CAPSULE, a=(0.3, -0.1, -0.02), b=(0.3, 0.1, -0.02), radius=0.01, friction=0.4),
        oracle.make_prim(oracle.BOX, a=(0.8, 0, -0.03), half=(0.05, 0.05, 0.02), rot=(0.9659258, 0, 0, 0.258819),
                         friction=0.6)])
    out.append(("contact", lambda kind, prec: _clamped(kind, prec, 32), world,
                oracle.Params(dt=1e-3, substeps=8, iterations=1), 80))

    def rope(kind, prec):
        return oracle.Rod.centerline(helix_points(), (2e6, 0.35, 1000.0, 5e-3), (0, 0, 1), 0.0, precision=prec,
                                     kind=kind)

    selfw = oracle.World(prims=[oracle.make_prim(oracle.PLANE, a=(0, 0, 0), normal=(0, 0, 1), friction=0.6)],
                         self_collision=True, self_friction=0.3, hash_table_size=256)
    out.append(("self", rope, selfw, oracle.Params(dt=1e-3, substeps=16, iterations=4, lin_damp=1, ang_damp=1),
                40))
    return out


def _clamped(kind, prec, n):
    r = oracle.Rod.straight(n, 1.0, MAT, precision=prec, kind=kind)
    r.clamp_end(False)
    return r


def helix_points(turns=3.0, per_turn=20, radius=0.03, pitch=0.0095, z0=0.02):
    """A coiled rope whose neighbouring turns overlap (pitch < 2r), so it is
    in self-contact from the first substep, as the reference's coiled pile is
    after it falls (src/gpu/gpu_cases.cpp:421-511)."""
    n = int(turns * per_turn)
    t = np.arange(n + 1) / per_turn * 2 * np.pi
    return np.stack([radius * np.cos(t), radius * np.sin(t), z0 + pitch * t / (2 * np.pi)], axis=-1)


@pytest.mark.skipif(not oracle.available("ref64"), reason="reference CPU build (oracle/_ref) not present")
@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("coloured", [False, True])
@pytest.mark.parametrize("scenario", _scenarios(), ids=lambda s: s[0])
def test_oracle_bitwise_equals_reference_build(scenario, coloured, prec):
    name, build, world, params, steps = scenario
    params.coloured = coloured
    a = build("oracle", prec)
    b = build("ref64" if prec == 64 else "ref32", prec)
    a.step(params, world, steps)
    b.step(params, world, steps)
    sa, sb = a.get(), b.get()
    for k in ("x", "v", "q", "omega", "sLam", "bLam"):
        assert np.array_equal(sa[k].view(np.uint64), sb[k].view(np.uint64)), (name, k)
    assert a.contact_counts() == b.contact_counts()
    if name == "self":
        # Self-collision must change the outcome (reference gpu_cases.cpp:537-541).
        c = build("oracle", prec)
        w2 = oracle.World(prims=world.prims, self_collision=False)
        c.step(params, w2, steps)
        assert np.max(np.abs(c.get()["x"] - sa["x"])) > 1e-3
    if name == "contact":
        assert a.contact_counts()[0] >= 5


@pytest.mark.skipif(not oracle.available("ref64"), reason="reference CPU build (oracle/_ref) not present")
def test_builders_bitwise_equal_reference_build():
    pts = [(math.cos(t) * 0.3, math.sin(t) * 0.3, 0.01 * t) for t in np.linspace(0, 6, 41)]
    for kind in ("ref64",):
        for make in (lambda k: oracle.Rod.straight(33, 0.7, MAT, (0.1, -0.2, 0.3), (0.2, 0.5, 1.0), kind=k),
                     lambda k: oracle.Rod.centerline(pts, MAT, (0, 1, 0), 0.8, kind=k)):
            a, b = make("oracle"), make(kind)
            a.clamp_end(False)
            b.clamp_end(False)
            a.clamp_end(True)
            b.clamp_end(True)
            sa, sb = a.get(), b.get()
            for k in sa:
                assert np.array_equal(np.asarray(sa[k]), np.asarray(sb[k])), k


def test_cholesky_matches_lapack():
    rng = np.random.default_rng(1)
    for _ in range(200):
        M = rng.normal(size=(3, 3))
        A = M @ M.T + 0.1 * np.eye(3)
        b = rng.normal(size=3)
        ok, y = oracle.chol_solve3(A, b)
        assert ok
        np.testing.assert_allclose(y, np.linalg.solve(A, b), rtol=1e-10, atol=1e-12)
    # Not positive definite: the solve refuses, as the reference's does.
    ok, _ = oracle.chol_solve3(np.diag([1.0, -1.0, 1.0]), [1, 1, 1])
    assert not ok


def test_signed_distances_analytic():
    plane = oracle.make_prim(oracle.PLANE, a=(0, 0, 1), normal=(0, 0, 1))
    d, n = oracle.signed_distance(plane, (3, -2, 1.25))
    assert d == pytest.approx(0.25) and np.allclose(n, (0, 0, 1))
    sph = oracle.make_prim(oracle.SPHERE, a=(1, 1, 1), radius=0.5)
    d, n = oracle.signed_distance(sph, (1, 1, 3))
    assert d == pytest.approx(1.5) and np.allclose(n, (0, 0, 1))
    cap = oracle.make_prim(oracle.CAPSULE, a=(0, 0, 0), b=(2, 0, 0), radius=0.1)
    d, n = oracle.signed_distance(cap, (1, 0.5, 0))        # beside the axis
    assert d == pytest.approx(0.4) and np.allclose(n, (0, 1, 0))
    d, n = oracle.signed_distance(cap, (3, 0, 0))          # beyond end B
    assert d == pytest.approx(0.9) and np.allclose(n, (1, 0, 0))
    c, s = math.cos(0.3), math.sin(0.3)
    q = (math.cos(0.15), 0, 0, math.sin(0.15))             # 0.3 rad about z
    box = oracle.make_prim(oracle.BOX, a=(1, 2, 3), half=(0.5, 0.2, 0.1), rot=q)
    # A point 0.3 beyond the local +x face.
    p = np.array([1, 2, 3]) + np.array([c, s, 0]) * 0.8
    d, n = oracle.signed_distance(box, p)
    assert d == pytest.approx(0.3) and np.allclose(n, (c, s, 0))
    # Inside, nearest the local +z face: depth 0.05.
    d, n = oracle.signed_distance(box, (1, 2, 3.05))
    assert d == pytest.approx(-0.05) and np.allclose(n, (0, 0, 1))
    # The FP32 build agrees to FP32 rounding.
    d32, _ = oracle.signed_distance(box, p, precision=32)
    assert d32 == pytest.approx(0.3, abs=1e-6)


def test_segment_closest_points_brute_force():
    rng = np.random.default_rng(2)
    u = np.linspace(0, 1, 401)
    for _ in range(60):
        p1, q1, p2, q2 = rng.normal(size=(4, 3))
        s, t = oracle.seg_seg(p1, q1, p2, q2)
        assert 0 <= s <= 1 and 0 <= t <= 1
        d = np.linalg.norm((p1 + s * (q1 - p1)) - (p2 + t * (q2 - p2)))
        A = p1[None, None, :] + u[:, None, None] * (q1 - p1)[None, None, :]
        B = p2[None, None, :] + u[None, :, None] * (q2 - p2)[None, None, :]
        brute = np.min(np.linalg.norm(A - B, axis=-1))
        assert d <= brute + 1e-12
        assert d >= brute - 2e-2 * max(np.linalg.norm(q1 - p1), np.linalg.norm(q2 - p2))
    # Parallel segments and a degenerate one.
    s, t = oracle.seg_seg((0, 0, 0), (1, 0, 0), (0.5, 1, 0), (1.5, 1, 0))
    assert abs(np.linalg.norm(np.array([s, 0, 0]) - np.array([0.5 + t, 1, 0])) - 1.0) < 1e-12
    s, t = oracle.seg_seg((0, 0, 0), (0, 0, 0), (-1, 1, 0), (1, 1, 0))
    assert s == 0 and t == pytest.approx(0.5)


def test_hash_is_integer_exact_and_floor_correct():
    # Teschner hash with unsigned wrap-around; the buckets of negative cells.
    for ix, iy, iz, T in [(0, 0, 0, 256), (-1, 2, -3, 256), (1000, -1000, 77, 4096), (-5, -5, -5, 97)]:
        h = ((ix * 73856093) & 0xFFFFFFFF) ^ ((iy * 19349663) & 0xFFFFFFFF) ^ ((iz * 83492791) & 0xFFFFFFFF)
        assert oracle.hash_cell(ix, iy, iz, T) == h % T
    assert oracle.cell_coord(-0.01, 0.1) == -1 and oracle.cell_coord(0.0, 0.1) == 0
    assert oracle.cell_coord(0.35, 0.1) == 3 and oracle.cell_coord(-0.35, 0.1) == -4


def test_colouring_two_colours_conflict_free():
    rows = list(csv.DictReader(open(os.path.join(GOLDEN, "coloring.csv"))))
    for row in rows:
        n = int(row["segments"])
        r = oracle.Rod.straight(n, 1.0, MAT)
        r.clamp_end(False)
        a = r.get()
        for fam in (0, 1):
            order, start = r.colouring(fam)
            assert len(start) - 1 == int(row["stretch_colors" if fam == 0 else "bend_colors"]) == 2
            for c in range(len(start) - 1):
                used = set()
                for k in order[start[c]:start[c + 1]]:
                    res = ([a["sP0"][k], a["sP1"][k], 10**6 + a["sSeg"][k]] if fam == 0
                           else [10**6 + a["bA"][k], 10**6 + a["bB"][k]])
                    assert not (used & set(res))
                    used |= set(res)
            assert sorted(order.tolist()) == list(range(len(order)))


def test_material_closed_forms():
    """Compliance of the stretch and bend blocks from the closed forms
    G = E / 2(1+nu), A = pi r^2, I = pi r^4 / 4, J = 2I,
    kappa = 6(1+nu)/(7+6nu), alpha_s = diag(1/kGA, 1/kGA, 1/EA)/l,
    alpha_b = diag(1/EI, 1/EI, 1/GJ)/lbar."""
    E, nu, rho, rr = 3e6, 0.4, 1200.0, 4e-3
    r = oracle.Rod.straight(10, 0.5, (E, nu, rho, rr))
    a = r.get()
    l = 0.05
    G = E / (2 * (1 + nu))
    A = np.pi * rr**2
    I = np.pi * rr**4 / 4
    k = 6 * (1 + nu) / (7 + 6 * nu)
    np.testing.assert_allclose(a["sComp"][3], [1 / (k * G * A * l), 1 / (k * G * A * l), 1 / (E * A * l)], rtol=1e-12)
    np.testing.assert_allclose(a["bComp"][3], [1 / (E * I * l), 1 / (E * I * l), 1 / (G * 2 * I * l)], rtol=1e-12)
    np.testing.assert_allclose(a["mass"].sum(), rho * A * 0.5, rtol=1e-12)
    m = rho * A * l
    np.testing.assert_allclose(a["inertia"][3], [m * (3 * rr**2 + l**2) / 12, m * (3 * rr**2 + l**2) / 12,
                                                 0.5 * m * rr**2], rtol=1e-12)


def test_momentum_conserved_in_free_flight():
    r = oracle.Rod.straight(20, 1.0, MAT, (0, 0, 0), (1, 0.2, 0.1))
    a = r.get()
    rng = np.random.default_rng(3)
    a["v"] = rng.normal(size=a["v"].shape) * 0.1
    r.set(a)
    p0 = (a["mass"][:, None] * a["v"]).sum(0)
    r.step(oracle.Params(dt=1e-3, substeps=8, gravity=(0, 0, 0), coloured=True), None, 100)
    b = r.get()
    p1 = (b["mass"][:, None] * b["v"]).sum(0)
    assert np.max(np.abs(p1 - p0)) < 1e-12


def test_rest_state_is_a_fixed_point():
    r = oracle.Rod.straight(16, 1.0, MAT)
    r.clamp_end(False)
    x0 = r.get()["x"].copy()
    r.step(oracle.Params(gravity=(0, 0, 0), coloured=True), None, 50)
    assert np.max(np.abs(r.get()["x"] - x0)) < 1e-13


def test_flop_counts_frozen():
    """The per-segment-substep FP32 operation counts bench.py uses: the
    contact-free step, and the routing cables with their world and
    self-collision from the initial pose."""
    import bench
    from ropephysics_amd import scenes

    rod = rp.make_straight_rod(64, 0.5, REFERENCE_MATERIAL, (0, 0, 0.1), (1, 0, 0))
    rp.clamp_root_exact(rod)
    o = oracle_twin(rod, 64)
    c = o.count_ops(oracle_params(rp.SolverParams(dt=1e-3, substeps=10), precision=64), None, 10)
    flops = (c["add"] + c["mul"] + c["div"] + c["sqrt"]) / (64 * 10 * 10)
    assert flops == pytest.approx(bench.FLOPS_PER_SEGMENT_SUBSTEP, abs=0.01)

    w = scenes.routing_rollouts()
    ow = oracle_world(w.world)
    total = segs = 0
    for r in range(0, w.num_rods, 64):
        c = oracle_twin(w.rods[r], 64).count_ops(oracle_params(w.params, precision=64), ow, 10)
        total += c["add"] + c["mul"] + c["div"] + c["sqrt"]
        segs += w.rods[r].num_segments
    assert total / (segs * 10 * 10) == pytest.approx(bench.FLOPS_PER_SEGMENT_SUBSTEP_ROUTING, abs=0.01)