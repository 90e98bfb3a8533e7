"""The coloured self-contact order of the oracle (World.self_order = 1), on the
CPU: an opt-in order of this build. The reference projects self contacts in
list order only (src/core/solver.cpp:163-216), so nothing in the reference
fixes this order. It is pinned by the properties that define it, checked on the
oracle's exported colours and ranks rather than restated:
  * a proper colouring: no two self contacts of one colour (below 32) share a
    particle;
  * greedy in list order: a contact's colour c is the lowest one free, so for
    every b < c an earlier contact sharing one of its particles holds b;
  * the projection order is by colour, then by list position, the tail last,
    after the world contacts;
  * when no two self contacts share a particle, the coloured order is the list
    order, and the run is bitwise the list-order run;
  * the reference build refuses the coloured order.
"""
import numpy as np
import pytest

import oracle
from test_oracle_pins import MAT, helix_points

SELF_PARAMS = dict(dt=1e-3, substeps=16, iterations=4, gravity=(0, 0, -9.81), lin_damp=1.0, ang_damp=1.0)


def coil(prec=64, kind="oracle"):
    return oracle.Rod.centerline(helix_points(), MAT, precision=prec, kind=kind)


def world(order, prims=True):
    p = [oracle.make_prim(0, (0, 0, 0), (0, 0, 0), (0, 0, 1), (0, 0, 0), (1, 0, 0, 0), 0.0, 0.6)] if prims else []
    return oracle.World(prims=p, self_collision=True, self_friction=0.3, hash_table_size=256, self_order=order)


@pytest.mark.parametrize("prec", [64, 32])
def test_colouring_is_proper_greedy_and_orders_by_colour(prec):
    # One substep: the list is the one generated from the initial pose, whose
    # neighbouring turns overlap (the contacts push them apart within a few
    # substeps).
    r = coil(prec)
    r.step(oracle.Params(**dict(SELF_PARAMS, substeps=1)), world(1), 1)
    sc = r.self_contacts()
    idx, col, rank = sc["idx"], sc["colour"], sc["rank"]
    n = len(col)
    assert n > 50  # the coil is in dense self-contact
    assert np.all((col >= 0) & (col <= 32))
    assert sorted(rank.tolist()) == list(range(n))
    parts = [set(int(p) for p in idx[i] if p >= 0) for i in range(n)]
    assert max(col) >= 2  # chains of contacts sharing particles need several colours
    for i in range(n):
        earlier = {int(col[j]) for j in range(i) if parts[i] & parts[j]}
        if col[i] < 32:
            # proper: no earlier contact of the same colour shares a particle ...
            assert col[i] not in earlier, i
            # ... and greedy: every lower colour is taken by such a contact
            assert all(b in earlier for b in range(col[i])), i
        else:
            assert earlier >= set(range(32)), i
    # projection order: by colour, then list position
    by_rank = np.argsort(rank)
    keys = [(int(col[i]), int(i)) for i in by_rank]
    assert keys == sorted(keys)


def test_list_order_runs_give_no_colours():
    r = coil()
    r.step(oracle.Params(**dict(SELF_PARAMS, substeps=1)), world(0), 1)
    sc = r.self_contacts()
    assert len(sc["colour"]) > 0 and np.all(sc["colour"] == -1) and np.all(sc["rank"] == -1)


def crossing(prec):
    """A rod laid as a polyline with particles every 5 cm: a leg along x at
    z = 24 mm, then two legs along y at x = +-0.15 m and z = 16 mm that pass
    under it mid-segment. Radius 5 mm: two self contacts (8 mm apart at each
    crossing), which share no particle."""
    way = np.array([(-0.325, 0, 0.024), (0.325, 0, 0.024), (0.325, 0.3, 0.02), (0.15, 0.3, 0.016),
                    (0.15, -0.3, 0.016), (-0.15, -0.3, 0.016), (-0.15, 0.3, 0.016)])
    seg = np.linalg.norm(np.diff(way, axis=0), axis=1)
    cum = np.concatenate([[0.0], np.cumsum(seg)])
    s = np.arange(0.0, cum[-1], 0.05)
    pts = np.stack([np.interp(s, cum, way[:, k]) for k in range(3)], axis=1)
    return oracle.Rod.centerline(pts, MAT, precision=prec)


@pytest.mark.parametrize("prec", [64, 32])
def test_disjoint_contacts_keep_list_order_bitwise(prec):
    p = oracle.Params(dt=1e-3, substeps=1, iterations=2, gravity=(0, 0, 0))
    a, b = crossing(prec), crossing(prec)
    a.step(p, world(0, prims=False), 1)
    b.step(p, world(1, prims=False), 1)
    sc = b.self_contacts()
    parts = [set(int(q) for q in row) for row in sc["idx"]]
    n = len(parts)
    assert n == 2
    assert all(not (parts[i] & parts[j]) for i in range(n) for j in range(i))
    assert np.all(sc["colour"] == 0) and list(sc["rank"]) == list(range(n))
    sa, sb = a.get(), b.get()
    for k in ("x", "v", "q", "omega"):
        assert np.array_equal(sa[k].view(np.uint64), sb[k].view(np.uint64)), k


@pytest.mark.skipif(not oracle.available("ref64"), reason="reference CPU build (oracle/_ref) not present")
def test_reference_build_refuses_the_coloured_order():
    r = coil(64, "ref64")
    with pytest.raises(RuntimeError):
        r.step(oracle.Params(**SELF_PARAMS), world(1), 1)
