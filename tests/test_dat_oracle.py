"""CPU pins of the oracle's Planar-DAT truncation (oracle/oracle.cpp,
datReducePair / datApply / the pass order): the reference's analytic
division-plane tests (PlanarDATTests.swift:272-1003, 474-650, 1312-1432)
restated on the oracle, and a no-pass-through property of two sheets fired
at each other far faster than the gap allows."""
import numpy as np
import pytest

import oracle
from dat_fixtures import (RELAX, bilateral, edge_edge, fraction, place, sheets_fired, small_triangles, strictest,
                          two_triangles)


def _run(scene, hooks=None, setup=None):
    o = oracle.OracleSolver(scene)
    if hooks:
        o.setDatTestHooks(**hooks)
    if setup:
        setup(o)
    o.step()
    return o


def _rq(o):
    return o.clothInfo()["dat_query_radius"]


def test_vt_predictor_matches_analytic_division_plane():
    # :272-303 -- the static target takes the minimum 5 % of the gap
    s, _, src = two_triangles(0.2, (0, 0, -0.4))
    o = _run(s)
    hit = (0.2 - 0.2 * 0.05) / 0.4
    t = fraction(hit)
    assert abs(0.4 * t) < 0.5 * RELAX * _rq(o)          # the plane, not the rq cap, binds
    z = o.positions()[src, 2]
    np.testing.assert_allclose(z, 0.2 - 0.4 * t, atol=2e-5)
    assert o.clothStats()["dat_vt_pairs"] >= 3
    assert o.datStats()["dat_truncations"] >= 3


def test_bilateral_plane_allocates_relative_motion():
    # :768-798 -- relative approach 0.4, the target gets 25 %: both hit at 0.5
    s, tgt, src = bilateral()
    o = _run(s)
    t = 0.5 * RELAX
    p = o.positions()
    np.testing.assert_allclose(p[src, 2], 0.2 - 0.3 * t, atol=2e-5)
    np.testing.assert_allclose(p[tgt, 2], 0.1 * t, atol=2e-5)


def test_exact_vt_plane_hit_is_strictly_truncated():
    # :800-818 -- landing exactly on z = 0.01 at t = 1 is still truncated
    s, _, src = two_triangles(0.2, (0, 0, -0.19))
    z = _run(s).positions()[src, 2]
    np.testing.assert_allclose(z, 0.2 - 0.19 * RELAX, atol=2e-5)
    assert np.all(z > 0.01)


def test_strictest_simultaneous_plane_wins():
    # :820-850
    s, src = strictest()
    hit = (0.2 - 0.105) / 0.4
    z = _run(s).positions()[src, 2]
    np.testing.assert_allclose(z, 0.2 - 0.4 * fraction(hit), atol=2e-5)
    assert np.all(z > 0.105)


def test_tangential_motion_keeps_its_progress():
    # :852-882 (planar half) -- motion parallel to the plane is not truncated
    s, _, src = two_triangles(0.05, (0.1, 0, 0))
    starts = np.array([s.bodies[i]["position"] for i in src], np.float32)
    p = _run(s).positions()[src]
    assert np.all(p[:, 0] - starts[:, 0] > 0.095)
    np.testing.assert_allclose(p[:, 2], 0.05, atol=1e-5)


def test_vt_truncation_is_translation_invariant():
    # :884-914
    off = (10000.0, -10000.0, 10000.0)
    s, _, src = two_triangles(0.2, (0, 0, -0.4), offset=off)
    hit = 0.95 * 0.2 / 0.4
    z = _run(s).positions()[src, 2] - off[2]
    np.testing.assert_allclose(z, 0.2 - 0.4 * fraction(hit), atol=0.002)


@pytest.mark.parametrize("vz,expected", [(-0.4, 0.2 - 0.4 * fraction(0.95 * 0.2 / 0.4)), (-0.19, 0.2 - 0.19 * RELAX)])
def test_edge_edge_predictor_matches_analytic_plane(vz, expected):
    # :916-975 -- E-E only (no V-T pair within rq = 0.5); exact hit truncated
    s, src = edge_edge(velocity_z=vz)
    o = _run(s, hooks=dict(queryRadius=0.5))
    cs = o.clothStats()
    assert cs["dat_vt_pairs"] == 0 and cs["dat_pairs"] > 0
    z = o.positions()[src[:2], 2]
    np.testing.assert_allclose(z, expected, atol=2e-5)
    assert np.all(z > 0.01)


def test_intersecting_edge_anchor_uses_conservative_fallback():
    # :977-1001 -- crossing edges at the anchor: fallback normal, frozen, no failure
    s, src = edge_edge(height=0.0, velocity_z=0.0)
    o = _run(s, hooks=dict(queryRadius=0.3))
    np.testing.assert_allclose(o.positions()[src, 2], 0.0, atol=1e-6)
    d = o.datStats()
    assert d["dat_ee_degeneracies"] > 0 and d["dat_nonfinite"] == 0 and d["dat_invalid_anchors"] == 0


def test_nonfinite_prediction_freezes_at_finite_anchor():
    # :652-680 -- infinite velocity: invalid anchor, the surface stays at the start
    s, _, src = two_triangles(0.2, (0, 0, 0))
    o = _run(s, setup=lambda o: place(o, s, src, velocity=(0, 0, np.inf)))
    d = o.datStats()
    assert d["dat_invalid_anchors"] > 0 and d["dat_nonfinite"] > 0
    z = o.positions()[src, 2]
    assert np.all(np.isfinite(z))
    np.testing.assert_allclose(z, 0.2, atol=1e-6)


def test_element_span_overflow_freezes():
    # :1312-1337 -- a target vertex 100 m away overflows the 4-cell cover
    s, tgt, src = two_triangles(0.2, (0, 0, 0))
    o = _run(s, setup=lambda o: place(o, s, [tgt[0]], x=-100.0))
    assert o.datStats()["overflow"] > 0
    p = o.positions()
    np.testing.assert_allclose(p[tgt[0], 0], -100.0, atol=1e-6)
    np.testing.assert_allclose(p[src, 2], 0.2, atol=1e-6)


def test_pair_overflow_freezes_particles():
    # :1363-1393 (capacity 1 here: the backend's override caps at > 0)
    s, _, src = two_triangles(0.2, (0, 0, -0.4), iterations=2)
    o = _run(s, hooks=dict(pairCapacity=1))
    assert o.clothStats()["dat_pairs"] > 1
    np.testing.assert_allclose(o.positions()[src, 2], 0.2, atol=1e-6)


def test_soft_overflow_restores_step_start_pose():
    # :1395-1432 (capacity 1)
    s, _, src = two_triangles(0.04, (0.1, 0, 0), iterations=2)
    starts = np.array([s.bodies[i]["position"] for i in src], np.float32)
    o = _run(s, hooks=dict(softCapacity=1))
    assert o.clothStats()["soft_candidates"] > 1
    np.testing.assert_allclose(o.positions()[src], starts, atol=1e-6)


def test_accepted_pose_query_drops_pair_that_moved_outside_radius():
    # :474-515 -- moving away: only the rq cap applies; the start query saw
    # the pair (peak demand), the accepted-pose query dropped it
    s, _, src = two_triangles(0.5, (0, 0, 0))
    o = oracle.OracleSolver(s)
    rq = _rq(o)
    place(o, s, src, z=0.8 * rq, velocity=(0, 0, 2 * rq))
    o.step()
    assert o.clothStats()["dat_pairs"] > 0
    assert len(o.datPairs()) == 0
    z = o.positions()[src, 2]
    assert np.all(z > rq)
    np.testing.assert_allclose(z, 0.8 * rq + 0.5 * RELAX * rq, atol=2e-5)


def test_accepted_pose_query_finds_entering_contact():
    # :517-563 -- the pair enters rq during the prediction: contacts at the
    # accepted pose, the displacement capped at 0.5 * 0.85 * rq
    s, _, src = small_triangles()
    o = oracle.OracleSolver(s)
    rq = _rq(o)
    start = 1.01 * rq
    place(o, s, src, z=start, velocity=(0, 0, -2 * rq))
    o.step()
    assert len(o.softContacts()) >= 3
    z = o.positions()[src, 2]
    assert np.all(z < rq)
    np.testing.assert_allclose(z, start - 0.5 * RELAX * rq, atol=2e-5)


def test_accepted_pose_contact_c0_is_linearized_at_step_start():
    # :565-650 -- C0 back-linearized to the step-start pose, not divided by 1 - alpha
    s, tgt, src = small_triangles()
    o = oracle.OracleSolver(s)
    rq = _rq(o)
    start = 1.01 * rq
    place(o, s, src, z=start, velocity=(0, 0, -2 * rq))
    starts = o.positions().astype(np.float64)
    o.step()
    acc = o.positions().astype(np.float64)
    sc = o.softContacts()
    sel = [k for k in range(len(sc)) if list(sc["ids"][k]) == [src[2]] + list(tgt)]
    assert sel, "the accepted-pose query must emit the V-T contact of the third source vertex"
    k = sel[0]
    n = sc["normal"][k][:3].astype(np.float64)
    w = sc["weights"][k].astype(np.float64)
    ids = sc["ids"][k]
    start_rel = sum(w[i] * starts[ids[i]] for i in range(4))
    acc_rel = sum(w[i] * acc[ids[i]] for i in range(4))
    radius = lambda b: s.bodies[b]["size"][0] / 2
    h = radius(ids[0]) + max(radius(ids[1]), radius(ids[2]), radius(ids[3]))
    margin = o.clothInfo()["elem_margin"]
    expected = float(n @ start_rel) - h + margin
    disp = float(n @ (acc_rel - start_rel))
    assert abs(disp) > 1e-4
    c0 = float(sc["C0"][k][0])
    assert abs(c0 - expected) <= 2e-5
    assert abs(expected) > 1e-3
    assert abs(c0 - expected / (1 - 0.99)) > 10 * abs(expected)


def test_fired_sheets_do_not_pass_through():
    # the upper sheet moves 5 cm per step toward the lower one 9 mm below
    s, (lower, upper) = sheets_fired()
    o = oracle.OracleSolver(s)
    for _ in range(6):
        o.step()
        p = o.positions()
        # every upper vertex over the lower sheet stays above it
        assert p[upper, 2].min() > p[lower, 2].max() - 1e-4
    assert o.datStats()["dat_invalid_anchors"] == 0
