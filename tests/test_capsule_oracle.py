"""CPU pins of the oracle's capsule narrow phase (capsule vs sphere,
capsule, box; 30_narrowphase.metal:510-745): rest heights against the
closed form (within the collision-margin band the contact formulation
settles in, as the reference's box-rest gate 0.5 +- 0.03), the two
reference capsule tests restated (GaudiFunicularTests.swift:110-150,
GPUSolverTests.swift:577-620), and contact normals against a brute-force
segment distance."""
import numpy as np

import oracle
import capsule_fixtures as F


def test_lying_capsule_rests_on_its_side():
    s, c = F.lying_capsule()
    o = oracle.OracleSolver(s)
    o.step(180)
    p, r, v, _ = o.bodyStates()
    assert abs(p[c, 2] - 0.1) < 0.015
    np.testing.assert_allclose(np.abs(r[c]), np.abs(np.array(F.ALONG_X, np.float32)), atol=1e-3)
    assert np.linalg.norm(v[c]) < 0.01
    h, _, _ = o.manifolds()
    assert [list(x) for x in h] == [[0, 1, 3]]      # three seeds, three ground contacts


def test_sphere_rests_on_capsule_bar():
    s, cap, sph = F.sphere_on_capsule()
    o = oracle.OracleSolver(s)
    o.step(180)
    p, _, v, _ = o.bodyStates()
    assert abs(p[sph, 2] - (0.5 + 0.1 + 0.15)) < 0.015
    assert abs(p[sph, 0]) < 1e-4 and abs(p[sph, 1]) < 1e-4
    assert np.linalg.norm(v[sph]) < 0.01


def test_capsule_rests_across_two_bars():
    s, bars, up = F.crossed_capsules()
    o = oracle.OracleSolver(s)
    o.step(180)
    p, _, v, _ = o.bodyStates()
    assert abs(p[up, 2] - (0.4 + 0.05 + 0.07)) < 0.015
    assert np.linalg.norm(v[up]) < 0.01
    h, nrm, _ = o.manifolds()
    active = {(int(a), int(b)): k for k, (a, b, n) in enumerate(h) if n > 0}
    for bar in bars:
        k = active[(bar, up)]
        assert h[k][2] == 1
        # normal from B (the resting capsule) toward A (the bar): -z
        np.testing.assert_allclose(nrm[k], [0, 0, -1], atol=1e-4)


def test_rope_rope_and_bag_rope_contacts():
    # GaudiFunicularTests.testRopeRopeAndBagRopeNarrowphase: contacts on the
    # first frame, separated within 3 mm after 240 frames, at rest
    s, a, b, bag = F.ropes()
    o = oracle.OracleSolver(s)
    o.step()
    h, _, _ = o.manifolds()
    active = {(int(x), int(y)) for x, y, n in h if n > 0}
    assert (a, b) in active and (a, bag) in active
    o.step(240)
    p, r, v, _ = o.bodyStates()
    A0, A1 = F.capsule_segment(p[a], r[a], 0.8)
    B0, B1 = F.capsule_segment(p[b], r[b], 0.8)
    depth = -(F.segment_distance(A0, A1, B0, B1) - 0.08 + s.settings.collisionMargin)
    assert depth < 0.003
    assert np.all(np.isfinite(p)) and np.all(np.linalg.norm(v[[a, b, bag]], axis=1) < 0.1)


def test_compound_capsule_foot_holds_under_sub_coulomb_load():
    # GPUSolverTests.testCompoundCapsuleFootHoldsUnderSubCoulombActuatorLoad
    s, foot = F.capsule_foot()
    o = oracle.OracleSolver(s)
    o.step(600)
    settled = o.positions()[foot].copy()
    for _ in range(1200):
        q = o.positions()[foot]
        o.setJointWorldAnchors([0], [[q[0] + 0.04, q[1], q[2]]])
        o.step()
    p, _, v, _ = o.bodyStates()
    assert abs(p[foot, 0] - settled[0]) < 0.001
    assert np.linalg.norm(v[foot]) < 0.03


def test_capsule_capsule_normal_matches_brute_force_closest_points():
    # skew capsules just touching: the first-frame normal is the unit vector
    # between the brute-force closest points of the two axes
    from avbd_amd.scene import PhysicsScene, SimSettings
    rng = np.random.default_rng(7)
    for trial in range(6):
        s = PhysicsScene("skew", SimSettings(gravity=0.0, iterations=1))
        q1 = F.quat(rng.uniform(0, np.pi), rng.standard_normal(3))
        q2 = F.quat(rng.uniform(0, np.pi), rng.standard_normal(3))
        c1 = np.zeros(3)
        A0, A1 = F.capsule_segment(c1, q1, 0.6)
        # place capsule 2 so that its axis passes 0.09 m from capsule 1's axis
        B0, B1 = F.capsule_segment(np.zeros(3), q2, 0.6)
        u = np.cross(A1 - A0, B1 - B0)
        u /= np.linalg.norm(u)
        c2 = u * 0.09 + 0.05 * (A1 - A0)
        a = s.addCapsule(0.6, 0.05, 0.0, 0.5, tuple(c1), rotation=q1)
        b = s.addCapsule(0.6, 0.05, 1.0, 0.5, tuple(c2), rotation=q2)
        o = oracle.OracleSolver(s, precision=64)
        o.step()
        h, nrm, _ = o.manifolds()
        assert [list(x) for x in h] == [[a, b, 1]], trial
        B0, B1 = F.capsule_segment(c2, q2, 0.6)
        # brute-force closest points
        ss = np.linspace(0, 1, 2001)
        P = A0[None] + (A1 - A0)[None] * ss[:, None]
        Q = B0[None] + (B1 - B0)[None] * ss[:, None]
        d = np.linalg.norm(P[:, None] - Q[None], axis=2)
        i, j = np.unravel_index(np.argmin(d), d.shape)
        want = (P[i] - Q[j]) / np.linalg.norm(P[i] - Q[j])      # from B toward A
        assert np.dot(nrm[0], want) > 0.999, trial


def test_convex_hull_rests_on_box_terrain():
    # GPUSolverTests.testConvexHullColliderRestsOnBoxTerrain (:537-570)
    s, ground, body = F.hull_cube()
    o = oracle.OracleSolver(s)
    o.step(360)
    p, r, _, _ = o.bodyStates()
    assert abs(p[body, 2] - 0.5) < 0.06
    assert abs(r[body, 0]) < 0.05 and abs(r[body, 1]) < 0.05
    h, nrm, _ = o.manifolds()
    act = [(int(a), int(b), int(n)) for a, b, n in h if n > 0]
    assert (ground, body, 4) in act                  # the four bottom vertices


def test_hull_pairs_other_than_box_are_refused():
    s, body, sph = F.hull_wedge()
    o = oracle.OracleSolver(s)
    o.step()
    h, _, _ = o.manifolds()
    pairs = {(int(a), int(b)): int(n) for a, b, n in h}
    assert pairs.get((body, sph), pairs.get((sph, body))) == 0     # overlapping, but no contacts
    o.step(240)
    p = o.positions()
    assert np.all(np.isfinite(p))
    h, _, _ = o.manifolds()
    ground_contacts = [int(n) for a, b, n in h if {int(a), int(b)} == {0, body}]
    assert ground_contacts and 2 <= ground_contacts[0] <= 8


def _rt_contacts(o):
    sc = o.softContacts()
    keyA = sc["lambda"][:, 3].view(np.uint32)
    return sc[(keyA >> 28) == 2]


def test_capsule_rigid_features_against_a_triangle():
    """25_cloth.metal:957-970: a capsule meets cloth triangles through three
    spheres of its radius, at its ends (features 9, 11) and middle (10);
    a sphere is used when its center is within the triangle's reach."""
    for x, feat in ((0.5, 11), (-0.5, 9), (0.0, 10)):
        gap, rp, rc = 0.003, 0.002, 0.1
        s, cap, ids = F.triangle_over_capsule(x=x, gap=gap, rp=rp, rc=rc)
        o = oracle.OracleSolver(s)
        o.step()
        rt = _rt_contacts(o)
        assert len(rt) == 1, (x, len(rt))
        c = rt[0]
        assert c["lambda"][3:4].view(np.uint32)[0] == (2 << 28) | cap
        assert c["penalty"][3:4].view(np.uint32)[0] == 0 * 16 + feat
        assert list(c["ids"]) == [cap] + ids
        np.testing.assert_allclose(abs(float(c["normal"][2])), 1.0, atol=1e-6)
        # gap between the feature sphere and the triangle's particle shell
        margin = o.clothInfo()["elem_margin"]
        np.testing.assert_allclose(float(c["C0"][0]), gap - rp + margin, atol=2e-6)
