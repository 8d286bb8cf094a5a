"""GPU parity of the capsule narrow phase (30_narrowphase.metal:510-745)
through the C ABI: capsule mass and inertia against the closed form; pair
stream, manifold ids / counts / normals and one iteration from an identical
state against the oracle on a pile of capsules, spheres and boxes; a
60-step pile trajectory against the FP32 and FP64 oracle; the closed-form
rest cases; and the reference's two capsule tests
(GaudiFunicularTests.swift:110-150, GPUSolverTests.swift:577-620)."""
import numpy as np
import pytest

import oracle
import capsule_fixtures as F
from avbd_amd import GPUSolver
from conftest import flush_print

pytestmark = pytest.mark.gpu


def _case(name):
    flush_print(f"[gpu case] {name}")


def test_capsule_mass_and_inertia():
    _case("capsule mass / inertia (GPUSolver.swift:940-945)")
    s, c = F.lying_capsule(length=0.8, radius=0.07)
    g = GPUSolver(s)
    m, inertia = F.capsule_mass(0.8, 0.07, 1.0)
    assert abs(g.bodyMass(c) - m) <= 1e-6 * m


@pytest.mark.parametrize("lanes", [8, 32])
def test_pile_one_iteration_parity(lanes):
    _case(f"capsule pile: pairs, manifolds, one iteration from an identical state, lanes={lanes}")
    s = F.pile()
    src = oracle.OracleSolver(s)
    src.step(90)                                     # a settled, contact-rich state
    g = GPUSolver(s, lanesPerBody=lanes, sweepMode=1)
    o = oracle.OracleSolver(s)
    p, r, v, w = src.bodyStates()
    idx = np.arange(len(s.bodies), dtype=np.uint32)
    g.setBodyStates(idx, p, r, v, w)
    o.setBodyStates(idx, p, r, v, w)
    st = g.settings
    st.iterations = 1
    g.settings = st
    o.set_settings(st)
    g.step()
    o.step()
    assert np.array_equal(g.pairs(), o.pairs())
    gm = g.manifolds()
    h, nrm, _ = o.manifolds()
    assert np.array_equal(gm["header"][:, :3], h)
    act = h[:, 2] > 0
    ncap = sum(1 for b in s.bodies[1:] if b["shape"] == 3)
    flush_print(f"  {len(h)} pairs, {int(act.sum())} active manifolds, {ncap} capsules")
    np.testing.assert_allclose(gm["basisN"][act, :3], nrm[act], atol=1e-5)
    err = float(np.abs(g.positions() - o.positions()).max())
    flush_print(f"  max|x_gpu - x_oracle| = {err:.3e} m")
    assert err <= 1e-5


def test_pile_trajectory_60_steps():
    _case("capsule pile: 60-step trajectory vs oracle FP32 / FP64")
    s = F.pile()
    g = GPUSolver(s)
    o = oracle.OracleSolver(s)
    o64 = oracle.OracleSolver(s, precision=64)
    for _ in range(60):
        g.step()
        o.step()
        o64.step()
    pg, po, p64 = g.positions(), o.positions(), o64.bodyStates64()[0]
    rms = float(np.sqrt(np.mean(np.sum((pg - po) ** 2, axis=1))))
    e64 = float(np.sqrt(np.mean(np.sum((po - p64) ** 2, axis=1))))
    flush_print(f"  RMS gpu-oracle32 {rms:.3e} m, oracle32-oracle64 {e64:.3e} m")
    assert np.all(np.isfinite(pg))
    assert rms < max(2e-3, 3.0 * e64)


def test_closed_form_rests():
    _case("capsule rests: lying on the ground, sphere on a bar, across two bars")
    s, c = F.lying_capsule()
    g = GPUSolver(s)
    g.step(180)
    assert abs(g.positions()[c, 2] - 0.1) < 0.015
    s, cap, sph = F.sphere_on_capsule()
    g = GPUSolver(s)
    g.step(180)
    assert abs(g.positions()[sph, 2] - 0.75) < 0.015
    s, bars, up = F.crossed_capsules()
    g = GPUSolver(s)
    g.step(180)
    p, _, v, _ = g.bodyStates()
    assert abs(p[up, 2] - 0.52) < 0.015 and np.linalg.norm(v[up]) < 0.01


def test_reference_capsule_contracts():
    _case("reference capsule tests: rope/rope + bag/rope contacts, compound capsule foot")
    s, a, b, bag = F.ropes()
    g = GPUSolver(s)
    g.step()
    m = g.manifolds()
    active = {(int(x[0]), int(x[1])) for x in m["header"] if x[2] > 0}
    assert (a, b) in active and (a, bag) in active
    g.step(240)
    p, r, v, _ = g.bodyStates()
    A0, A1 = F.capsule_segment(p[a], r[a], 0.8)
    B0, B1 = F.capsule_segment(p[b], r[b], 0.8)
    assert -(F.segment_distance(A0, A1, B0, B1) - 0.08 + 0.002) < 0.003
    assert np.all(np.linalg.norm(v[[a, b, bag]], axis=1) < 0.1)

    s, foot = F.capsule_foot()
    g = GPUSolver(s)
    g.step(600)
    settled = g.positions()[foot].copy()
    for _ in range(1200):
        q = g.positions()[foot]
        g.setJointWorldAnchors([0], [[q[0] + 0.04, q[1], q[2]]])
        g.step()
    p, _, v, _ = g.bodyStates()
    flush_print(f"  foot creep {p[foot, 0] - settled[0]:.2e} m, speed {np.linalg.norm(v[foot]):.2e}")
    assert abs(p[foot, 0] - settled[0]) < 0.001 and np.linalg.norm(v[foot]) < 0.03


def test_convex_hull_contacts():
    _case("convex hull vs box terrain (GPUSolverTests.swift:537-570), hull vs sphere refused, parity")
    s, ground, body = F.hull_cube()
    g = GPUSolver(s)
    g.step(360)
    p, r, _, _ = g.bodyStates()
    assert abs(p[body, 2] - 0.5) < 0.06 and abs(r[body, 0]) < 0.05 and abs(r[body, 1]) < 0.05
    m = g.manifolds()
    act = [k for k in range(len(m)) if m["header"][k][2] > 0]
    assert len(act) == 1 and tuple(m["header"][act[0]][:3]) == (ground, body, 4)
    feats = sorted(int(x) for x in m["c"][act[0]]["rA"][:4, 3].view(np.uint32))
    assert feats == [(4 << 24) | v for v in (0, 1, 2, 3)]      # the four bottom hull vertices

    s, body, sph = F.hull_wedge()
    g = GPUSolver(s)
    o = oracle.OracleSolver(s)
    for k in range(90):
        g.step()
        o.step()
        gm = g.manifolds()
        h, nrm, _ = o.manifolds()
        assert np.array_equal(gm["header"][:, :3], h), k
        for a, b, n in h:
            if {int(a), int(b)} == {body, sph}:
                assert n == 0, k
    err = float(np.abs(g.positions() - o.positions()).max())
    flush_print(f"  wedge after 90 steps: max|x_gpu - x_oracle| = {err:.3e} m")
    assert err <= 1e-3


def test_capsule_rigid_features_against_cloth():
    _case("capsule rigid features vs cloth triangles (25_cloth.metal:957-970): soft contacts vs the oracle")
    for x in (0.5, -0.5, 0.0):
        s, cap, ids = F.triangle_over_capsule(x=x)
        with GPUSolver(s) as g:
            o = oracle.OracleSolver(s)
            g.step()
            o.step()
            sg, so = g.softContacts(), o.softContacts()
            assert len(sg) == len(so) == 1
            assert np.array_equal(sg["ids"], so["ids"])
            for f in ("lambda", "penalty", "anchorA"):
                assert np.array_equal(sg[f][:, 3].view(np.uint32), so[f][:, 3].view(np.uint32)), f
            np.testing.assert_allclose(sg["C0"], so["C0"], rtol=1e-5, atol=1e-7)
    s = F.sheet_over_capsule()
    src = oracle.OracleSolver(s)
    src.step(40)
    p, r, v, w = src.bodyStates()
    idx = np.arange(len(s.bodies), dtype=np.uint32)
    with GPUSolver(s) as g:
        o = oracle.OracleSolver(s)
        g.setBodyStates(idx, p, r, v, w)
        o.setBodyStates(idx, p, r, v, w)
        st = g.settings
        st.iterations = 1
        g.settings = st
        o.set_settings(st)
        g.step()
        o.step()
        sg, so = g.softContacts(), o.softContacts()
        keyA = so["lambda"][:, 3].view(np.uint32)
        flush_print(f"  sheet: {len(so)} element contacts, {int(((keyA >> 28) == 2).sum())} capsule-feature contacts")
        assert int(((keyA >> 28) == 2).sum()) > 0
        assert len(sg) == len(so)
        assert np.array_equal(sg["ids"], so["ids"])
        for f in ("lambda", "penalty", "anchorA"):
            assert np.array_equal(sg[f][:, 3].view(np.uint32), so[f][:, 3].view(np.uint32)), f
        err = float(np.abs(g.positions() - o.positions()).max())
        flush_print(f"  max|x_gpu - x_oracle| = {err:.3e} m")
        assert err <= 1e-5
