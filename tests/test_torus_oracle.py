"""The oracle's torus narrow phase (30_narrowphase.metal:292-336, 749-965),
kinematic spinners (GPUSolver.swift:628-637, 4418-4427) and the spinner
surface shift (30_narrowphase.metal:274-284), pinned by closed forms and
brute force rather than by restating the formulas:

* torus - sphere: the separation is the distance from the sphere center to
  the spine circle minus both radii; the normal is the radial direction;
* torus - box, flat: every spine sample is equally deep, so the greedy
  choice (lowest sample on ties, neighbours within 3 masked) takes samples
  0, 4, 8, 12, whose points are R (sin t, -cos t) in the spine frame;
* torus - capsule through the hole: the gap is R - e - r - rc;
* a zoo of torus pairs at seeded poses: the deepest separation never
  undercuts the brute-force surface gap (each hit is a real point pair) and
  lies within the two-projection refinement error of it;
* spinners: a static spinning body's orientation after n frames is the
  rotation by n omega dt about the spinner axis applied on the left;
* the reference's contracts (TorusAndMachineTests.swift): a torus rests
  flat, a ball rests in an upright ring, linked tori carry a weight, the
  torus hangs on a capsule axle at the analytic offset, a spinning roller
  conveys a box in the direction of its surface motion (and the reverse
  spin the other way), the roller conveyor carries its cargo, the chain mail
  hangs on its pegs and catches the drops."""
import math

import numpy as np
import pytest

import oracle
import torus_fixtures as F
from avbd_amd import PhysicsScene


def _records(s, steps=1):
    o = oracle.OracleSolver(s)
    o.step(steps)
    return o, o.manifoldRecords()


def _torus_first(rec, torus_body):
    """(torus is collider A, sign that turns the manifold normal into other -> torus)."""
    a = int(rec["header"][0])
    return a == torus_body, (1.0 if a == torus_body else -1.0)


def test_torus_sphere_closed_form():
    for off, rot in (((0.0, 0.0, 0.25), (0, 0, 0, 1)), ((-0.1, 0.12, 0.15), F.quat(0.7, (1, 2, 0.5)))):
        s, t, b = F.torus_sphere(offset=off, rot=rot)
        R, r, rs = 0.5, 0.1, 0.1
        _, m = _records(s)
        assert len(m) == 1 and m["header"][0, 2] == 1
        rec = m[0]
        tA, sg = _torus_first(rec, t)
        # exact geometry in float64: nearest spine point to the sphere center
        c = np.asarray(s.bodies[b]["position"])
        ax = F.qrot(rot, (0, 0, 1))
        d = c - ax * (c @ ax)
        spine = d / np.linalg.norm(d) * R
        gapv = c - spine
        dist = np.linalg.norm(gapv)
        n_ts = gapv / dist                                  # torus -> sphere
        sep = dist - r - rs
        margin = s.settings.collisionMargin
        assert abs(float(rec["c"]["C0"][0, 0]) - margin - sep) < 2e-6
        np.testing.assert_allclose(sg * rec["basisN"][:3], -n_ts, atol=2e-6)
        xT = spine + n_ts * r                               # world points (both colliders world-round)
        xS = c - n_ts * rs
        pA = np.asarray(s.bodies[int(rec["header"][0])]["position"])
        pB = np.asarray(s.bodies[int(rec["header"][1])]["position"])
        np.testing.assert_allclose(rec["c"]["rA"][0, :3] + pA, xT if tA else xS, atol=2e-6)
        np.testing.assert_allclose(rec["c"]["rB"][0, :3] + pB, xS if tA else xT, atol=2e-6)
        assert rec["c"]["rA"][0, 3:4].view(np.uint32)[0] == 0


def test_torus_box_flat_picks_four_quarter_spread_samples():
    h, R, r = 0.15, 0.45, 0.12
    s, t = F.torus_over_box(h=h, R=R, r=r)
    _, m = _records(s)
    assert len(m) == 1
    rec = m[0]
    n = int(rec["header"][2])
    keys = rec["c"]["rA"][:n, 3].copy().view(np.uint32)
    assert n == 4 and list(keys) == [0, 4, 8, 12]
    tA, sg = _torus_first(rec, t)
    np.testing.assert_allclose(sg * rec["basisN"][:3], (0, 0, 1), atol=1e-6)   # box -> torus: up
    margin = s.settings.collisionMargin
    np.testing.assert_allclose(rec["c"]["C0"][:n, 0] - margin, h - r, atol=2e-6)
    # spine frame of a torus with axis z: u = (0, -1, 0), v = (1, 0, 0)
    ang = 2 * math.pi * keys / 32.0
    want = np.stack([R * np.sin(ang), -R * np.cos(ang), np.full(n, h - r)], axis=1)
    side = "rA" if tA else "rB"
    got = rec["c"][side][:n, :3] + np.asarray(s.bodies[t]["position"])   # world-round torus anchors
    np.testing.assert_allclose(got, want, atol=2e-6)


@pytest.mark.parametrize("e", [0.0, 0.01, 0.03])
def test_torus_capsule_axle_gap(e):
    R, r, rc = 0.27, 0.115, 0.11
    s, cap, t = F.torus_capsule(e=e, R=R, r=r, rc=rc)
    _, m = _records(s)
    assert len(m) == 1 and m["header"][0, 2] >= 1
    rec = m[0]
    n = int(rec["header"][2])
    margin = s.settings.collisionMargin
    seps = rec["c"]["C0"][:n, 0] - margin
    assert abs(float(seps.min()) - (R - e - r - rc)) < 2e-6
    if e > 0:
        _, sg = _torus_first(rec, t)
        # the deepest contact sits where the spine passes nearest the axle (+x)
        np.testing.assert_allclose(sg * rec["basisN"][:3], (1, 0, 0), atol=1e-5)


def test_zoo_separations_against_brute_force():
    s = F.zoo()
    _, m = _records(s)
    margin = s.settings.collisionMargin
    checked = {0: 0, 1: 0, 2: 0, 3: 0}
    for rec in m:
        a, b, n = (int(v) for v in rec["header"][:3])
        if n == 0:
            continue
        it, io = (a, b) if s.bodies[a]["shape"] == 2 else (b, a)
        kind = s.bodies[io]["shape"]
        g = F.min_gap(s, it, io)
        sep = float((rec["c"]["C0"][:n, 0] - margin).min())
        assert sep >= g - 1e-5, (a, b)
        assert sep <= g + (5e-3 if kind == 0 else 1e-4), (a, b, sep, g)
        np.testing.assert_allclose(np.linalg.norm(rec["basisN"][:3]), 1.0, atol=1e-5)
        checked[kind] += 1
    assert min(checked.values()) >= 8, checked


def test_spinner_orientation_closed_form():
    s = F.spinning_statics()
    o = oracle.OracleSolver(s, precision=64)
    n = 50
    o.step(n)
    _, rot, _, _ = o.bodyStates64()
    dt = s.settings.dt
    by_body = {}
    for b, axis, omega in s.spinners:
        by_body.setdefault(b, []).append((np.asarray(axis), omega))
    for b, sp in by_body.items():
        q = np.asarray(s.bodies[b]["rotation"], np.float64)
        q = q / np.linalg.norm(q)
        for _ in range(n):
            for axis, omega in sp:
                h = 0.5 * omega * dt
                dq = (*(axis * math.sin(h)), math.cos(h))
                q = np.asarray(F.qmul(dq, q))
        assert abs(abs(float(np.dot(q, rot[b]))) - 1.0) < 1e-6, b


def _spin_box_vx(omega, steps=60):
    s, roller, box = F.spinner_roller(omega)
    o = oracle.OracleSolver(s)
    o.step(steps)
    return float(o.bodyStates()[2][box][0])


def test_spinner_conveys_in_surface_direction():
    # omega about +y moves the roller's top face toward +x
    assert _spin_box_vx(0.5) > 0.3                 # TorusAndMachineTests.swift:60-71
    assert _spin_box_vx(-0.5) < -0.3
    assert abs(_spin_box_vx(0.0)) < 0.05


def test_contract_torus_rests_flat():
    s, t = F.rests_flat()
    o = oracle.OracleSolver(s)
    o.step(400)
    p, q, _, _ = o.bodyStates()
    assert abs(float(p[t][2]) - 0.12) < 0.1
    assert abs(float(F.qrot(q[t], (0, 0, 1))[2])) > 0.95


def test_contract_linked_tori_hold_weight():
    s, w = F.interlocked()
    o = oracle.OracleSolver(s)
    o.step(420)
    assert float(o.positions()[w][2]) > 3.0


def test_contract_sphere_caught_in_ring():
    s, b = F.ring_catch()
    o = oracle.OracleSolver(s)
    o.step(240)
    assert float(o.positions()[b][2]) > 0.8


def test_contract_capsule_torus_bearing():
    s, t = F.bearing()
    o = oracle.OracleSolver(s)
    o.step(300)
    assert abs(float(o.positions()[t][2]) - (3 - 0.045)) < 0.02


def test_contract_roller_conveyor():
    s = F.treadmill(5)
    o = oracle.OracleSolver(s)
    x0 = 0.0
    for f in range(360):
        o.step()
        if f == 89:
            x0 = float(o.positions()[-5:, 0].mean())
    assert float(o.positions()[-5:, 0].mean()) - x0 > 1.0


def test_contract_chainmail_hangs_on_pegs():
    s = F.chainmail(5, 3)
    rings = [i for i, b in enumerate(s.bodies) if b["shape"] == 2]
    o = oracle.OracleSolver(s)
    o.step(700)
    p = o.positions()
    assert float(p[rings, 2].min()) > 4.5
    assert int((p[-3:, 2] > 4.0).sum()) >= 2


def test_oracle_rejects_spinner_on_missing_body():
    s = PhysicsScene("bad")
    s.addBody((1, 1, 1), 1.0, 0.5, (0, 0, 1))
    s.addSpinner(0, (0, 0, 1), 1.0)
    s.spinners[0] = (5, s.spinners[0][1], s.spinners[0][2])
    with pytest.raises(Exception):
        oracle.OracleSolver(s)
