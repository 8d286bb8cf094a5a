"""GPU parity of the torus narrow phase (30_narrowphase.metal:292-336,
749-965), kinematic spinners (GPUSolver.swift:628-637, 4418-4427) and the
spinner surface shift (30_narrowphase.metal:274-284) through the C ABI:
torus mass and inertia against the closed form; a zoo of torus pairs whose
pair stream, manifold headers, feature keys and flags equal the oracle's
bit for bit, with normals and C0 within float rounding, on the first frame
and over warm-started one-iteration frames; spinner orientations bit-exact
against the oracle; the closed-form single contacts; and the reference's
torus and machine contracts (TorusAndMachineTests.swift) run on the GPU."""
import math

import numpy as np
import pytest

import oracle
import torus_fixtures as F
from avbd_amd import GPUSolver
from conftest import flush_print

pytestmark = pytest.mark.gpu


def _case(name):
    flush_print(f"[gpu case] {name}")


def _keys(m):
    k = m["c"]["rA"][:, :, 3].copy().view(np.uint32)
    k[np.arange(8)[None, :] >= m["header"][:, 2][:, None]] = 0xFFFFFFFF
    return k


def _compare_manifolds(gm, om, atol=2e-6):
    assert np.array_equal(gm["header"], om["header"])
    assert np.array_equal(gm["colliderPair"][:, :2], om["colliderPair"][:, :2])
    assert np.array_equal(_keys(gm), _keys(om))
    act = gm["header"][:, 2] > 0
    np.testing.assert_allclose(gm["basisN"][act], om["basisN"][act], atol=atol)
    for k in range(8):
        on = gm["header"][:, 2] > k
        if on.any():
            np.testing.assert_allclose(gm["c"]["C0"][on, k, :3], om["c"]["C0"][on, k, :3], atol=atol)
            np.testing.assert_allclose(gm["c"]["rA"][on, k, :3], om["c"]["rA"][on, k, :3], atol=atol)
            np.testing.assert_allclose(gm["c"]["rB"][on, k, :3], om["c"]["rB"][on, k, :3], atol=atol)
    return int(act.sum())


def test_torus_mass_and_inertia():
    _case("torus mass / inertia (GPUSolver.swift:933-939)")
    s, t = F.rests_flat()
    with GPUSolver(s) as g:
        m, inertia = F.torus_mass(0.45, 0.12, 1.0)
        assert abs(g.bodyMass(t) - m) <= 1e-6 * m
        np.testing.assert_allclose(g.bodyDiagonalInertia(t), inertia, rtol=1e-6)


def test_zoo_first_frame_bit_exact():
    _case("torus zoo: pairs, manifold headers / keys bit-exact, normals and C0 within rounding")
    s = F.zoo()
    with GPUSolver(s) as g:
        o = oracle.OracleSolver(s)
        g.step()
        o.step()
        assert np.array_equal(g.pairs(), o.pairs())
        n = _compare_manifolds(g.manifolds(), o.manifoldRecords())
        flush_print(f"  {len(s.bodies) // 2} torus pairs, {n} active manifolds")
        assert n >= 40


def test_zoo_warm_started_frames():
    _case("torus zoo: 30 warm-started one-iteration frames in lockstep")
    s = F.zoo()
    st = s.settings
    st.iterations = 1
    with GPUSolver(s) as g:
        o = oracle.OracleSolver(s)
        g.settings = st
        o.set_settings(st)
        worst = 0.0
        for k in range(30):
            g.step()
            o.step()
            err = float(np.abs(g.positions() - o.positions()).max())
            worst = max(worst, err)
            assert err <= 1e-5, (k, err)
        gm, om = g.manifolds(), o.manifoldRecords()
        assert np.array_equal(gm["header"][:, :3], om["header"][:, :3])
        act = gm["header"][:, 2] > 0
        np.testing.assert_allclose(gm["c"]["penalty"][act, 0, 0], om["c"]["penalty"][act, 0, 0], rtol=1e-3)
        flush_print(f"  worst max|x_gpu - x_oracle| = {worst:.3e} m, {int(act.sum())} manifolds")


def test_spinner_orientations_bit_exact():
    _case("spinners: 50 frames of static spinning bodies, orientations bit-exact vs the oracle")
    s = F.spinning_statics()
    with GPUSolver(s) as g:
        o = oracle.OracleSolver(s)
        g.step(50)
        o.step(50)
        rg, ro = g.bodyStates()[1], o.bodyStates()[1]
        assert np.array_equal(rg.view(np.uint32), ro.view(np.uint32))
        # the step size change reaches the spinners
        st = g.settings
        st.dt = st.dt * 0.5
        g.settings = st
        o.set_settings(st)
        g.step(10)
        o.step(10)
        assert np.array_equal(g.bodyStates()[1].view(np.uint32), o.bodyStates()[1].view(np.uint32))


def test_closed_form_single_contacts():
    _case("torus - sphere / flat box / capsule axle: single-frame contacts equal the oracle's")
    for s in (F.torus_sphere()[0], F.torus_sphere(offset=(-0.1, 0.12, 0.15), rot=F.quat(0.7, (1, 2, 0.5)))[0],
              F.torus_over_box()[0], F.torus_capsule(e=0.01)[0]):
        with GPUSolver(s) as g:
            o = oracle.OracleSolver(s)
            g.step()
            o.step()
            assert _compare_manifolds(g.manifolds(), o.manifoldRecords()) == 1
    s, t = F.torus_over_box()
    with GPUSolver(s) as g:
        g.step()
        assert list(_keys(g.manifolds())[0, :4]) == [0, 4, 8, 12]


@pytest.mark.parametrize("name", ["spinner", "treadmill", "interlocked"])
def test_machine_lockstep_parity(name):
    _case(f"{name}: one iteration from an identical settled state, manifolds and poses vs the oracle")
    s = {"spinner": lambda: F.spinner_roller()[0], "treadmill": lambda: F.treadmill(5),
         "interlocked": lambda: F.interlocked()[0]}[name]()
    src = oracle.OracleSolver(s)
    src.step(60)
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
        assert np.array_equal(g.pairs(), o.pairs())
        n = _compare_manifolds(g.manifolds(), o.manifoldRecords(), atol=1e-5)
        err = float(np.abs(g.positions() - o.positions()).max())
        flush_print(f"  {n} active manifolds, max|x_gpu - x_oracle| = {err:.3e} m")
        assert err <= 1e-5


def test_contract_torus_rests_flat():
    _case("testTorusRestsFlat (TorusAndMachineTests.swift:7-15)")
    s, t = F.rests_flat()
    with GPUSolver(s) as g:
        g.step(400)
        z = float(g.bodyPosition(t)[2])
        up = float(F.qrot(g.bodyRotation(t), (0, 0, 1))[2])
        flush_print(f"  z = {z:.4f}, axis z = {up:.5f}")
        assert abs(z - 0.12) < 0.1
        assert abs(up) > 0.95


def test_contract_linked_tori_hold_weight():
    _case("testInterlockedToriHoldWeight (TorusAndMachineTests.swift:17-30)")
    s, w = F.interlocked()
    with GPUSolver(s) as g:
        g.step(420)
        assert float(g.bodyPosition(w)[2]) > 3.0


def test_contract_sphere_caught_in_ring():
    _case("testSphereCaughtInRing (TorusAndMachineTests.swift:32-41)")
    s, b = F.ring_catch()
    with GPUSolver(s) as g:
        g.step(240)
        assert float(g.bodyPosition(b)[2]) > 0.8


def test_contract_capsule_torus_bearing():
    _case("testCapsuleTorusBearing (TorusAndMachineTests.swift:91-104)")
    s, t = F.bearing()
    with GPUSolver(s) as g:
        g.step(300)
        z = float(g.bodyPosition(t)[2])
        flush_print(f"  hub z = {z:.4f} (analytic {3 - 0.045:.4f})")
        assert abs(z - (3 - 0.045)) < 0.02


def test_contract_spinner_conveys():
    _case("testSpinnerFrictionConveys (TorusAndMachineTests.swift:60-71), both spin directions")
    for omega, sign in ((0.5, 1.0), (-0.5, -1.0)):
        s, roller, box = F.spinner_roller(omega)
        with GPUSolver(s) as g:
            g.step(60)
            assert sign * float(g.bodyVelocity(box)[0]) > 0.3


def test_contract_roller_conveyor():
    _case("testRollerConveyor (TorusAndMachineTests.swift:73-87)")
    s = F.treadmill(5)
    with GPUSolver(s) as g:
        x0 = 0.0
        for f in range(360):
            g.step()
            if f == 89:
                x0 = float(g.positions()[-5:, 0].mean())
        dx = float(g.positions()[-5:, 0].mean()) - x0
        flush_print(f"  cargo conveyed {dx:.3f} m")
        assert dx > 1.0


def test_contract_chainmail_hangs_on_pegs():
    _case("testChainmailHangsOnPegs (TorusAndMachineTests.swift:43-58)")
    s = F.chainmail(5, 3)
    rings = [i for i, b in enumerate(s.bodies) if b["shape"] == 2]
    with GPUSolver(s) as g:
        g.step(700)
        p = g.positions()
        flush_print(f"  lowest ring z = {float(p[rings, 2].min()):.3f}, drops z = {p[-3:, 2]}")
        assert float(p[rings, 2].min()) > 4.5
        assert int((p[-3:, 2] > 4.0).sum()) >= 2
