"""The reference's GPU solver tests (Tests/PhysicsAVBDTests/GPUSolverTests.swift)
restated with their own thresholds, run on the oracle (CPU) and on the HIP
backend through the C ABI (GPU): per-body gravity scale, rigid linear /
angular damping against exp(-2), particles exempt from rigid damping, soft
particles and tets free-falling with particle damping, offset compound
colliders, fast boxes and particles against floors, walls and each other,
the structural soft-contact capacity, explicit inertia; and from
ConvergenceTests.swift / SphereTests.swift the friction stop distance and
Coulomb deceleration, a welded compound dissipating its sliding energy,
sphere rest, rolling without slip (v = w r, 5/7 g sin theta), a sliding
sphere staying on the floor and two spheres not interpenetrating; the
static-friction sweep, ratio stack, high-mass-ratio pendulum, chain error
decay, spring stiffness ratio, low-iteration stack and determinism of
RigidCorrectnessTests.swift / ConvergenceTests.swift on their demo scenes."""
import numpy as np
import pytest

import oracle
import reference_scenes as R

BACKENDS = ["oracle", pytest.param("gpu", marks=pytest.mark.gpu)]


def make(kind, scene):
    if kind == "oracle":
        return oracle.OracleSolver(scene)
    from avbd_amd import GPUSolver
    from conftest import flush_print
    flush_print(f"[gpu case] {scene.name}")
    return GPUSolver(scene)


def states(solver):
    return solver.bodyStates()


@pytest.mark.parametrize("kind", BACKENDS)
def test_per_body_gravity_scale(kind):
    s = R.gravity_scale()
    x = make(kind, s)
    x.step(10)
    p = x.positions()
    assert abs(p[0, 2] - 2.0) < 1e-5
    assert p[1, 2] < 1.95
    o = oracle.OracleSolver(s)
    o.step(10)
    assert abs(p[1, 2] - o.positions()[1, 2]) < 2e-3


@pytest.mark.parametrize("kind", BACKENDS)
def test_rigid_linear_and_angular_damping(kind):
    x = make(kind, R.linear_damping())
    x.step(120)
    assert abs(states(x)[2][0, 0] - R.EXP_M2) < 1e-4
    s = R.angular_damping()
    x = make(kind, s)
    x.setBodyStates(np.array([0], np.uint32), np.zeros((1, 3), np.float32), np.array([[0, 0, 0, 1]], np.float32),
                    np.zeros((1, 3), np.float32), np.array([[0, 0, 1]], np.float32))
    x.step(120)
    assert abs(states(x)[3][0, 2] - R.EXP_M2) < 2e-3
    x = make(kind, R.particle_undamped())
    x.step(120)
    assert abs(states(x)[2][0, 0] - 1.0) < 1e-4


@pytest.mark.parametrize("kind", BACKENDS)
def test_soft_particles_free_fall_like_rigid_bodies(kind):
    s, rigid, particle, tet = R.soft_free_fall()
    x = make(kind, s)
    x.step(40)
    p, _, v, _ = states(x)
    z = R.free_fall_z(40)
    assert abs(p[rigid, 2] - z) < 0.02
    assert abs(p[particle, 2] - z) < 0.02
    assert abs(p[tet, 2].mean() - z) < 0.02
    assert abs(v[particle, 2] - v[rigid, 2]) < 0.02


@pytest.mark.parametrize("kind", BACKENDS)
def test_offset_compound_colliders_support_one_body(kind):
    s, b = R.offset_compound()
    x = make(kind, s)
    x.step(360)
    p, r, _, _ = states(x)
    assert abs(p[b, 2] - 0.90) < 0.08
    assert abs(r[b, 0]) < 0.08 and abs(r[b, 1]) < 0.08
    assert len(x.pairs()) >= 2


@pytest.mark.parametrize("kind", BACKENDS)
def test_fast_bodies_do_not_tunnel(kind):
    s, b = R.fast_box_floor()
    x = make(kind, s)
    zmin = np.inf
    for _ in range(20):
        x.step()
        zmin = min(zmin, float(x.positions()[b, 2]))
    assert zmin > 0.45
    assert np.linalg.norm(states(x)[3][b]) < 0.05            # no spurious spin
    s, b = R.fast_box_wall()
    x = make(kind, s)
    xmax = -np.inf
    for _ in range(20):
        x.step()
        xmax = max(xmax, float(x.positions()[b, 0]))
    assert xmax < -0.55
    s, left, right = R.fast_opposing_boxes()
    x = make(kind, s)
    order = np.inf
    for _ in range(20):
        x.step()
        p = x.positions()
        order = min(order, float(p[right, 0] - p[left, 0]))
    p = x.positions()
    assert order > 0 and p[right, 0] - p[left, 0] > 0.99
    s, q = R.fast_particle_floor()
    x = make(kind, s)
    zmin = np.inf
    for _ in range(20):
        x.step()
        zmin = min(zmin, float(x.positions()[q, 2]))
    assert zmin > 0.035


@pytest.mark.parametrize("kind", BACKENDS)
def test_disconnected_soft_mesh_contact_capacity(kind):
    x = make(kind, R.disconnected_soft_mesh())
    assert x.clothInfo()["max_soft"] == 2600


@pytest.mark.gpu
def test_explicit_inertia_overrides_density():
    from avbd_amd import GPUSolver
    g = GPUSolver(R.explicit_inertia())
    assert abs(g.bodyMass(0) - 5.39) < 1e-6
    np.testing.assert_allclose(g.bodyDiagonalInertia(0), [0.0490211, 0.0445821, 0.00824619], atol=1e-7)


@pytest.mark.parametrize("kind", BACKENDS)
def test_friction_stop_distance_and_coulomb_deceleration(kind):
    # ConvergenceTests.swift:51-97
    stops = []
    for mu in (0.2, 0.8):
        x = make(kind, R.slide(mu))
        x.step(300)
        p, _, v, _ = states(x)
        assert np.linalg.norm(v[1]) < 0.1
        want = 25.0 / (2 * mu * 10)
        assert abs(p[1, 0] - want) < 0.3 * want
        stops.append(p[1, 0])
    assert stops[0] > 2 * stops[1]
    x = make(kind, R.slide(0.3, vx=6.0))
    x.step(5)
    v0 = states(x)[2][1, 0]
    x.step(30)
    v1 = states(x)[2][1, 0]
    decel = (v0 - v1) / (30 / 60.0)
    assert abs(decel - 3.0) < 0.25 * 3.0


@pytest.mark.parametrize("kind", BACKENDS)
def test_welded_compound_dissipates_sliding_energy(kind):
    # ConvergenceTests.swift:99-145
    s, bar, stem = R.welded_pusht()
    x = make(kind, s)
    p, r, _, _ = states(x)
    idx = np.array([bar, stem], np.uint32)
    x.setBodyStates(idx, p[idx], r[idx], np.array([[0.8, 0, 0]] * 2, np.float32),
                    np.array([[0, 0, 0.46]] * 2, np.float32))
    x.step(64)
    _, _, v, w = states(x)
    assert np.linalg.norm(0.5 * (v[bar] + v[stem])) < 0.1
    assert np.linalg.norm(0.5 * (w[bar] + w[stem])) < 0.1


@pytest.mark.parametrize("kind", BACKENDS)
def test_sphere_contracts(kind):
    # SphereTests.swift:7-65
    x = make(kind, R.sphere_on_ground())
    x.step(240)
    p, _, v, _ = states(x)
    assert abs(p[1, 2] - 0.5) < 0.04 and np.linalg.norm(v[1]) < 0.05
    x = make(kind, R.sphere_incline())
    x.step(20)
    v0 = np.linalg.norm(states(x)[2][1])
    x.step(60)
    _, _, v, w = states(x)
    v1, w1 = np.linalg.norm(v[1]), np.linalg.norm(w[1])
    assert abs(v1 / max(w1, 0.01) - 0.5) < 0.05                      # v = w r
    accel = (v1 - v0) / (60 / 60.0)
    analytic = 5.0 / 7.0 * 10 * np.sin(0.25)
    assert abs(accel - analytic) < 0.2 * analytic
    x = make(kind, R.sliding_sphere())
    x.step(300)
    assert abs(x.positions()[1, 2] - 0.35) < 0.05
    x = make(kind, R.sphere_sphere())
    x.step(300)
    p = x.positions()
    assert np.linalg.norm(p[1] - p[2]) > 0.93
    assert abs(p[1, 2] - 0.5) < 0.05 and abs(p[2, 2] - 0.5) < 0.05


@pytest.mark.parametrize("kind", BACKENDS)
def test_static_friction_sweep_and_ratio_stack(kind):
    # RigidCorrectnessTests.swift:14-47
    s = R.slopefriction(8)
    x = make(kind, s)
    p0 = x.positions().copy()
    x.step(400)
    p = x.positions()
    for k in (0, 1):
        assert p[2 + k, 0] - p0[2 + k, 0] > 0.5, k             # slick boxes slide
    for k in (6, 7):
        assert np.linalg.norm(p[2 + k] - p0[2 + k]) < 0.08, k   # grippy boxes hold
    s = R.ratiostack(5)
    x = make(kind, s)
    top = len(s.bodies) - 1
    z0 = float(x.positions()[top, 2])
    x.step(600)
    p = x.positions()[top]
    assert abs(p[2] - z0) < 0.1 * z0 and np.hypot(p[0], p[1]) < 0.25


@pytest.mark.parametrize("kind", BACKENDS)
def test_chains_springs_and_low_iteration_stacks(kind):
    # ConvergenceTests.swift:13-49, 147-162, 182-193
    s, az = R.pendulum(20, 1000.0)
    s.settings.iterations = 20
    x = make(kind, s)
    x.step(240)
    bob = x.positions()[len(s.bodies) - 1]
    assert np.linalg.norm(bob - np.array([0, 0, 10])) < (20 * 0.5 + 0.5) * 1.05
    if kind == "gpu":
        assert x.maxConstraintError() < 0.15
    s = R.chain(20)
    s.settings.iterations = 5
    x = make(kind, s)
    x.step(240)
    if kind == "gpu":
        assert x.maxConstraintError() < 0.02
    s, zt = R.spring_ratio()
    s.settings.iterations = 10
    x = make(kind, s)
    x.step(600)
    p, _, v, _ = states(x)
    assert abs(np.linalg.norm(p[1] - np.array([0, 0, 10])) - 2.0) < 0.005
    assert np.linalg.norm(v[1]) < 0.05
    s = R.stack(8)
    s.settings.iterations = 2
    x = make(kind, s)
    x.step(600)
    p = x.positions()
    for i in range(8):
        assert np.linalg.norm(p[1 + i]) < 50 and p[1 + i, 2] > -1


@pytest.mark.gpu
def test_pyramid_bitwise_determinism():
    # ConvergenceTests.testDeterminismGPU (:221-240)
    from avbd_amd import GPUSolver
    a, b = GPUSolver(R.pyramid(5)), GPUSolver(R.pyramid(5))
    a.step(60)
    b.step(60)
    assert np.array_equal(a.positions(), b.positions())


@pytest.mark.parametrize("kind", BACKENDS)
def test_particle_rods_and_pinned_tet(kind):
    # SoftBodyTests.swift:28-68
    x = make(kind, R.particle_rod_chain())
    x.step(300)
    p, r, _, _ = states(x)
    d = np.linalg.norm(p[1:12] - p[:11], axis=1)
    assert np.all(np.abs(d - 0.25) < 0.01), d
    assert abs(r[5, 3] - 1.0) < 1e-4                      # particles carry no rotation
    s, ids = R.pinned_tet()
    x = make(kind, s)
    x.step(300)
    a, b, c, d = x.positions()[ids].astype(np.float64)
    v = abs(np.dot(b - a, np.cross(c - a, d - a))) / 6
    assert abs(v * 6 / 0.5 ** 3 - 1.0) < 0.05


@pytest.mark.parametrize("kind", BACKENDS)
def test_ball_rests_on_soft_block(kind):
    # SoftBodyTests.testRigidRestsOnSoftBlock (:70-83)
    s, ball = R.ball_on_soft_block()
    x = make(kind, s)
    x.step(500)
    assert x.positions()[ball, 2] > 0.8


@pytest.mark.parametrize("kind", BACKENDS)
def test_soft_blocks_do_not_interpenetrate(kind):
    # SoftBodyTests.testSoftBlocksDontInterpenetrate (:98-123)
    s, top = R.stacked_soft_blocks()
    x = make(kind, s)
    x.step(400)
    p = x.positions()[top]
    assert p[:, 2].min() > 0.95 and p[:, 2].mean() > 1.15


@pytest.mark.parametrize("kind", BACKENDS)
def test_combined_tet_cloth_rigid_stable(kind):
    # ClothTests.testCombinedTetClothRigidStable (:156-178): tets, a membrane
    # cloth and rigid bodies in one solve come to rest
    from avbd_amd import demos
    s = R.clothcombo(16)
    g = make(kind, s)
    g.step(600)
    p, _, v, _ = g.bodyStates()
    parts = [b for b in range(len(s.bodies)) if s.bodies[b]["particle"]]
    mass = [s.bodies[b]["density"] * 4.0 / 3.0 * np.pi * (s.bodies[b]["size"][0] / 2) ** 3 for b in parts]
    ke = sum(0.5 * m * float(np.dot(v[b], v[b])) for m, b in zip(mass, parts))
    gap, stretch = demos.cloth_metrics(s, p, g.collisionTris())
    assert ke < 0.2 and gap > -0.06 and stretch < 0.02
    for b in (len(s.bodies) - 2, len(s.bodies) - 1):
        assert p[b, 2] > 0.1 and np.linalg.norm(v[b]) < 0.2


@pytest.mark.gpu
def test_heavy_box_on_cloth_and_hammock():
    # ClothTests.testHeavyBoxRestsOnDrapedCloth (:36-56) and
    # testHammockInextensible (:115-128)
    from avbd_amd import demos
    s = R.boxoncloth(20, 8.0)
    g = make("gpu", s)
    g.step(480)
    p, _, v, _ = g.bodyStates()
    box = len(s.bodies) - 1
    gap, _ = demos.cloth_metrics(s, p, g.collisionTris())
    assert p[box, 2] > 1.05 and max(abs(p[box, 0]), abs(p[box, 1])) < 0.3 and np.linalg.norm(v[box]) < 0.1
    assert gap > -0.05 and g.stats().num_soft > 10
    s = R.hammock(16)
    g = make("gpu", s)
    g.step(480)
    p = g.positions()
    gap, stretch = demos.cloth_metrics(s, p, g.collisionTris())
    assert stretch < 0.02 and gap > -0.06 and p[len(s.bodies) - 1, 2] > 0.5


@pytest.mark.gpu
def test_folded_cloth_and_whipping_flag():
    # ClothTests.testFoldedClothNoInterpenetration (:12-32) and
    # testFlagWhipNoEnergyGrowth (:131-152)
    from avbd_amd import demos
    from conftest import flush_print
    s = R.clothfold(36)
    g = make("gpu", s)
    g.step(600)
    p, _, v, _ = g.bodyStates()
    gap, stretch = demos.cloth_metrics(s, p, g.collisionTris())
    ke = R.particle_ke(s, v)
    flush_print(f"  fold: gap {gap:.4f}, stretch {stretch:.4f}, soft {g.stats().num_soft}, KE {ke:.2e}")
    assert gap > -0.023 and stretch < 0.05 and g.stats().num_soft > 100 and ke < 0.01
    s = R.flagwhip(14)
    g = make("gpu", s)
    windows, cur = [], 0.0
    for f in range(1800):
        g.step()
        if f % 10 == 0:
            cur = max(cur, R.particle_ke(s, g.bodyStates()[2]))
        if (f + 1) % 600 == 0:
            windows.append(cur)
            cur = 0.0
    _, stretch = demos.cloth_metrics(s, g.positions(), g.collisionTris())
    flush_print(f"  flag: KE window maxima {windows}, stretch {stretch:.4f}")
    assert windows[2] < windows[0] * 1.05 and stretch < 0.02


@pytest.mark.gpu
def test_dragged_strips_do_not_pass_at_corners():
    # ClothTests.testDraggedStripsDontPassAtCorners (:58-113): strip B is
    # dragged 1.5 m across taut strip A; no B node may ever sit under an
    # interior point of an A triangle, and E-E contacts must fire
    from conftest import flush_print
    s, drag, aNodes, bNodes = R.eecross()
    g = make("gpu", s)
    g.step(120)
    anchors = np.array([s.joints[j].rA for j in drag], np.float64)
    aSet = set(aNodes)
    aTris = np.array([t[0] for t in s.tris if all(v in aSet for v in t[0])], np.int64)
    bNodes = np.array(bNodes)
    violations = sawEE = 0
    for f in range(300):
        anchors[:, 1] -= 0.30 / 60
        g.setJointWorldAnchors(drag, anchors)
        g.step()
        x = g.positions().astype(np.float64)
        p = x[bNodes]
        p = p[(np.abs(p[:, 0]) < 1.0) & (np.abs(p[:, 1]) < 0.25) & (np.abs(p[:, 2] - 0.12) < 0.25)]
        a, b, c = x[aTris[:, 0]], x[aTris[:, 1]], x[aTris[:, 2]]
        n = np.cross(b - a, c - a)
        n /= np.linalg.norm(n, axis=1)[:, None]
        for q in p:
            sd = np.sum((q - a) * n, axis=1)
            proj = q - n * sd[:, None]
            v0, v1, v2 = b - a, c - a, proj - a
            d00, d01, d11 = np.sum(v0 * v0, 1), np.sum(v0 * v1, 1), np.sum(v1 * v1, 1)
            d20, d21 = np.sum(v2 * v0, 1), np.sum(v2 * v1, 1)
            den = d00 * d11 - d01 * d01
            ok = np.abs(den) > 1e-12
            v = np.where(ok, (d11 * d20 - d01 * d21) / np.where(ok, den, 1), 0)
            w = np.where(ok, (d00 * d21 - d01 * d20) / np.where(ok, den, 1), 0)
            violations += int(np.sum((sd < -1e-4) & ok & (v > 0.1) & (w > 0.1) & (1 - v - w > 0.1)))
        if f % 30 == 0:
            sc = g.softContacts()
            kinds = (sc["anchorA"][:, 3].view(np.uint32) >> 2) & 7
            sawEE += int(np.sum(kinds == 3))
    flush_print(f"  violations {violations}, E-E contacts seen {sawEE}")
    assert violations == 0 and sawEE > 0


@pytest.mark.parametrize("kind", BACKENDS)
def test_flag_stays_attached_to_pole(kind):
    # SoftBodyTests.testFlagStaysAttachedToPole (:384-397)
    s, attach = R.flag_on_pole(14)
    x = make(kind, s)
    x.step(500)
    assert x.positions()[attach, 2] > 3.5


@pytest.mark.parametrize("kind", BACKENDS)
def test_legacy_rate_motor_advances_wrapped_target(kind):
    # GPUSolverTests.testLegacyRateMotorAdvancesWrappedServoTarget (:148-168),
    # and the wrap to [-pi, pi] over later frames against float64
    x = make(kind, R.legacy_rate_motor())
    assert x.motorTargets(1)[0] == 0.0
    x.step()
    assert abs(x.motorTargets(1)[0] - 0.5) <= 1e-6
    want = 0.5
    for _ in range(12):
        x.step()
        want += 0.5
        want -= 2 * np.pi * np.floor(want / (2 * np.pi) + 0.5)
        assert abs(x.motorTargets(1)[0] - want) < 1e-5
        assert -np.pi - 1e-6 <= x.motorTargets(1)[0] <= np.pi + 1e-6
