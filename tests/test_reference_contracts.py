"""The reference's GPU solver tests (Tests/PhysicsAVBDTests/GPUSolverTests.swift)
restated with their own thresholds, run on the oracle (CPU) and on the HIP
backend through the C ABI (GPU): per-body gravity scale, rigid linear /
angular damping against exp(-2), particles exempt from rigid damping, soft
particles and tets free-falling with particle damping, offset compound
colliders, fast boxes and particles against floors, walls and each other,
the structural soft-contact capacity, explicit inertia."""
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
