"""Hinge motors and limits (40_solver.metal:741-908, warm start :546-655)
against closed forms, on the oracle (CPU) and on the HIP backend (GPU).

* Velocity motor (motorMode 2): the torque kd (w_target - w), clamped to the
  effort limit, is evaluated at the frame start and held over the frame, so
  a body spinning freely about a principal axis follows
  w_{n+1} = w_n + dt * clamp(kd (w_t - w_n), +-limit) / I  exactly: the
  unsaturated motor converges geometrically onto its rate, the saturated one
  accelerates at limit / I.
* Implicit PD motor (motorMode 0): the motor adds kp/2 (theta - theta_t)^2
  to the frame's energy, so a hinged bar held against gravity settles where
  kp (theta - theta_t) = m g (L/2) cos(theta)  (theta measured from the
  horizontal, positive downwards).
* Limit: outside [lo, hi] the hinge adds kL/2 over^2 with kL = 4e4, so a bar
  falling onto its upper limit rests at  kL (theta - hi) = m g (L/2) cos(theta).

A wrong sign, a dropped clamp or a changed stiffness moves these answers by
far more than the tolerances."""
import math

import numpy as np
import pytest
from scipy.optimize import brentq

import oracle
from avbd_amd.scene import PhysicsScene, SceneJoint, SimSettings

BACKENDS = ["oracle", pytest.param("gpu", marks=pytest.mark.gpu)]
LIMIT_STIFFNESS = 4.0e4


def make(kind, scene):
    if kind == "oracle":
        return oracle.OracleSolver(scene)
    from avbd_amd import GPUSolver
    from conftest import flush_print
    flush_print(f"[gpu case] {scene.name}")
    return GPUSolver(scene)


def twist_about(q, axis):
    """Rotation angle of quaternion q (x, y, z, w) about a unit axis."""
    v = q[0] * axis[0] + q[1] * axis[1] + q[2] * axis[2]
    w = q[3]
    if w < 0:
        v, w = -v, -w
    return 2.0 * math.atan2(v, w)


def spinner(mode_limit, kd, target, size=(1.0, 0.4, 0.1), density=10.0, dt=1.0 / 60.0):
    """A box hinged about world z through its centre, no gravity."""
    s = PhysicsScene("velocity motor", SimSettings(dt=dt, gravity=0.0, iterations=20))
    b = s.addBody(size, density, 0.0, (0, 0, 1.0), collisionEnabled=False)
    s.addJoint(SceneJoint(-1, b, (0, 0, 1.0), (0, 0, 0), stiffnessLin=math.inf, stiffnessAng=math.inf,
                          hingeAxis=(0, 0, 1), motorMode=2, motorTarget=target, motorTorque=mode_limit,
                          motorDamping=kd))
    m = np.float32(density) * np.float32(size[0]) * np.float32(size[1]) * np.float32(size[2])
    inertia = float(m) * (size[0] ** 2 + size[1] ** 2) / 12.0
    return s, b, inertia


def bar(length=1.0, kp=0.0, kd=0.0, limits=(1.0, -1.0), mass=1.0, ang_damping=0.0, iterations=20):
    """A bar along x hinged about world y at its left end (pivot at z = 1),
    starting horizontal; gravity 10 pulls its free end down (positive twist)."""
    s = PhysicsScene("hinged bar", SimSettings(iterations=iterations, rigidAngularDamping=ang_damping))
    w = 0.1
    density = mass / (length * w * w)
    b = s.addBody((length, w, w), density, 0.0, (length / 2, 0, 1.0), collisionEnabled=False)
    s.addJoint(SceneJoint(-1, b, (0, 0, 1.0), (-length / 2, 0, 0), stiffnessLin=math.inf, stiffnessAng=math.inf,
                          hingeAxis=(0, 1, 0), motorMode=0, motorTarget=0.0, motorStiffness=kp,
                          motorDamping=kd, motorTorque=1.0e4 if kp > 0 else 0.0,
                          limitLo=limits[0], limitHi=limits[1]))
    m = float(np.float32(density) * np.float32(length) * np.float32(w) * np.float32(w))
    return s, b, m


@pytest.mark.parametrize("kind", BACKENDS)
def test_velocity_motor_follows_explicit_rate_update(kind):
    dt = 1.0 / 60.0
    # unsaturated: geometric approach to the target rate
    s, b, inertia = spinner(mode_limit=100.0, kd=0.2, target=2.0, dt=dt)
    assert dt * 0.2 / inertia < 0.1               # a stable, non-oscillating explicit update
    x = make(kind, s)
    w = 0.0
    for n in range(120):
        x.step()
        w += dt * float(np.clip(0.2 * (2.0 - w), -100.0, 100.0)) / inertia
        got = x.bodyStates()[3][b, 2]
        assert abs(got - w) <= 2e-3 * max(1.0, abs(w)), (n, got, w)
    assert abs(got - 2.0) < 5e-3
    # saturated: constant angular acceleration limit / I until the target
    s, b, inertia = spinner(mode_limit=0.02, kd=1.0e3, target=5.0, dt=dt)
    x = make(kind, s)
    for n in range(1, 61):
        x.step()
        want = min(5.0, n * dt * 0.02 / inertia)
        got = x.bodyStates()[3][b, 2]
        assert abs(got - want) <= 2e-3 * max(1.0, want), (n, got, want)
    # the twist integrates the rate (checks the motor acts about the hinge axis)
    ang = twist_about(x.bodyStates()[1][b], (0, 0, 1))
    assert abs(ang) > 0.1


@pytest.mark.parametrize("kind", BACKENDS)
def test_pd_motor_static_angle(kind):
    kp, L, m_req, g = 20.0, 1.0, 1.0, 10.0
    s, b, m = bar(length=L, kp=kp, kd=4.0, mass=m_req)
    x = make(kind, s)
    x.step(420)
    theta = twist_about(x.bodyStates()[1][b], (0, 1, 0))
    want = brentq(lambda t: kp * t - m * g * (L / 2) * math.cos(t), 0.0, math.pi / 2)
    rate = float(np.linalg.norm(x.bodyStates()[3][b]))
    assert rate < 1e-2, rate
    assert abs(theta - want) < 1e-4, (theta, want)
    # the same bar without the motor would hang straight down
    assert want < 0.5


@pytest.mark.parametrize("kind", BACKENDS)
def test_limit_holds_bar_at_upper_stop(kind):
    L, g, hi = 1.0, 10.0, 0.3
    s, b, m = bar(length=L, limits=(-0.3, hi), ang_damping=3.0)
    x = make(kind, s)
    x.step(420)
    theta = twist_about(x.bodyStates()[1][b], (0, 1, 0))
    want = brentq(lambda t: LIMIT_STIFFNESS * (t - hi) - m * g * (L / 2) * math.cos(t), hi, hi + 0.1)
    assert want - hi > 5e-5                      # the stop's compliance is resolved by the tolerance below
    assert abs(theta - want) < 1e-5, (theta, want, hi)
