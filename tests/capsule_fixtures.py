"""Capsule scenes shared by the oracle (CPU) and backend (GPU) tests, after
the reference's capsule tests: crossing rope capsules and a box bag
(GaudiFunicularTests.swift:110-150), the compound capsule foot under a
sub-Coulomb load (GPUSolverTests.swift:577-620), plus closed-form rest
cases (lying capsule on the ground, sphere on a capsule, crossed capsules)
and a mixed pile for trajectory parity."""
import math

import numpy as np

from avbd_amd.demos import SplitMix64
from avbd_amd.scene import BOX, CAPSULE, SPHERE, PhysicsScene, SceneJoint, SimSettings


def quat(angle, axis):
    a = np.asarray(axis, np.float64)
    a = a / np.linalg.norm(a)
    s = math.sin(angle / 2)
    return (float(a[0] * s), float(a[1] * s), float(a[2] * s), float(math.cos(angle / 2)))


ALONG_X = quat(math.pi / 2, (0, 1, 0))   # capsule axis (local z) along world x
ALONG_Y = quat(math.pi / 2, (1, 0, 0))   # along world y


def ropes():
    """testRopeRopeAndBagRopeNarrowphase (GaudiFunicularTests.swift:110-150)."""
    s = PhysicsScene("gaudi-contact-contract", SimSettings(gravity=0.0, dt=1.0 / 120.0, iterations=20,
                                                           collisionMargin=0.002, rigidLinearDamping=3.0,
                                                           rigidAngularDamping=3.0))
    a = s.addCapsule(0.8, 0.04, 1.0, 0.3, (0, 0, 0), rotation=ALONG_X)
    b = s.addCapsule(0.8, 0.04, 1.0, 0.3, (0, 0, 0.055), rotation=ALONG_Y)
    bag = s.addBody((0.12, 0.12, 0.12), 1.0, 0.5, (0.30, 0, 0.075))
    return s, a, b, bag


def capsule_foot():
    """testCompoundCapsuleFootHoldsUnderSubCoulombActuatorLoad
    (GPUSolverTests.swift:577-620): three capsule colliders on a 5 kg foot
    pulled by a 500 N/m world joint whose anchor leads it by 4 cm (20 N
    against ~55 N of available Coulomb friction)."""
    s = PhysicsScene("compound-capsule-static-friction", SimSettings(dt=1.0 / 120.0, iterations=20))
    s.addBody((20, 8, 0.20), 0.0, 1.1, (0, 0, -0.10))
    foot = s.addBody((0.35, 0.20, 0.20), 0.0, 1.1, (0, 0, 0.36), mass=5.0, diagonalInertia=(0.08, 0.12, 0.08),
                     collisionEnabled=False)
    for y in (-0.07, 0.07):
        s.addCollider(foot, (0.24, 0.035, 0), friction=1.1, localPosition=(0, y, -0.31), localRotation=ALONG_X,
                      shape=CAPSULE)
    s.addCollider(foot, (0.14, 0.035, 0), friction=1.1, localPosition=(0.14, 0, -0.31),
                  localRotation=quat(math.pi / 2, (1, 0, 0)), shape=CAPSULE)
    s.addJoint(SceneJoint(-1, foot, (0, 0, 0.36), (0, 0, 0), stiffnessLin=500.0, stiffnessAng=0.0))
    return s, foot


def lying_capsule(length=1.0, radius=0.1, z=0.3):
    s = PhysicsScene("lying capsule", SimSettings(iterations=10))
    s.addBody((20, 20, 2), 0.0, 0.6, (0, 0, -1))
    c = s.addCapsule(length, radius, 1.0, 0.6, (0, 0, z), rotation=ALONG_X)
    return s, c


def sphere_on_capsule(rc=0.1, rs=0.15):
    s = PhysicsScene("sphere on capsule", SimSettings(iterations=10))
    s.addBody((20, 20, 2), 0.0, 0.6, (0, 0, -1))
    cap = s.addCapsule(2.0, rc, 0.0, 0.6, (0, 0, 0.5), rotation=ALONG_X)     # static bar
    sph = s.addSphere(2 * rs, 1.0, 0.6, (0.0, 0.0, 0.5 + rc + rs + 0.05))
    return s, cap, sph


def crossed_capsules(r1=0.05, r2=0.07):
    """A capsule along y laid across two static parallel bars along x: two
    capsule-capsule contacts, rest height 0.4 + r1 + r2."""
    s = PhysicsScene("crossed capsules", SimSettings(iterations=10))
    s.addBody((20, 20, 2), 0.0, 0.6, (0, 0, -1))
    bars = [s.addCapsule(1.5, r1, 0.0, 0.6, (0, y, 0.4), rotation=ALONG_X) for y in (-0.3, 0.3)]
    up = s.addCapsule(1.0, r2, 1.0, 0.6, (0.1, 0.0, 0.4 + r1 + r2 + 0.03), rotation=ALONG_Y)
    return s, bars, up


def pile(n=40, seed=21):
    """Capsules, spheres and boxes dropped in a loose column onto the ground."""
    s = PhysicsScene("capsule pile", SimSettings(iterations=10))
    s.addBody((20, 20, 2), 0.0, 0.6, (0, 0, -1))
    rng = SplitMix64(seed)
    for i in range(n):
        x, y = rng.uniform(-0.4, 0.4), rng.uniform(-0.4, 0.4)
        z = 0.3 + 0.25 * i
        kind = i % 3
        q = quat(rng.uniform(0, math.pi), (rng.uniform(-1, 1), rng.uniform(-1, 1), 1.0))
        if kind == 0:
            s.addCapsule(rng.uniform(0.2, 0.6), rng.uniform(0.04, 0.1), 1.0, 0.5, (x, y, z), rotation=q)
        elif kind == 1:
            s.addSphere(rng.uniform(0.1, 0.25), 1.0, 0.5, (x, y, z))
        else:
            e = rng.uniform(0.1, 0.3)
            s.addBody((e, e, e), 1.0, 0.5, (x, y, z), rotation=q)
    return s


def box_pile(n=36, seed=33):
    """Boxes of mixed aspect ratios and orientations dropped in a loose
    column: face-face, face-edge and edge-edge box contacts."""
    s = PhysicsScene("box pile", SimSettings(iterations=10))
    s.addBody((20, 20, 2), 0.0, 0.6, (0, 0, -1))
    rng = SplitMix64(seed)
    for i in range(n):
        x, y = rng.uniform(-0.35, 0.35), rng.uniform(-0.35, 0.35)
        q = quat(rng.uniform(0, math.pi), (rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(-1, 1)))
        size = (rng.uniform(0.1, 0.5), rng.uniform(0.1, 0.3), rng.uniform(0.05, 0.3))
        s.addBody(size, 1.0, 0.5, (x, y, 0.3 + 0.3 * i), rotation=q)
    return s


def capsule_mass(length, radius, density):
    """GPUSolver.swift:940-945 in float64."""
    m = math.pi * radius * radius * (length + 4 * radius / 3) * density
    return m, (m * (length ** 2 / 12 + radius ** 2 / 4), m * (length ** 2 / 12 + radius ** 2 / 4), 0.5 * m * radius ** 2)


def segment_distance(p0, p1, q0, q1):
    """Closest distance of two segments by dense sampling refinement (float64)."""
    p0, p1, q0, q1 = (np.asarray(v, np.float64) for v in (p0, p1, q0, q1))
    s = np.linspace(0, 1, 401)
    P = p0[None] + (p1 - p0)[None] * s[:, None]
    Q = q0[None] + (q1 - q0)[None] * s[:, None]
    d = np.linalg.norm(P[:, None] - Q[None], axis=2)
    return float(d.min())


def capsule_segment(pos, rot, length):
    x, y, z, w = rot
    ax = np.array([2 * (x * z + w * y), 2 * (y * z - w * x), 1 - 2 * (x * x + y * y)])
    c = np.asarray(pos, np.float64)
    return c - ax * length / 2, c + ax * length / 2


# -- convex hulls ---------------------------------------------------------------
CUBE = [(-0.5, -0.5, -0.5), (0.5, -0.5, -0.5), (-0.5, 0.5, -0.5), (0.5, 0.5, -0.5),
        (-0.5, -0.5, 0.5), (0.5, -0.5, 0.5), (-0.5, 0.5, 0.5), (0.5, 0.5, 0.5)]


def hull_cube():
    """testConvexHullColliderRestsOnBoxTerrain (GPUSolverTests.swift:537-570)."""
    s = PhysicsScene("convex-hull-box-contact", SimSettings(dt=1.0 / 120.0, gravity=-9.81, iterations=20))
    ground = s.addBody((20, 20, 0.2), 0.0, 1.0, (0, 0, -0.1))
    body = s.addBody((1, 1, 1), 0.0, 1.0, (0, 0, 2), mass=1.0, diagonalInertia=(1 / 6, 1 / 6, 1 / 6),
                     collisionEnabled=False)
    s.addConvexCollider(body, CUBE, friction=1.0)
    return s, ground, body


def hull_wedge():
    """A triangular-prism hull (6 vertices, an off-centre AABB) dropped tilted
    onto the ground next to a sphere it overlaps: hull-box contacts with
    up to 8 chosen vertices, the hull-sphere pair refused."""
    s = PhysicsScene("hull wedge", SimSettings(dt=1.0 / 120.0, iterations=20))
    s.addBody((20, 20, 0.2), 0.0, 0.8, (0, 0, -0.1))
    body = s.addBody((0.6, 0.4, 0.3), 0.0, 0.8, (0, 0, 0.6), rotation=quat(0.3, (1, 0.5, 0)), mass=2.0,
                     diagonalInertia=(0.05, 0.08, 0.1), collisionEnabled=False)
    wedge = [(-0.3, -0.2, 0.0), (0.3, -0.2, 0.0), (-0.3, 0.2, 0.0), (0.3, 0.2, 0.0), (-0.3, 0.0, 0.3),
             (0.3, 0.0, 0.3)]
    s.addConvexCollider(body, wedge, friction=0.8, localPosition=(0.05, 0, 0))
    sph = s.addSphere(0.3, 1.0, 0.5, (0.05, 0.0, 0.75))
    return s, body, sph


def triangle_over_capsule(x=0.5, gap=0.003, side=0.08, rp=0.002, rc=0.1):
    """One free particle triangle (flat, edge `side`, particle radius rp) with
    its centroid above x on a static capsule along x (length 1, radius rc,
    axis at z = 0.5), its plane `gap` above the capsule top; zero gravity.
    The capsule's rigid features against the triangle are the spheres at its
    two ends (x = -0.5, 0.5) and middle (x = 0)."""
    s = PhysicsScene("triangle over capsule", SimSettings(gravity=0.0, iterations=4))
    cap = s.addCapsule(1.0, rc, 0.0, 0.6, (0, 0, 0.5), rotation=ALONG_X)
    z = 0.5 + rc + gap
    h = side * math.sqrt(3.0) / 2
    ids = [s.addParticle(rp, 0.001, (x + dx, dy, z)) for dx, dy in ((-side / 2, -h / 3), (side / 2, -h / 3),
                                                                     (0.0, 2 * h / 3))]
    s.addTri(tuple(ids))
    return s, cap, ids


def sheet_over_capsule(n=14, size=0.5, rp=0.003):
    """A particle sheet (springs + collision triangles) dropping across a
    static capsule bar and onto the ground."""
    s = PhysicsScene("sheet over capsule", SimSettings(iterations=20, particleDamping=1.5))
    s.addBody((20, 20, 2), 0.0, 0.6, (0, 0, -1))
    s.addCapsule(0.8, 0.06, 0.0, 0.6, (0, 0, 0.1), rotation=ALONG_X)
    h = size / (n - 1)
    base = len(s.bodies)
    for i in range(n):
        for j in range(n):
            s.addParticle(rp, 0.3 * h * h, (i * h - size / 2, j * h - size / 2, 0.2), friction=0.6)
    idx = lambda i, j: base + i * n + j
    for i in range(n):
        for j in range(n):
            if i + 1 < n:
                s.addSpring(idx(i, j), idx(i + 1, j), (0, 0, 0), (0, 0, 0), 1e4)
            if j + 1 < n:
                s.addSpring(idx(i, j), idx(i, j + 1), (0, 0, 0), (0, 0, 0), 1e4)
    for i in range(n - 1):
        for j in range(n - 1):
            a, b, c, d = idx(i, j), idx(i + 1, j), idx(i, j + 1), idx(i + 1, j + 1)
            s.addTri((a, b, d), 300, 300, 1e-3)
            s.addTri((a, d, c), 300, 300, 1e-3)
    return s
