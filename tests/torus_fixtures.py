"""Torus and kinematic-spinner scenes shared by the oracle (CPU) and backend
(GPU) tests, after the reference's TorusAndMachineTests.swift: a torus lying
on the ground, a ball caught in an upright ring, two linked tori carrying a
weight, the capsule-axle bearing, the spinning roller, the roller conveyor
and the chain mail; plus single-contact closed-form cases, a mixed zoo of
torus pairs near contact, and the brute-force distances that pin the
narrow phase's separations."""
import math

import numpy as np

from avbd_amd import demos
from avbd_amd.demos import SplitMix64
from avbd_amd.scene import BOX, CAPSULE, SPHERE, TORUS, PhysicsScene, SceneJoint, SimSettings


def quat(angle, axis):
    a = np.asarray(axis, np.float64)
    a = a / np.linalg.norm(a)
    s = math.sin(angle / 2)
    return (float(a[0] * s), float(a[1] * s), float(a[2] * s), float(math.cos(angle / 2)))


def qmul(a, b):
    ax, ay, az, aw = a
    bx, by, bz, bw = b
    q = np.array([aw * bx + ax * bw + ay * bz - az * by, aw * by - ax * bz + ay * bw + az * bx,
                  aw * bz + ax * by - ay * bx + az * bw, aw * bw - ax * bx - ay * by - az * bz])
    return tuple(float(v) for v in q / np.linalg.norm(q))


def qrot(q, v):
    x, y, z, w = (float(c) for c in q)
    u = np.array([x, y, z])
    v = np.asarray(v, np.float64)
    t = 2.0 * np.cross(u, v)
    return v + w * t + np.cross(u, t)


UPRIGHT_X = quat(math.pi / 2, (1, 0, 0))    # torus axis (local z) along world -y: spine in the xz plane


def torus_mass(R, r, density):
    """GPUSolver.swift:933-939: solid torus mass and principal moments."""
    m = 2.0 * math.pi * math.pi * R * r * r * density
    return m, (m * (R * R / 2 + 5 * r * r / 8), m * (R * R / 2 + 5 * r * r / 8), m * (R * R + 3 * r * r / 4))


def rests_flat():
    """testTorusRestsFlat (TorusAndMachineTests.swift:7-15)."""
    s = PhysicsScene("t")
    s.addBody((100, 100, 2), 0.0, 0.5, (0, 0, -1))
    t = s.addTorus(0.45, 0.12, 1.0, 0.5, (0, 0, 2))
    return s, t


def interlocked():
    """testInterlockedToriHoldWeight (TorusAndMachineTests.swift:17-30)."""
    s = PhysicsScene("link")
    t1 = s.addTorus(0.45, 0.12, 1.0, 0.4, (0, 0, 6), rotation=UPRIGHT_X)
    t2 = s.addTorus(0.45, 0.12, 1.0, 0.4, (0, 0, 5.34), rotation=qmul(quat(math.pi / 2, (0, 0, 1)), UPRIGHT_X))
    s.addJoint(SceneJoint(-1, t1, (0, 0, 6.4), (0.45, 0, 0)))
    w = s.addBody((0.5, 0.5, 0.5), 8.0, 0.5, (0, 0, 4.4))
    s.addJoint(SceneJoint(t2, w, (0.45, 0, 0), (0, 0, 0.45)))
    return s, w


def ring_catch():
    """testSphereCaughtInRing (TorusAndMachineTests.swift:32-41)."""
    s = PhysicsScene("h")
    s.addBody((100, 100, 2), 0.0, 0.5, (0, 0, -1))
    s.addTorus(0.6, 0.15, 0.0, 0.5, (0, 0, 1.5), rotation=UPRIGHT_X)
    b = s.addSphere(0.7, 1.0, 0.5, (0, 0, 1.52))
    return s, b


def spinner_roller(omega=0.5):
    """testSpinnerFrictionConveys (TorusAndMachineTests.swift:60-71)."""
    s = PhysicsScene("roll")
    roller = s.addBody((2, 6, 2), 0.0, 1.0, (0, 0, 1))
    s.addSpinner(roller, (0, 1, 0), omega)
    box = s.addBody((0.6, 0.6, 0.6), 1.0, 1.0, (0, 0, 2.4))
    return s, roller, box


def bearing():
    """testCapsuleTorusBearing (TorusAndMachineTests.swift:91-104): the hub
    hangs on the axle at (major - minor) - axle radius = 0.045 below it."""
    s = PhysicsScene("rb", SimSettings(iterations=20, betaLin=20000.0))
    s.addCapsule(2.0, 0.11, 0.0, 0.02, (0, 0, 3), rotation=UPRIGHT_X)
    t = s.addTorus(0.27, 0.115, 1.5, 0.02, (0, 0, 3), rotation=UPRIGHT_X)
    return s, t


def spinning_statics(n=5, seed=7):
    """Static bodies with one or two spinners each (mixed axes and rates),
    nothing in contact: their orientations are the spinners' rotations alone."""
    s = PhysicsScene("spinners", SimSettings(gravity=0.0))
    rng = SplitMix64(seed)
    for k in range(n):
        q = quat(rng.uniform(0, 3), (rng.uniform(-1, 1), rng.uniform(-1, 1), 1.0))
        b = s.addBody((0.4, 0.3, 0.2), 0.0, 0.5, (3.0 * k, 0, 1), rotation=q)
        s.addSpinner(b, (rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(-1, 1)), rng.uniform(-4, 4))
        if k % 2 == 1:
            s.addSpinner(b, (0, 0, 1), rng.uniform(-2, 2))
    return s


def torus_over_box(h=0.15, R=0.45, r=0.12):
    """A flat static-free torus with its spine h above the ground's top face,
    zero gravity: every spine sample is h from the box."""
    s = PhysicsScene("torus over box", SimSettings(gravity=0.0))
    s.addBody((100, 100, 2), 0.0, 0.5, (0, 0, -1))
    t = s.addTorus(R, r, 1.0, 0.5, (0, 0, h))
    return s, t


def torus_sphere(offset=(0.0, 0.0, 0.25), R=0.5, r=0.1, rs=0.1, rot=(0, 0, 0, 1)):
    """A static torus at the origin and a sphere of radius rs whose center
    sits at spine point (R, 0, 0) + offset (torus frame), zero gravity."""
    s = PhysicsScene("torus sphere", SimSettings(gravity=0.0))
    t = s.addTorus(R, r, 0.0, 0.5, (0, 0, 0), rotation=rot)
    c = qrot(rot, (R + offset[0], offset[1], offset[2]))
    b = s.addSphere(2 * rs, 1.0, 0.5, tuple(float(v) for v in c))
    return s, t, b


def torus_capsule(e=0.01, R=0.27, r=0.115, rc=0.11):
    """A static capsule along the torus axis through its hole, offset e
    toward +x: the nearest spine point is R - e from the axle."""
    s = PhysicsScene("torus capsule", SimSettings(gravity=0.0))
    cap = s.addCapsule(1.0, rc, 0.0, 0.5, (e, 0, 0))
    t = s.addTorus(R, r, 1.0, 0.5, (0, 0, 0))
    return s, cap, t


def zoo(n=48, seed=31):
    """Torus pairs near contact at seeded poses, zero gravity: each cell holds
    a torus and one partner (sphere, box, capsule, torus) placed so the tube
    passes within reach of it; cells are far apart."""
    s = PhysicsScene("torus zoo", SimSettings(gravity=0.0, iterations=4))
    rng = SplitMix64(seed)
    for k in range(n):
        base = np.array([4.0 * (k % 8), 4.0 * (k // 8), 2.0])
        R = rng.uniform(0.3, 0.6)
        r = rng.uniform(0.06, 0.15)
        qt = quat(rng.uniform(0, math.pi), (rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(-1, 1)))
        s.addTorus(R, r, 1.0, 0.4, tuple(base), rotation=qt)
        kind = k % 4
        ang = rng.uniform(0, 2 * math.pi)
        spine = qrot(qt, (R * math.cos(ang), R * math.sin(ang), 0.0))
        out = spine / np.linalg.norm(spine)
        up = qrot(qt, (0, 0, 1))
        side = rng.uniform(-1, 1)
        d = out * side + up * (1 - abs(side))
        d = d / np.linalg.norm(d)
        q2 = quat(rng.uniform(0, math.pi), (rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(-1, 1)))
        gap = rng.uniform(-0.02, 0.05)
        if kind == 0:
            rs = rng.uniform(0.08, 0.25)
            c = base + spine + d * (r + rs + gap)
            s.addSphere(2 * rs, 1.0, 0.4, tuple(c))
        elif kind == 1:
            e = (rng.uniform(0.1, 0.3), rng.uniform(0.1, 0.3), rng.uniform(0.1, 0.3))
            c = base + spine + d * (r + min(e) / 2 + gap)
            s.addBody(e, 1.0, 0.4, tuple(c), rotation=q2)
        elif kind == 2:
            L, rc = rng.uniform(0.2, 0.8), rng.uniform(0.04, 0.12)
            c = base + spine + d * (r + rc + gap)
            s.addCapsule(L, rc, 1.0, 0.4, tuple(c), rotation=q2)
        else:
            R2, r2 = rng.uniform(0.25, 0.5), rng.uniform(0.06, 0.12)
            # a linked partner: its spine passes through the first torus's hole
            ax2 = np.cross(qrot(qt, (0, 0, 1)), out)
            c = base + spine - out * (R2 - r - r2 - gap)
            q2 = _axis_to(ax2)
            s.addTorus(R2, r2, 1.0, 0.4, tuple(c), rotation=q2)
    return s


def _axis_to(a):
    """A rotation taking local z to the unit direction a."""
    a = np.asarray(a, np.float64) / np.linalg.norm(a)
    z = np.array([0.0, 0.0, 1.0])
    c = float(np.dot(z, a))
    if c < -0.999999:
        return (1.0, 0.0, 0.0, 0.0)
    ax = np.cross(z, a)
    s = math.sqrt(2.0 * (1.0 + c))
    return (float(ax[0] / s), float(ax[1] / s), float(ax[2] / s), float(s / 2))


def treadmill(boxes=5):
    return demos.treadmill_scene(boxes)


def chainmail(rings=5, drops=3):
    return demos.chainmail_scene(rings, drops)


# ---- brute-force distances (float64) ----
def spine_points(center, rot, R, n=20000):
    t = np.linspace(0.0, 2.0 * math.pi, n, endpoint=False)
    local = np.stack([R * np.cos(t), R * np.sin(t), np.zeros_like(t)], axis=1)
    x, y, z, w = rot
    u = np.array([x, y, z])
    tt = 2.0 * np.cross(u, local)
    return np.asarray(center) + local + w * tt + np.cross(u, tt)


def box_distance(p, center, rot, half):
    x, y, z, w = rot
    u = -np.array([x, y, z])
    d = p - np.asarray(center)
    tt = 2.0 * np.cross(u, d)
    loc = d + w * tt + np.cross(u, tt)
    cl = np.clip(loc, -np.asarray(half), np.asarray(half))
    return np.linalg.norm(loc - cl, axis=1)


def segment_distance(p, a, b):
    ab = np.asarray(b) - np.asarray(a)
    t = np.clip(((p - a) @ ab) / max(ab @ ab, 1e-12), 0.0, 1.0)
    return np.linalg.norm(p - (a + t[:, None] * ab), axis=1)


def spine_distance(p, center, rot, R):
    """Distance from points p to the spine circle (closed form per point)."""
    ax = qrot(rot, (0, 0, 1))
    d = p - np.asarray(center)
    h = d @ ax
    radial = np.linalg.norm(d - h[:, None] * ax, axis=1)
    return np.sqrt(h * h + (radial - R) ** 2)


def min_gap(scene, it, io):
    """Brute-force surface gap between torus body `it` and body `io` (their
    colliders are the bodies' own shapes) at the scene's initial poses."""
    bt, bo = scene.bodies[it], scene.bodies[io]
    R, r = bt["size"][0], bt["size"][1]
    pts = spine_points(bt["position"], bt["rotation"], R)
    if bo["shape"] == SPHERE:
        d = np.linalg.norm(pts - np.asarray(bo["position"]), axis=1)
        return float(d.min() - r - bo["size"][0] / 2)
    if bo["shape"] == BOX:
        return float(box_distance(pts, bo["position"], bo["rotation"], np.asarray(bo["size"]) / 2).min() - r)
    if bo["shape"] == CAPSULE:
        ax = qrot(bo["rotation"], (0, 0, 1)) * bo["size"][0] / 2
        a, b = np.asarray(bo["position"]) - ax, np.asarray(bo["position"]) + ax
        return float(segment_distance(pts, a, b).min() - r - bo["size"][1])
    assert bo["shape"] == TORUS
    return float(spine_distance(pts, bo["position"], bo["rotation"], bo["size"][0]).min() - r - bo["size"][1])
