"""Planar-DAT truncation fixtures restated from the reference's analytic
tests (Tests/PhysicsAVBDTests/PlanarDATTests.swift): small triangle scenes
with dt = 1, no gravity and no sweeps, so the predictor and its full
truncation pass fix the result in closed form.  Shared by the oracle tests
(CPU) and the backend tests (GPU).
"""
import numpy as np

from avbd_amd.scene import PhysicsScene, SimSettings

RELAX = 0.85          # params.planarDATRelaxation (GPUSolver.swift:1660)
TIME_MARGIN = 1.0e-3  # DAT_TIME_MARGIN (27_planar_dat.metal:10)


def fraction(hit):
    """Accepted fraction of a displacement whose plane hit is at `hit`
    (datPlaneFraction, 27_planar_dat.metal:536-560)."""
    return min(RELAX * hit, hit - TIME_MARGIN)


def add_triangle(scene, a, b, c, mass, velocity=(0, 0, 0), radius=0.01):
    """addTriangle (PlanarDATTests.swift:31-45): three particles, one triangle."""
    ids = [scene.addParticle(radius, mass, tuple(float(x) for x in p), friction=0.0,
                             velocity=tuple(float(x) for x in velocity)) for p in (a, b, c)]
    scene.addTri(tuple(ids))
    return ids


def _scene(name, iterations=0, dt=1.0):
    return PhysicsScene(name, SimSettings(dt=dt, gravity=0.0, iterations=iterations))


def two_triangles(height, velocity, iterations=0, offset=(0.0, 0.0, 0.0)):
    """twoTriangleScene (PlanarDATTests.swift:48-70): a large static target
    at z = 0 and a small moving source triangle at `height`."""
    s = _scene("planar-dat-two-triangle", iterations)
    o = np.asarray(offset, np.float64)
    target = add_triangle(s, o + (-1, -1, 0), o + (1, -1, 0), o + (0, 1, 0), 0.0)
    source = add_triangle(s, o + (-0.15, -0.10, height), o + (0.15, -0.10, height), o + (0, 0.15, height), 1.0,
                          velocity)
    return s, target, source


def bilateral():
    """testBilateralVertexTrianglePlaneAllocatesRelativeMotion (:768-798)."""
    s = _scene("planar-dat-bilateral-vt")
    target = add_triangle(s, (-1, -1, 0), (1, -1, 0), (0, 1, 0), 1.0, (0, 0, 0.1))
    source = add_triangle(s, (-0.15, -0.10, 0.2), (0.15, -0.10, 0.2), (0, 0.15, 0.2), 1.0, (0, 0, -0.3))
    return s, target, source


def strictest():
    """testStrictestSimultaneousDivisionPlaneWins (:820-850)."""
    s = _scene("planar-dat-strictest-plane")
    add_triangle(s, (-1, -1, 0), (1, -1, 0), (0, 1, 0), 0.0)
    add_triangle(s, (-1, -1, 0.1), (1, -1, 0.1), (0, 1, 0.1), 0.0)
    source = add_triangle(s, (-0.15, -0.10, 0.2), (0.15, -0.10, 0.2), (0, 0.15, 0.2), 1.0, (0, 0, -0.4))
    return s, source


def edge_edge(height=0.2, velocity_z=-0.4):
    """testEdgeEdgePredictorMatchesAnalyticDivisionPlane (:916-944) and its
    exact-hit variant (:946-975); the query radius is set to 0.5."""
    s = _scene("planar-dat-ee")
    add_triangle(s, (0, -1, 0), (0, 1, 0), (0.02, -1, 0), 0.0)
    source = add_triangle(s, (-1, 0, height), (1, 0, height), (-1, 0.02, height), 1.0, (0, 0, velocity_z))
    return s, source


def small_triangles():
    """The small-edge scene of testAcceptedPoseQueryFindsEnteringContact
    (:517-563) and testAcceptedPoseVTContactStoresStartLinearizedC0
    (:565-650): rq comes from the contact reach, not the edge length."""
    s = _scene("planar-dat-entering-contact")
    target = add_triangle(s, (-0.04, -0.04, 0), (0.04, -0.04, 0), (0, 0.04, 0), 0.0)
    source = add_triangle(s, (-0.01, -0.007, 0.5), (0.01, -0.007, 0.5), (0, 0.01, 0.5), 1.0)
    return s, target, source


def place(solver, scene, ids, z=None, velocity=None, x=None):
    """setBodyStates of `ids`: z (and x) replaced, velocity set, identity rotation."""
    n = len(scene.bodies)
    idx = np.asarray(ids, np.uint32)
    p = np.array([scene.bodies[i]["position"] for i in ids], np.float32).reshape(-1, 3)
    if z is not None:
        p[:, 2] = z
    if x is not None:
        p[:, 0] = x
    r = np.tile(np.array([0, 0, 0, 1], np.float32), (len(ids), 1))
    v = np.zeros((len(ids), 3), np.float32)
    if velocity is not None:
        v[:] = np.asarray(velocity, np.float32)
    solver.setBodyStates(idx, p, r, v, np.zeros_like(v))
    assert n == solver.bodyCount


def sheets_fired(gap=0.009, n=10, speed=-3.0):
    """Two sheets, the upper fired at the lower at `speed` m/s: one step
    (dt 1/60) carries it 5 cm, far beyond the gap, so only the truncation
    keeps the sheets apart."""
    s = PhysicsScene("fired sheets", SimSettings(iterations=10, gravity=0.0))
    size, r = 0.12, 0.003
    h = size / (n - 1)
    layers = []
    for layer, z in enumerate((0.02, 0.02 + gap)):
        base = len(s.bodies)
        vel = (0, 0, speed) if layer == 1 else (0, 0, 0)
        for i in range(n):
            for j in range(n):
                s.addParticle(r, 0.3 * h * h, (i * h - size / 2 + 0.3 * h * layer, j * h - size / 2, z),
                              friction=0.3, velocity=vel)
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
        layers.append(np.arange(base, base + n * n))
    return s, layers
