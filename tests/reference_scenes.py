"""Small scenes of the reference's GPU solver tests
(Tests/PhysicsAVBDTests/GPUSolverTests.swift), restated with their
thresholds; shared by the oracle (CPU) and backend (GPU) tests."""
import math

from avbd_amd.scene import SPHERE, PhysicsScene, SimSettings

EXP_M2 = math.exp(-2.0)


def gravity_scale():
    """testPerBodyGravityScaleMatchesCPUAndGPU (:7-30)."""
    s = PhysicsScene("gravity-scale", SimSettings(dt=1.0 / 100.0, iterations=4))
    s.addBody((0.1, 0.1, 0.1), 1000.0, 0.0, (0, 0, 2), shape=SPHERE, gravityScale=0.0)
    s.addBody((0.1, 0.1, 0.1), 1000.0, 0.0, (2, 0, 2), shape=SPHERE, gravityScale=1.0)
    return s


def linear_damping():
    """testRigidLinearDampingMatchesCPUAndGPU (:32-53): v = exp(-2) after 1 s."""
    s = PhysicsScene("rigid-damping", SimSettings(gravity=0.0, dt=1.0 / 120.0, rigidLinearDamping=2.0))
    s.addBody((0.2, 0.2, 0.2), 1000.0, 0.0, (0, 0, 0), velocity=(1, 0, 0), shape=SPHERE)
    return s


def angular_damping():
    """testRigidAngularDampingMatchesCPUAndGPU (:55-84): w = exp(-2) after 1 s."""
    s = PhysicsScene("rigid-angular-damping", SimSettings(gravity=0.0, dt=1.0 / 120.0, rigidAngularDamping=2.0))
    s.addBody((0.3, 0.2, 0.1), 1000.0, 0.0, (0, 0, 0))
    return s


def particle_undamped():
    """testRigidDampingDoesNotDampParticles (:86-106)."""
    s = PhysicsScene("rigid-damping-particle-isolation",
                     SimSettings(gravity=0.0, dt=1.0 / 120.0, rigidLinearDamping=2.0, rigidAngularDamping=2.0))
    s.addParticle(0.04, 0.01, (0, 0, 0), friction=0.0, velocity=(1, 0, 0))
    return s


def disconnected_soft_mesh():
    """testDisconnectedSoftMeshUsesStructuralContactCapacity (:386-403): maxSoft 2600."""
    s = PhysicsScene("disconnected-soft-capacity")
    for t in range(50):
        ids = [s.addParticle(0.01, 0.01, (2.0 * t, float(c), 1.0), friction=0.0) for c in range(3)]
        s.addTri(tuple(ids))
    return s


def explicit_inertia():
    """testExplicitImportedInertiaOverridesCollisionPrimitiveDensity (:486-506)."""
    s = PhysicsScene("explicit-inertia")
    s.addBody((0.3, 0.2, 0.1), 0.0, 0.5, (0, 0, 10), mass=5.39, diagonalInertia=(0.0490211, 0.0445821, 0.00824619))
    return s


def offset_compound():
    """testOffsetCompoundCollidersSupportOneRigidBody (:508-535): COM rests near 0.90 m."""
    s = PhysicsScene("compound-collider-support", SimSettings(dt=1.0 / 120.0, iterations=20))
    s.addBody((20, 20, 0.1), 0.0, 1.0, (0, 0, -0.05))
    b = s.addBody((0.4, 0.4, 0.4), 0.0, 0.9, (0, 0, 1.5), mass=2.0, diagonalInertia=(0.25, 0.25, 0.25),
                  collisionEnabled=False)
    s.addCollider(b, (0.30, 0.30, 0.20), localPosition=(-0.35, 0, -0.80))
    s.addCollider(b, (0.30, 0.30, 0.20), localPosition=(0.35, 0, -0.80))
    return s, b


def soft_free_fall():
    """testSoftParticlesFreeFallLikeRigidBodiesWithDamping (:623-656)."""
    s = PhysicsScene("soft-rigid-freefall", SimSettings(particleDamping=20.0))
    rigid = s.addBody((0.4, 0.4, 0.4), 1.0, 0.5, (-2, 0, 20))
    particle = s.addParticle(0.08, 0.02, (0, 0, 20))
    tet = [s.addParticle(0.05, 0.02, p) for p in ((2, 0, 19.75), (2.4, 0, 19.75), (2, 0.4, 19.75), (2, 0, 20.75))]
    s.addTet(tuple(tet), 2000.0, 20000.0)
    return s, rigid, particle, tet


def free_fall_z(steps, dt=1.0 / 60.0, g=-10.0, z0=20.0):
    v, z = 0.0, z0
    for _ in range(steps):
        v += g * dt
        z += v * dt
    return z


def _floor(s):
    s.addBody((200, 200, 2), 0.0, 0.7, (0, 0, -1))


def fast_box_floor():
    """testFastBoxDoesNotStartBelowFloor / ...DoesNotAddSpuriousSpin (:666-693)."""
    s = PhysicsScene("fast-floor-box")
    _floor(s)
    b = s.addBody((1, 1, 1), 1.0, 0.5, (0, 0, 0.72), velocity=(0, 0, -18))
    return s, b


def fast_box_wall():
    """testFastBoxDoesNotTunnelThroughWall (:695-710)."""
    s = PhysicsScene("fast-wall-box", SimSettings(gravity=0.0))
    s.addBody((0.2, 20, 20), 0.0, 0.7, (0, 0, 0))
    b = s.addBody((1, 1, 1), 1.0, 0.5, (-0.72, 0, 0), velocity=(18, 0, 0))
    return s, b


def fast_opposing_boxes():
    """testFastOpposingDynamicBoxesDoNotTunnelThroughEachOther (:712-739)."""
    s = PhysicsScene("fast-opposing-boxes", SimSettings(gravity=0.0, iterations=20))
    left = s.addBody((1, 1, 1), 1.0, 0.0, (-0.6, 0, 0), velocity=(42, 0, 0))
    right = s.addBody((1, 1, 1), 1.0, 0.0, (0.6, 0, 0), velocity=(-42, 0, 0))
    return s, left, right


def fast_particle_floor():
    """testFastParticleDoesNotStartBelowFloor (:741-755)."""
    s = PhysicsScene("fast-floor-particle")
    _floor(s)
    p = s.addParticle(0.05, 0.02, (0, 0, 0.16), friction=0.5, velocity=(0, 0, -8))
    return s, p


def _quat(angle, axis):
    n = math.sqrt(sum(a * a for a in axis))
    s = math.sin(angle / 2)
    return (axis[0] / n * s, axis[1] / n * s, axis[2] / n * s, math.cos(angle / 2))


def slide(mu, vx=5.0):
    """testFrictionStopsSlidingBox / testFrictionMagnitudeMatchesCoulomb
    (ConvergenceTests.swift:51-97)."""
    s = PhysicsScene("slide")
    s.addBody((200, 200, 2), 0.0, mu, (0, 0, -1))
    s.addBody((1, 1, 1), 1.0, mu, (0, 0, 0.5), velocity=(vx, 0, 0))
    return s


def welded_pusht():
    """testWeldedPushTCompoundDissipatesSlidingEnergy (ConvergenceTests.swift:99-145)."""
    from avbd_amd.scene import SceneJoint
    s = PhysicsScene("welded-pusht-friction", SimSettings(iterations=16, betaLin=20000.0, lambdaMax=600.0))
    s.addBody((20, 20, 2), 0.0, 1.3, (0, 0, -1))
    bar = s.addBody((1, 0.25, 0.18), 3.0, 1.1, (0, 0.125, 0.09), collisionGroup=1)
    stem = s.addBody((0.25, 0.65, 0.18), 3.0, 1.1, (0, -0.325, 0.09), collisionGroup=1)
    pb, ps = s.bodies[bar]["position"], s.bodies[stem]["position"]
    mid = tuple(0.5 * (a + b) for a, b in zip(pb, ps))
    s.addJoint(SceneJoint(bar, stem, tuple(m - a for m, a in zip(mid, pb)), tuple(m - b for m, b in zip(mid, ps)),
                          stiffnessLin=math.inf, stiffnessAng=math.inf))
    s.addJoint(SceneJoint(bar, stem, (0, 0, 0), (0, 0, 0), stiffnessLin=0.0, stiffnessAng=0.0))
    return s, bar, stem


def sphere_on_ground():
    """testSphereRestsOnGround (SphereTests.swift:7-17)."""
    s = PhysicsScene("s")
    s.addBody((100, 100, 2), 0.0, 0.5, (0, 0, -1))
    s.addSphere(1.0, 1.0, 0.5, (0, 0, 2))
    return s


def sphere_incline(theta=0.25):
    """testSphereRollsWithoutSlip (SphereTests.swift:19-39)."""
    s = PhysicsScene("r")
    s.addBody((40, 10, 1), 0.0, 0.6, (0, 0, 0), rotation=_quat(theta, (0, 1, 0)))
    s.addSphere(1.0, 1.0, 0.6, (0, 0, 1.2))
    return s


def sliding_sphere():
    """testSlidingSphereStaysOnGround (SphereTests.swift:41-49)."""
    s = PhysicsScene("sl")
    s.addBody((200, 200, 2), 0.0, 0.3, (0, 0, -1))
    s.addSphere(0.7, 1.0, 0.3, (0, 0, 1), velocity=(8, 0, 0))
    return s


def sphere_sphere():
    """testSphereSphereCollision (SphereTests.swift:51-65)."""
    s = PhysicsScene("ss")
    s.addBody((100, 100, 2), 0.0, 0.5, (0, 0, -1))
    s.addSphere(1.0, 1.0, 0.5, (0, 0, 0.5))
    s.addSphere(1.0, 1.0, 0.5, (0.1, 0, 1.6))
    return s


# -- Demos.swift scenes the ConvergenceTests / RigidCorrectnessTests use ----------
def _ground(s, friction=0.7):
    s.addBody((200, 200, 2), 0.0, friction, (0, 0, -1))


def _rot_act(q, v):
    x, y, z, w = q
    u = (x, y, z)
    t = [2 * (u[1] * v[2] - u[2] * v[1]), 2 * (u[2] * v[0] - u[0] * v[2]), 2 * (u[0] * v[1] - u[1] * v[0])]
    return (v[0] + w * t[0] + (u[1] * t[2] - u[2] * t[1]), v[1] + w * t[1] + (u[2] * t[0] - u[0] * t[2]),
            v[2] + w * t[2] + (u[0] * t[1] - u[1] * t[0]))


def slopefriction(count=8):
    """Demos.slopefriction (Demos.swift:245-276): a 30 degree mu = 1 ramp with
    boxes whose pair friction sqrt(mu) sweeps 0.30..0.78 around tan 30."""
    from avbd_amd.scene import PhysicsScene as PS
    s = PS("slopefriction", SimSettings(iterations=10))
    _ground(s, 0.8)
    tilt = math.pi / 6
    rot = _quat(tilt, (0, 1, 0))
    slope_t, slope_len = 0.4, 32.0
    cz = slope_len / 2 * math.sin(tilt) + 0.4
    s.addBody((slope_len, count * 1.3 + 1.5, slope_t), 0.0, 1.0, (0, 0, cz), rotation=rot)
    n = max(2, count)
    side = 0.8
    lift = _rot_act(rot, (0, 0, slope_t / 2 + side / 2 + 0.01))
    for k in range(n):
        pair = 0.30 + 0.48 * k / (n - 1)
        y = (k - (n - 1) / 2) * 1.3
        off = _rot_act(rot, (-slope_len * 0.32, y, 0))
        s.addBody((side, side, side), 1.0, pair * pair,
                  (off[0] + lift[0], off[1] + lift[1], cz + off[2] + lift[2]), rotation=rot)
    return s


def ratiostack(levels=5, growth=1.45):
    """Demos.ratiostack (Demos.swift:278-300)."""
    s = PhysicsScene("ratiostack", SimSettings(iterations=25, betaLin=40000.0))
    _ground(s, 0.8)
    side, z = 0.5, 0.0
    for _ in range(max(2, levels)):
        z += side / 2
        s.addBody((side * 0.98,) * 3, 1.0, 0.8, (0, 0, z))
        z += side / 2
        side *= growth
    return s


def pendulum(links, mass_ratio):
    """Demos.pendulum (Demos.swift:314-341): a chain of links ending in a heavy bob."""
    from avbd_amd.scene import SceneJoint
    s = PhysicsScene("pendulum")
    L = 0.5
    az = links * L + 2.0
    prev = -1
    for i in range(links):
        x = (i + 1) * L
        idx = s.addBody((L, 0.1, 0.1), 1.0, 0.5, (x - L / 2, 0, az))
        s.addJoint(SceneJoint(prev, idx, (0, 0, az) if prev < 0 else (L / 2, 0, 0), (-L / 2, 0, 0)))
        prev = idx
    bob = 1.0
    density = mass_ratio * (L * 0.1 * 0.1) / bob ** 3
    b = s.addBody((bob,) * 3, density, 0.5, (links * L + bob / 2, 0, az))
    s.addJoint(SceneJoint(prev, b, (L / 2, 0, 0), (-bob / 2, 0, 0)))
    return s, az


def chain(links):
    """Demos.chain (Demos.swift:343-358)."""
    from avbd_amd.scene import SceneJoint
    s = PhysicsScene("chain")
    _ground(s)
    L = 0.6
    prev = -1
    for i in range(links):
        z = 20 - i * L
        idx = s.addBody((0.15, 0.15, L), 1.0, 0.5, (0, 0, z - L / 2))
        s.addJoint(SceneJoint(prev, idx, (0, 0, 20) if prev < 0 else (0, 0, -L / 2), (0, 0, L / 2)))
        prev = idx
    return s


def spring_ratio(blocks=3):
    """Demos.springRatio (Demos.swift:385-400): springs of 1e6 / 1e2 stiffness."""
    s = PhysicsScene("spring")
    z_top = 4 + 2 * blocks
    prev = s.addBody((1, 1, 0.5), 0.0, 0.5, (0, 0, z_top))
    for i in range(1, blocks):
        b = s.addBody((1, 1, 0.5), 1.0, 0.5, (0, 0, z_top - i * 2))
        s.addSpring(prev, b, (0, 0, 0), (0, 0, 0), 1e6 if i % 2 == 1 else 1e2)
        prev = b
    return s, z_top


def stack(height):
    """Demos.stack (Demos.swift:214-222)."""
    s = PhysicsScene("stack")
    _ground(s)
    for i in range(height):
        s.addBody((1, 1, 1), 1.0, 0.7, (0, 0, 0.5 + i))
    return s


def pyramid(base):
    """Demos.pyramid (Demos.swift:302-312)."""
    s = PhysicsScene("pyramid")
    _ground(s)
    for j in range(base):
        for i in range(base - j):
            s.addBody((0.95, 0.95, 0.95), 1.0, 0.8, (i + j / 2 - base / 2, 0, 0.5 + j))
    return s


# -- SoftBodyTests.swift scenes --------------------------------------------------
def add_soft_lattice(s, center, nx, ny, nz, spacing, mu=3000.0, lam=30000.0, massPerNode=0.05, friction=0.8):
    """Demos.addSoftBlock (Demos5.swift:76-145): a particle lattice, five
    tets per cell with the split alternating on (i + j + k) parity, and inert
    (stiffness 0) joints between lattice neighbours as collision exclusions.
    Returns the particle ids in (i, j, k) order."""
    from avbd_amd.scene import SceneJoint
    half = [(n - 1) * spacing / 2 for n in (nx, ny, nz)]
    ids = {}
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                p = (center[0] + i * spacing - half[0], center[1] + j * spacing - half[1],
                     center[2] + k * spacing - half[2])
                ids[i, j, k] = s.addParticle(spacing * 0.32, massPerNode, p, friction=friction)
    for i in range(nx - 1):
        for j in range(ny - 1):
            for k in range(nz - 1):
                c = {(a, b, d): ids[i + a, j + b, k + d] for a in (0, 1) for b in (0, 1) for d in (0, 1)}
                if (i + j + k) % 2 == 0:
                    split = [(c[0, 0, 0], c[1, 0, 0], c[0, 1, 0], c[0, 0, 1]), (c[1, 1, 0], c[1, 0, 0], c[0, 1, 0], c[1, 1, 1]),
                             (c[1, 0, 1], c[1, 0, 0], c[0, 0, 1], c[1, 1, 1]), (c[0, 1, 1], c[0, 1, 0], c[0, 0, 1], c[1, 1, 1]),
                             (c[1, 0, 0], c[0, 1, 0], c[0, 0, 1], c[1, 1, 1])]
                else:
                    split = [(c[1, 0, 0], c[0, 0, 0], c[1, 1, 0], c[1, 0, 1]), (c[0, 1, 0], c[0, 0, 0], c[1, 1, 0], c[0, 1, 1]),
                             (c[0, 0, 1], c[0, 0, 0], c[1, 0, 1], c[0, 1, 1]), (c[1, 1, 1], c[1, 1, 0], c[1, 0, 1], c[0, 1, 1]),
                             (c[0, 0, 0], c[1, 1, 0], c[1, 0, 1], c[0, 1, 1])]
                for t in split:
                    s.addTet(t, mu, lam)
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                for di, dj, dk in ((1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (1, 0, 1), (0, 1, 1), (1, 1, 1)):
                    if i + di < nx and j + dj < ny and k + dk < nz:
                        s.addJoint(SceneJoint(ids[i, j, k], ids[i + di, j + dj, k + dk], (0, 0, 0), (0, 0, 0),
                                              stiffnessLin=0.0, stiffnessAng=0.0))
    return [ids[i, j, k] for i in range(nx) for j in range(ny) for k in range(nz)]


def particle_rod_chain():
    """testParticleRodChainInextensible (SoftBodyTests.swift:28-51): twelve
    particles joined by hard rods, the first pinned to the world."""
    from avbd_amd.scene import SceneJoint
    s = PhysicsScene("pc", SimSettings(iterations=20))
    prev = -1
    for k in range(12):
        p = s.addParticle(0.05, 0.02, (k * 0.25, 0, 4))
        if k == 0:
            s.addJoint(SceneJoint(-1, p, (0, 0, 4), (0, 0, 0)))
        else:
            s.addSpring(prev, p, (0, 0, 0), (0, 0, 0), 1e5, hard=True)
        prev = p
    return s


def pinned_tet():
    """testTetVolumePreservation (SoftBodyTests.swift:53-68): one Neo-Hookean
    tet of edge 0.5 hanging from a pinned corner."""
    from avbd_amd.scene import SceneJoint
    s = PhysicsScene("tet1", SimSettings(iterations=20))
    ids = [s.addParticle(0.05, 0.1, p) for p in ((0, 0, 3), (0.5, 0, 3), (0, 0.5, 3), (0, 0, 3.5))]
    s.addJoint(SceneJoint(-1, ids[0], (0, 0, 3), (0, 0, 0)))
    s.addTet(tuple(ids), 2000.0, 20000.0)
    return s, ids


def ball_on_soft_block():
    """testRigidRestsOnSoftBlock (SoftBodyTests.swift:70-83)."""
    s = PhysicsScene("ballonblock", SimSettings(iterations=20, betaLin=20000.0, lambdaMax=800.0))
    _ground(s)
    add_soft_lattice(s, (0, 0, 1.2), 4, 4, 4, 0.34, mu=4000.0, lam=40000.0)
    ball = s.addSphere(1.1, 3.0, 0.6, (0, 0, 4.2))
    return s, ball


def stacked_soft_blocks():
    """testSoftBlocksDontInterpenetrate (SoftBodyTests.swift:98-123)."""
    s = PhysicsScene("sbsb", SimSettings(iterations=20))
    _ground(s)
    add_soft_lattice(s, (0, 0, 0.7), 4, 4, 4, 0.34, mu=4000.0, lam=40000.0, massPerNode=0.05)
    top = add_soft_lattice(s, (0.15, 0.1, 2.3), 3, 3, 3, 0.34, mu=4000.0, lam=40000.0, massPerNode=0.05)
    return s, top


def add_cloth_grid(s, positions, thickness, massPerNode, friction=0.7, structuralK=5000.0, membraneMu=300.0,
                   membraneBend=5e-4):
    """Demos.addClothGrid (Demos7.swift:17-85) with hard structural rods and
    membranes (its defaults): particles, hard rods along both grid
    directions, two StVK / Bergou triangles per cell on alternating
    diagonals.  Returns the particle grid."""
    nu, nv = len(positions), len(positions[0])
    grid = [[s.addParticle(thickness, massPerNode, positions[i][j], friction=friction) for j in range(nv)]
            for i in range(nu)]
    for i in range(nu):
        for j in range(nv):
            if i + 1 < nu:
                s.addSpring(grid[i][j], grid[i + 1][j], (0, 0, 0), (0, 0, 0), structuralK, hard=True)
            if j + 1 < nv:
                s.addSpring(grid[i][j], grid[i][j + 1], (0, 0, 0), (0, 0, 0), structuralK, hard=True)
    for i in range(nu - 1):
        for j in range(nv - 1):
            v00, v10, v01, v11 = grid[i][j], grid[i + 1][j], grid[i][j + 1], grid[i + 1][j + 1]
            if (i + j) % 2 == 0:
                s.addTri((v00, v10, v11), membraneMu, membraneMu, membraneBend)
                s.addTri((v00, v11, v01), membraneMu, membraneMu, membraneBend)
            else:
                s.addTri((v00, v10, v01), membraneMu, membraneMu, membraneBend)
                s.addTri((v10, v11, v01), membraneMu, membraneMu, membraneBend)
    return grid


def clothcombo(res=16, friction=0.8, membraneMu=300.0, bend=5e-4, blockMu=4000.0):
    """Demos.clothcombo (Demos7.swift:520-565): a 7^3 soft block, a membrane
    cloth dropped over it and a box and a sphere dropped on top -- tets,
    cloth and rigid bodies in one solve."""
    s = PhysicsScene("clothcombo", SimSettings(iterations=20, betaLin=20000.0, particleDamping=1.5,
                                               clothViscosity=0.25, lambdaMax=1.0e6))
    _ground(s, friction)
    add_soft_lattice(s, (0, 0, 0.7), 7, 7, 7, 0.17, mu=blockMu, lam=10 * blockMu, massPerNode=0.0095)
    n = max(16, res)
    size = 2.6
    h = size / (n - 1)
    r = min(0.05, 0.3 * h)
    pos = [[(i * h - size / 2, j * h - size / 2, 1.6) for j in range(n)] for i in range(n)]
    add_cloth_grid(s, pos, r, 0.01, friction=friction, membraneMu=membraneMu, membraneBend=bend)
    s.addBody((0.45, 0.45, 0.45), 2.0, 0.7, (-0.1, 0.1, 2.6))
    s.addSphere(0.5, 1.5, 0.5, (0.5, -0.4, 3.1))
    return s


def boxoncloth(res=20, massRatio=8.0, friction=0.8, membraneMu=300.0, bend=5e-4):
    """Demos.boxoncloth (Demos7.swift:152-194): a cloth laid over a static
    pedestal and a box of `massRatio` times the cloth's mass dropped on it."""
    s = PhysicsScene("boxoncloth", SimSettings(iterations=20, betaLin=20000.0, particleDamping=1.5,
                                               clothViscosity=0.25))
    _ground(s, friction)
    s.addBody((1.2, 1.2, 0.8), 0.0, 0.6, (0, 0, 0.4))
    n = max(16, res)
    size = 2.6
    h = size / (n - 1)
    r = min(0.05, 0.3 * h)
    pos = [[(i * h - size / 2, j * h - size / 2, 0.8 + r + 0.02) for j in range(n)] for i in range(n)]
    add_cloth_grid(s, pos, r, 0.01, friction=friction, membraneMu=membraneMu, membraneBend=bend)
    side = 0.55
    s.addBody((side, side, side), massRatio * n * n * 0.01 / side ** 3, 0.7, (0, 0, 1.45))
    return s


def hammock(res=16, friction=0.8, membraneMu=300.0, bend=5e-4):
    """Demos.hammock (Demos7.swift:200-262) with one cargo box: a sagging
    cloth pinned to the world along both short ends."""
    from avbd_amd.demos import SplitMix64
    from avbd_amd.scene import SceneJoint
    s = PhysicsScene("hammock", SimSettings(iterations=20, betaLin=20000.0, particleDamping=1.5,
                                            clothViscosity=0.25))
    _ground(s, friction)
    nu, nv = max(16, res), max(10, res * 2 // 3)
    h = 0.14
    r = min(0.05, 0.3 * h)
    span = (nu - 1) * h
    sag = 0.22 * span
    spanZ = sag + 1.1
    pos = []
    for i in range(nu):
        t = i / (nu - 1) * 2 - 1
        dip = sag * (1 - t * t)
        pos.append([(i * h - span / 2, j * h - (nv - 1) * h / 2, spanZ - dip) for j in range(nv)])
    grid = add_cloth_grid(s, pos, r, 0.012, friction=friction, membraneMu=membraneMu, membraneBend=bend)
    for j in range(nv):
        for i in (0, nu - 1):
            node = grid[i][j]
            s.addJoint(SceneJoint(-1, node, s.bodies[node]["position"], (0, 0, 0)))
    rng = SplitMix64(7)
    jx = (rng.nextFloat() - 0.5) * 0.2
    jy = (rng.nextFloat() - 0.5) * 0.3
    angle = rng.nextFloat() * 0.5
    ax = (rng.nextFloat(), rng.nextFloat(), 1.0)
    s.addBody((0.5, 0.5, 0.5), 8.0, 0.6, (jx, jy, spanZ + 0.6), rotation=_quat(angle, ax))
    return s


def clothfold(res=36, friction=0.9, membraneMu=300.0, bend=5e-4):
    """Demos.clothfold (Demos7.swift:90-145): a 2.2 m towel laid in an S
    fold (two half-turns of radius rb) just above the ground."""
    s = PhysicsScene("clothfold", SimSettings(iterations=10, betaLin=20000.0, particleDamping=1.5,
                                              clothViscosity=0.25))
    _ground(s, friction)
    nu, nv = max(36, res), max(18, res // 2)
    total = 2.2
    h = total / (nu - 1)
    r = min(0.045, 0.36 * h)
    rb = max(0.055, 2.6 * r)
    straight = (total - 2 * math.pi * rb) / 3

    def profile(a):
        a1 = straight
        b1 = a1 + math.pi * rb
        a2 = b1 + straight
        b2 = a2 + math.pi * rb
        if a < a1:
            return a, 0.0
        if a < b1:
            t = (a - a1) / rb
            return a1 + rb * math.sin(t), rb - rb * math.cos(t)
        if a < a2:
            return a1 - (a - b1), 2 * rb
        if a < b2:
            t = (a - a2) / rb
            return a1 - straight - rb * math.sin(t), 2 * rb + rb * (1 - math.cos(t))
        return a1 - straight + (a - b2), 4 * rb

    pos = []
    for i in range(nu):
        x, z = profile(i * h)
        pos.append([(x - straight / 2, j * h - (nv - 1) * h / 2, z + r + 0.02) for j in range(nv)])
    add_cloth_grid(s, pos, r, 0.5 * h * h, friction=friction, membraneMu=membraneMu, membraneBend=bend)
    return s


def flagwhip(res=14, kick=3.0, membraneMu=300.0, bend=5e-4):
    """Demos.flagwhip (Demos7.swift:307-350): a hanging sheet pinned at its
    two top corners, kicked sideways with a velocity growing down the sheet."""
    from avbd_amd.scene import SceneJoint
    s = PhysicsScene("flagwhip", SimSettings(iterations=10, betaLin=20000.0, particleDamping=1.5,
                                             clothViscosity=0.25))
    _ground(s, 0.6)
    n = max(12, res)
    h = 2.0 / (n - 1)
    r = min(0.05, 0.3 * h)
    pos = [[(0.0, i * h - 1.0, 3.2 - j * h) for j in range(n)] for i in range(n)]
    grid = add_cloth_grid(s, pos, r, 0.008, friction=0.6, membraneMu=membraneMu, membraneBend=bend)
    for i in (0, n - 1):
        node = grid[i][0]
        s.addJoint(SceneJoint(-1, node, s.bodies[node]["position"], (0, 0, 0)))
    for i in range(n):
        for j in range(n):
            s.bodies[grid[i][j]]["velocity"] = (kick * j / (n - 1), 0.0, 0.0)
    return s


def particle_ke(scene, v):
    """Kinetic energy of the scene's particles (mass from the density the
    particle was given, as Scene.addParticle derives it)."""
    ke = 0.0
    for b, body in enumerate(scene.bodies):
        if body["particle"]:
            m = body["density"] * 4.0 / 3.0 * math.pi * (body["size"][0] / 2) ** 3
            ke += 0.5 * m * float(v[b][0] ** 2 + v[b][1] ** 2 + v[b][2] ** 2)
    return ke


def eecross(length=18, wid=4):
    """Demos.eecross (Demos7.swift:456-516): strip A pinned taut just above
    the ground, strip B across it held by four soft drag joints on its
    leading edge.  Returns (scene, drag joints, A nodes, B nodes)."""
    from avbd_amd.scene import SceneJoint
    s = PhysicsScene("eecross", SimSettings(iterations=20, betaLin=20000.0, particleDamping=1.5,
                                            clothViscosity=0.25))
    _ground(s, 0.3)
    h, r, zA = 0.12, 0.04, 0.12
    posA = [[(i * h - (length - 1) * h / 2, j * h - (wid - 1) * h / 2, zA) for j in range(wid)] for i in range(length)]
    gridA = add_cloth_grid(s, posA, r, 0.01, friction=0.3)
    for i in (0, length - 1):
        for j in range(wid):
            node = gridA[i][j]
            s.addJoint(SceneJoint(-1, node, s.bodies[node]["position"], (0, 0, 0)))
    posB = [[(j * h - (wid - 1) * h / 2, i * h - (length - 1) * h / 2, zA + 2 * r + 0.02) for j in range(wid)]
            for i in range(length)]
    gridB = add_cloth_grid(s, posB, r, 0.01, friction=0.3)
    drag = []
    for j in range(wid):
        node = gridB[length - 1][j]
        drag.append(len(s.joints))
        s.addJoint(SceneJoint(-1, node, s.bodies[node]["position"], (0, 0, 0), stiffnessLin=3000.0, stiffnessAng=0.0))
    return s, drag, [n for row in gridA for n in row], [n for row in gridB for n in row]


def add_cloth(s, origin, dirU, dirV, nu, nv, spacing, thickness=0.06, massPerNode=0.02, friction=0.7):
    """Demos.addCloth (Demos5.swift:15-74): a particle sheet spanned by
    dirU / dirV with StVK + Bergou triangles (mu = lambda = 300, bend 5e-4)
    on alternating diagonals and hard rods (cap 5000) along the grid; the
    skin radius stays under 0.38 spacing."""
    t = min(thickness, 0.38 * spacing)
    grid = [[s.addParticle(t, massPerNode, tuple(origin[k] + dirU[k] * i * spacing + dirV[k] * j * spacing
                                                 for k in range(3)), friction=friction)
             for j in range(nv)] for i in range(nu)]
    for i in range(nu - 1):
        for j in range(nv - 1):
            v00, v10, v01, v11 = grid[i][j], grid[i + 1][j], grid[i][j + 1], grid[i + 1][j + 1]
            if (i + j) % 2 == 0:
                s.addTri((v00, v10, v11), 300.0, 300.0, 5e-4)
                s.addTri((v00, v11, v01), 300.0, 300.0, 5e-4)
            else:
                s.addTri((v00, v10, v01), 300.0, 300.0, 5e-4)
                s.addTri((v10, v11, v01), 300.0, 300.0, 5e-4)
    for i in range(nu):
        for j in range(nv):
            if i + 1 < nu:
                s.addSpring(grid[i][j], grid[i + 1][j], (0, 0, 0), (0, 0, 0), 5000.0, hard=True)
            if j + 1 < nv:
                s.addSpring(grid[i][j], grid[i][j + 1], (0, 0, 0), (0, 0, 0), 5000.0, hard=True)
    return grid


def flag_on_pole(res=14):
    """Demos.cloth(res:ball: false) (Demos5.swift:243-305): a flag hard-attached
    at two points to a deformable pole of ten ball-jointed rigid segments
    with finite angular stiffness.  Returns (scene, top attachment particle)."""
    from avbd_amd.scene import SceneJoint
    s = PhysicsScene("cloth", SimSettings(iterations=20, betaLin=20000.0, particleDamping=1.5, clothViscosity=0.25,
                                          betaAng=20000.0, lambdaMax=500.0))
    _ground(s, 0.8)
    segs, segH = 10, 0.55
    prev = top = -1
    for k in range(segs):
        seg = s.addBody((0.16, 0.16, segH * 0.98), 40.0, 0.5, (0, 0, (k + 0.5) * segH))
        if k == 0:
            s.addJoint(SceneJoint(-1, seg, (0, 0, 0), (0, 0, -segH / 2), stiffnessLin=math.inf, stiffnessAng=math.inf))
        else:
            s.addJoint(SceneJoint(prev, seg, (0, 0, segH / 2), (0, 0, -segH / 2), stiffnessLin=math.inf,
                                  stiffnessAng=80000.0))
            s.addJoint(SceneJoint(prev, seg, (0, 0, 0), (0, 0, 0), stiffnessLin=0.0, stiffnessAng=0.0))
        prev = top = seg
    n = max(10, res)
    spacing = 2.6 / (n - 1)
    u = (0.55, 0.2, -0.81)
    un = math.sqrt(sum(c * c for c in u))
    grid = add_cloth(s, (0.12, 0, segs * segH - 0.15), tuple(c / un for c in u), (0, 0, -1), n, n, spacing,
                     massPerNode=0.007)
    s.addJoint(SceneJoint(top, grid[0][0], (0.12, 0, segH / 2 - 0.15), (0, 0, 0)))
    low = top - int(2.6 / segH)
    s.addJoint(SceneJoint(low, grid[0][n - 1], (0.12, 0, 0), (0, 0, 0)))
    for row in grid:
        for node in row:
            s.bodies[node]["velocity"] = (0.8, 0.5, 0.0)
    return s, grid[0][0]


def legacy_rate_motor(dt=0.25, rate=2.0):
    """testLegacyRateMotorAdvancesWrappedServoTarget (GPUSolverTests.swift:148-168)."""
    from avbd_amd.scene import SceneJoint
    s = PhysicsScene("legacy-rate-motor", SimSettings(dt=dt, gravity=0.0))
    b = s.addBody((0.2, 0.2, 0.2), 1.0, 0.0, (0, 0, 0))
    s.addJoint(SceneJoint(-1, b, (0, 0, 0), (0, 0, 0), stiffnessLin=math.inf, stiffnessAng=math.inf,
                          hingeAxis=(0, 0, 1), motorTorque=10.0, motorStiffness=400.0, motorRate=rate))
    return s


def sheets_and_block(n=7, h=0.06):
    """Two n x n cloth sheets, the upper turned a quarter and dropped on the
    lower, and a small tet block beside them: the tets select the isotropic
    limiter (GPUSolver.swift:93-99), the sheets give it authored V-T and E-E
    pairs (a test scene of this backend, built from the Demos helpers)."""
    s = PhysicsScene("sheets+block", SimSettings(iterations=8))
    _ground(s)
    lower = [[(x * h - 0.18, y * h - 0.18, 0.3) for y in range(n)] for x in range(n)]
    upper = [[(y * h - 0.17, x * h - 0.19, 0.36) for y in range(n)] for x in range(n)]
    add_cloth_grid(s, lower, 0.008, 0.004)
    add_cloth_grid(s, upper, 0.008, 0.004)
    add_soft_lattice(s, (1.0, 0, 0.4), 2, 2, 2, 0.2)
    return s
