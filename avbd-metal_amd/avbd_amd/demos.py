"""Scene generators for the benchmark workloads (conventions of
Sources/PhysicsAVBD/Demos/Demos.swift and Demos5.swift, own builders).

All randomness is SplitMix64 (Sources/SimCore/Math/Random.swift:3-21);
Gaussian samples use Box-Muller on nextFloat with u1 = 1 - nextFloat so the
logarithm never sees zero.
"""
import math

import numpy as np

from .scene import BOX, SPHERE, PhysicsScene, SceneJoint, SimSettings

_M64 = (1 << 64) - 1


class SplitMix64:
    """Random.swift:3-21."""

    def __init__(self, seed):
        self.state = seed & _M64

    def next(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & _M64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
        return z ^ (z >> 31)

    def nextFloat(self):
        return np.float32(self.next() >> 40) / np.float32(1 << 24)

    def uniform(self, lo, hi):
        return float(lo + (hi - lo) * float(self.nextFloat()))

    def gauss(self, sigma=1.0):
        u1 = 1.0 - float(self.nextFloat())
        u2 = float(self.nextFloat())
        return sigma * math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)


def add_ground(s, friction=0.7):
    """Demos.swift:191-195: a 200 x 200 x 2 static box whose top is z = 0."""
    return s.addBody((200, 200, 2), 0.0, friction, (0, 0, -1))


def stack_scene(height=10, seed=1, friction=0.5, iterations=10, jitter=0.01):
    """Rigid stack: unit cubes (density 1 -> mass 1) at z = 0.5 + i with seeded
    xy jitter on the static ground (Demos.swift:214-222 layout)."""
    s = PhysicsScene("stack", SimSettings(iterations=iterations))
    add_ground(s, friction)
    rng = SplitMix64(seed)
    for i in range(height):
        x = rng.uniform(-jitter, jitter)
        y = rng.uniform(-jitter, jitter)
        s.addBody((1, 1, 1), 1.0, friction, (x, y, 0.5 + i))
    return s


def kuhn_tets(nx, ny, nz, index):
    """6-tet Kuhn split of every lattice cell: the tets of one cell share its
    main diagonal c000-c111, one per axis permutation."""
    perms = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]
    out = []
    for i in range(nx - 1):
        for j in range(ny - 1):
            for k in range(nz - 1):
                for p in perms:
                    c = [i, j, k]
                    v = [index(*c)]
                    for a in p[:2]:
                        c[a] += 1
                        v.append(index(*c))
                    v.append(index(i + 1, j + 1, k + 1))
                    out.append(tuple(v))
    return out


def soft_block_scene(cells=20, edge=1.0, E=1.0e5, nu=0.4, density=1000.0, bottom=0.5, seed=2,
                     vel_sigma=0.1, iterations=20, ground_friction=0.7, particle_friction=0.5,
                     radius_factor=0.32, origin=(0.0, 0.0)):
    """Soft block: (cells+1)^3 particles on a lattice of spacing edge/cells,
    Kuhn 6-tet split, Stable Neo-Hookean with Lame parameters from (E, nu),
    lumped mass rho V / 4 per tet vertex, seeded initial velocity N(0, sigma)
    per component, dropped onto the ground.  Particles are ordered as
    addSoftBlock does (Demos5.swift:79-107): index = (i*ny + j)*nz + k."""
    n = cells + 1
    h = edge / cells
    mu = E / (2.0 * (1.0 + nu))
    lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    s = PhysicsScene("soft_block", SimSettings(iterations=iterations))
    add_ground(s, ground_friction)
    base = len(s.bodies)

    def index(i, j, k):
        return base + (i * n + j) * n + k

    tets = kuhn_tets(n, n, n, index)
    mass = np.zeros(n * n * n)
    tet_vol = h ** 3 / 6.0
    for t in tets:
        for v in t:
            mass[v - base] += density * tet_vol / 4.0
    rng = SplitMix64(seed)
    x0 = origin[0] - edge / 2
    y0 = origin[1] - edge / 2
    for i in range(n):
        for j in range(n):
            for k in range(n):
                vel = (rng.gauss(vel_sigma), rng.gauss(vel_sigma), rng.gauss(vel_sigma))
                p = (x0 + i * h, y0 + j * h, bottom + k * h)
                s.addParticle(radius_factor * h, float(mass[(i * n + j) * n + k]), p, friction=particle_friction,
                              velocity=vel)
    for t in tets:
        s.addTet(t, mu, lam)
    return s


def chain_scene(links=8, seed=5, iterations=10, torque_sigma=5.0):
    """Articulated chain of revolute-jointed boxes, the first hinged to the
    world, random unit axes and link lengths U(0.1, 0.4)."""
    rng = SplitMix64(seed)
    s = PhysicsScene("chain", SimSettings(iterations=iterations))
    add_ground(s)
    z = 3.0
    prev = -1
    top = (0.0, 0.0, z)
    for i in range(links):
        L = rng.uniform(0.1, 0.4)
        c = (top[0], top[1], top[2] - L / 2)
        b = s.addBody((0.1, 0.1, L), 1.0, 0.5, c)
        ax = np.array([rng.gauss(), rng.gauss(), rng.gauss()])
        ax /= max(1e-9, np.linalg.norm(ax))
        rA = (0.0, 0.0, -s.bodies[prev]["size"][2] / 2) if prev >= 0 else top
        s.addJoint(SceneJoint(prev, b, rA, (0.0, 0.0, L / 2), stiffnessLin=math.inf, stiffnessAng=math.inf,
                              hingeAxis=tuple(float(a) for a in ax)))
        prev = b
        top = (c[0], c[1], c[2] - L / 2)
    return s


def batched_scene(num_scenes=4096, pitch=3.0, iterations=10, stack_range=(4, 64), chain_range=(6, 19),
                  torque_sigma=5.0, seed_base=4 << 32):
    """Batched uneven rigid scenes in one solver (replica layout of
    Sources/RL/PushTEnv.swift:115-180): a grid of independent scenes, scene
    s in collision group s + 1 over one shared group-0 ground sized to the
    grid.  Scene s draws from SplitMix64(seed_base + s) either
      * a stack of n ~ U{4..64} cubes of one edge U(0.2, 1) m, density 1, or
      * a chain of n ~ U{6..19} hinge-jointed boxes (link length
        U(0.1, 0.4) m), the first hinged to the world, random unit axes,
        each hinge driven by a constant torque N(0, 5) N m: a velocity motor
        whose effort saturates at |tau| (target +-1000 rad/s, kd 1e6).
    Returns (scene, per-scene body ranges)."""
    cols = int(math.ceil(math.sqrt(num_scenes)))
    rows = int(math.ceil(num_scenes / cols))
    hx = 0.5 * (cols - 1) * pitch
    hy = 0.5 * (rows - 1) * pitch
    s = PhysicsScene("batched", SimSettings(iterations=iterations))
    s.addBody((2 * (hx + 3.5), 2 * (hy + 3.5), 2), 0.0, 0.7, (0, 0, -1))
    ranges = []
    for e in range(num_scenes):
        rng = SplitMix64(seed_base + e)
        cx = (e % cols) * pitch - hx
        cy = (e // cols) * pitch - hy
        grp = e + 1
        first = len(s.bodies)
        if rng.nextFloat() < 0.5:
            n = stack_range[0] + int(rng.next() % (stack_range[1] - stack_range[0] + 1))
            edge = rng.uniform(0.2, 1.0)
            for i in range(n):
                jx, jy = rng.uniform(-0.01, 0.01), rng.uniform(-0.01, 0.01)
                s.addBody((edge, edge, edge), 1.0, 0.5, (cx + jx, cy + jy, edge * (0.5 + i)), collisionGroup=grp)
        else:
            n = chain_range[0] + int(rng.next() % (chain_range[1] - chain_range[0] + 1))
            top = (cx, cy, 0.3 + 0.4 * n)
            prev = -1
            prevL = 0.0
            for i in range(n):
                L = rng.uniform(0.1, 0.4)
                c = (top[0], top[1], top[2] - L / 2)
                b = s.addBody((0.08, 0.08, L), 1.0, 0.5, c, collisionGroup=grp)
                ax = np.array([rng.gauss(), rng.gauss(), rng.gauss()])
                ax /= max(1e-9, float(np.linalg.norm(ax)))
                tau = rng.gauss(torque_sigma)
                rA = (0.0, 0.0, -prevL / 2) if prev >= 0 else top
                s.addJoint(SceneJoint(prev, b, rA, (0.0, 0.0, L / 2), stiffnessLin=math.inf,
                                      stiffnessAng=math.inf, hingeAxis=tuple(float(a) for a in ax),
                                      motorTarget=1000.0 if tau >= 0 else -1000.0, motorTorque=abs(tau),
                                      motorStiffness=0.0, motorDamping=1.0e6, motorMode=2))
                prev, prevL = b, L
                top = (c[0], c[1], c[2] - L / 2)
        ranges.append((first, len(s.bodies)))
    return s, ranges
