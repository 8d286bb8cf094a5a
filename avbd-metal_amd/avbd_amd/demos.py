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
    s = PhysicsScene("soft_block", SimSettings(iterations=iterations))
    add_ground(s, ground_friction)
    add_soft_block(s, cells, edge, E, nu, density, bottom, seed, vel_sigma, particle_friction, radius_factor, origin)
    return s


def add_soft_block(s, cells, edge=1.0, E=1.0e5, nu=0.4, density=1000.0, bottom=0.5, seed=2, vel_sigma=0.1,
                   particle_friction=0.5, radius_factor=0.32, origin=(0.0, 0.0)):
    """Adds one soft block (see soft_block_scene) centred at `origin` in xy;
    returns its body range."""
    n = cells + 1
    h = edge / cells
    mu = E / (2.0 * (1.0 + nu))
    lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
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
    return base, len(s.bodies)


def batched_soft_specs(num_scenes=64, cells_range=(3, 8), seed_base=5 << 32):
    """Per-scene draws of the batched soft blocks: scene e takes from
    SplitMix64(seed_base + e) cells ~ U{lo..hi}, edge U(0.3, 1) m, Young's
    modulus 10^U(4, 5) Pa, drop height U(0.1, 0.6) m and its velocity seed."""
    specs = []
    for e in range(num_scenes):
        rng = SplitMix64(seed_base + e)
        cells = cells_range[0] + int(rng.next() % (cells_range[1] - cells_range[0] + 1))
        edge = rng.uniform(0.3, 1.0)
        E = 10.0 ** rng.uniform(4.0, 5.0)
        bottom = rng.uniform(0.1, 0.6)
        specs.append(dict(cells=cells, edge=edge, E=E, bottom=bottom, seed=int(rng.next() & 0xFFFFFFFF)))
    return specs


def batched_soft_scene(num_scenes=64, cells_range=(3, 8), pitch=2.0, iterations=20, seed_base=5 << 32,
                       separate=False):
    """Uneven soft blocks (batched_soft_specs) on a grid of `pitch` m over
    one shared ground: one scene holding every block, or with separate=True
    one scene per block (each on its own ground) with the same bodies --
    the same scenes stepped one solver at a time.  Returns the scene (or the
    list of scenes)."""
    specs = batched_soft_specs(num_scenes, cells_range, seed_base)
    cols = int(math.ceil(math.sqrt(num_scenes)))
    rows = int(math.ceil(num_scenes / cols))
    hx = 0.5 * (cols - 1) * pitch
    hy = 0.5 * (rows - 1) * pitch
    if separate:
        out = []
        for e, sp in enumerate(specs):
            s = PhysicsScene(f"soft_{e}", SimSettings(iterations=iterations))
            cx, cy = (e % cols) * pitch - hx, (e // cols) * pitch - hy
            s.addBody((pitch, pitch, 2), 0.0, 0.7, (cx, cy, -1))
            add_soft_block(s, sp["cells"], sp["edge"], sp["E"], bottom=sp["bottom"], seed=sp["seed"], origin=(cx, cy))
            out.append(s)
        return out
    s = PhysicsScene("batched_soft", SimSettings(iterations=iterations))
    s.addBody((2 * hx + pitch, 2 * hy + pitch, 2), 0.0, 0.7, (0, 0, -1))
    for e, sp in enumerate(specs):
        cx, cy = (e % cols) * pitch - hx, (e // cols) * pitch - hy
        add_soft_block(s, sp["cells"], sp["edge"], sp["E"], bottom=sp["bottom"], seed=sp["seed"], origin=(cx, cy))
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


def cloth_scene(n=256, size=2.0, height=1.0, jitter=1.0e-3, seed=3, radius=2.5e-3, areal_density=0.5,
                stretch_k=1.0e4, hard_rods=True, membrane_mu=300.0, bend=1.0e-3, sphere_radius=0.5, sphere_z=0.45,
                iterations=30, friction=0.6, particle_damping=1.5, viscosity=0.25, ground=True):
    """Cloth sheet dropped onto a fixed sphere (grid layout of
    Demos7.swift:16-80 addClothGrid, drape settings of Demos7.swift:265-300):
    an n x n particle grid of side `size` at z = height + U(-jitter, jitter)
    (SplitMix64(seed), row-major); structural grid edges as addClothGrid
    builds them -- hard augmented-Lagrangian rods whose penalty is capped at
    `stretch_k` (soft springs of stiffness `stretch_k` with hard_rods=False);
    two triangles per quad with alternating diagonals carrying StVK
    membranes (mu = lambda = membrane_mu) and Bergou bending `bend`; per-node
    mass areal_density x spacing^2; particle radius `radius` (contact radius
    2 r).  The static sphere is centred at (0, 0, sphere_z) just below the
    sheet; the static ground's top touches the sphere's bottom."""
    s = PhysicsScene("cloth", SimSettings(iterations=iterations, particleDamping=particle_damping,
                                          clothViscosity=viscosity))
    if ground:
        s.addBody((200, 200, 2), 0.0, friction, (0, 0, sphere_z - sphere_radius - 1.0))
    s.addSphere(2.0 * sphere_radius, 0.0, friction, (0.0, 0.0, sphere_z))
    rng = SplitMix64(seed)
    h = size / (n - 1)
    mass = areal_density * h * h
    grid = np.zeros((n, n), np.int64)
    for i in range(n):
        for j in range(n):
            z = height + rng.uniform(-jitter, jitter)
            grid[i, j] = s.addParticle(radius, mass, (i * h - size / 2, j * h - size / 2, z), friction=friction)
    for i in range(n):
        for j in range(n):
            if i + 1 < n:
                s.addSpring(int(grid[i, j]), int(grid[i + 1, j]), (0, 0, 0), (0, 0, 0), stretch_k, hard=hard_rods)
            if j + 1 < n:
                s.addSpring(int(grid[i, j]), int(grid[i, j + 1]), (0, 0, 0), (0, 0, 0), stretch_k, hard=hard_rods)
    for i in range(n - 1):
        for j in range(n - 1):
            v00, v10, v01, v11 = (int(grid[i, j]), int(grid[i + 1, j]), int(grid[i, j + 1]), int(grid[i + 1, j + 1]))
            if (i + j) % 2 == 0:
                s.addTri((v00, v10, v11), membrane_mu, membrane_mu, bend)
                s.addTri((v00, v11, v01), membrane_mu, membrane_mu, bend)
            else:
                s.addTri((v00, v10, v01), membrane_mu, membrane_mu, bend)
                s.addTri((v10, v11, v01), membrane_mu, membrane_mu, bend)
    return s


def drape_scene(res=24, friction=0.6, membrane_mu=300.0, bend=5.0e-4, structural_k=5000.0, hard_rods=True):
    """Demos.drape (Demos7.swift:265-300) built with addClothGrid defaults
    (Demos7.swift:16-80): an n x n sheet (n = max(20, res)) of side 2.8 m at
    z = 1.75 over a static sphere of diameter 1.1 at z = 1, structural edges as
    hard AL rods capped at `structural_k`, StVK membranes and Bergou bending on
    alternating-diagonal triangles, 8 g nodes, r = min(0.05, 0.3 spacing),
    10 iterations, betaLin 2e4, particle damping 1.5, cloth viscosity 0.25."""
    s = PhysicsScene("drape", SimSettings(iterations=10, betaLin=20000.0, particleDamping=1.5, clothViscosity=0.25))
    add_ground(s, friction)
    s.addSphere(1.1, 0.0, friction, (0.0, 0.0, 1.0))
    n = max(20, res)
    size = 2.8
    h = size / (n - 1)
    r = min(0.05, 0.3 * h)
    grid = [[s.addParticle(r, 0.008, (i * h - size / 2, j * h - size / 2, 1.75), friction=friction)
             for j in range(n)] for i in range(n)]
    for i in range(n):
        for j in range(n):
            if i + 1 < n:
                s.addSpring(grid[i][j], grid[i + 1][j], (0, 0, 0), (0, 0, 0), structural_k, hard=hard_rods)
            if j + 1 < n:
                s.addSpring(grid[i][j], grid[i][j + 1], (0, 0, 0), (0, 0, 0), structural_k, hard=hard_rods)
    for i in range(n - 1):
        for j in range(n - 1):
            v00, v10, v01, v11 = grid[i][j], grid[i + 1][j], grid[i][j + 1], grid[i + 1][j + 1]
            if (i + j) % 2 == 0:
                s.addTri((v00, v10, v11), membrane_mu, membrane_mu, bend)
                s.addTri((v00, v11, v01), membrane_mu, membrane_mu, bend)
            else:
                s.addTri((v00, v10, v01), membrane_mu, membrane_mu, bend)
                s.addTri((v10, v11, v01), membrane_mu, membrane_mu, bend)
    return s


def cloth_metrics(scene, positions, tris=None):
    """debugClothMetrics (GPUSolver.swift:2687-2768): the minimum surface gap
    |p - closest(p, T)| - (r_p + r_T) over particles and non-adjacent
    triangles, and the largest relative elongation of stiff springs
    (stiffness >= 1000 or hard).  `tris`: the collision triangles (a
    solver's collisionTris(), tet boundary faces included); the scene's
    authored triangles when omitted."""
    x = np.asarray(positions, np.float64)
    rad = np.array([b["size"][0] / 2 for b in scene.bodies])
    tris = np.array([t[0] for t in scene.tris] if tris is None else tris, np.int64)
    nbr = {}
    for t in tris:
        for a in t:
            for b in t:
                if a != b:
                    nbr.setdefault(int(a), set()).add(int(b))
    parts = sorted({int(v) for t in tris for v in t if scene.bodies[int(v)]["particle"]})
    A, B, Cc = x[tris[:, 0]], x[tris[:, 1]], x[tris[:, 2]]
    rT = rad[tris].max(axis=1)
    gap = np.inf
    for v in parts:
        p = x[v]
        skip = np.isin(tris, list(nbr.get(v, set()) | {v})).any(axis=1)
        q = _closest_points(p, A, B, Cc)
        d = np.linalg.norm(p - q, axis=1) - (rad[v] + rT)
        d[skip] = np.inf
        gap = min(gap, float(d.min()))
    stretch = 0.0
    for sp in scene.springs:
        if not (sp["k"] >= 1000 or sp["hard"]):
            continue
        a, b = sp["a"], sp["b"]
        rest = sp["rest"] if sp["rest"] >= 0 else np.linalg.norm(np.subtract(scene.bodies[a]["position"],
                                                                             scene.bodies[b]["position"]))
        if rest > 1e-6:
            stretch = max(stretch, max(np.linalg.norm(x[a] - x[b]) / rest - 1.0, 0.0))
    return (0.0 if gap == np.inf else gap), stretch


def _closest_points(p, a, b, c):
    """Closest point of p on each triangle (rows of a, b, c), Ericson 5.1.5, vectorised."""
    ab, ac, ap = b - a, c - a, p - a
    d1, d2 = np.einsum("ij,ij->i", ab, ap), np.einsum("ij,ij->i", ac, ap)
    bp = p - b
    d3, d4 = np.einsum("ij,ij->i", ab, bp), np.einsum("ij,ij->i", ac, bp)
    cp = p - c
    d5, d6 = np.einsum("ij,ij->i", ab, cp), np.einsum("ij,ij->i", ac, cp)
    vc, vb, va = d1 * d4 - d3 * d2, d5 * d2 - d1 * d6, d3 * d6 - d5 * d4
    den = va + vb + vc
    with np.errstate(divide="ignore", invalid="ignore"):
        v = vb / den
        w = vc / den
        out = a + ab * v[:, None] + ac * w[:, None]
        e_bc = (d4 - d3) / ((d4 - d3) + (d5 - d6))
        out = np.where(((va <= 0) & ((d4 - d3) >= 0) & ((d5 - d6) >= 0))[:, None], b + (c - b) * e_bc[:, None], out)
        e_ac = d2 / (d2 - d6)
        out = np.where(((vb <= 0) & (d2 >= 0) & (d6 <= 0))[:, None], a + ac * e_ac[:, None], out)
        out = np.where(((d6 >= 0) & (d5 <= d6))[:, None], c, out)
        e_ab = d1 / (d1 - d3)
        out = np.where(((vc <= 0) & (d1 >= 0) & (d3 <= 0))[:, None], a + ab * e_ab[:, None], out)
        out = np.where(((d3 >= 0) & (d4 <= d3))[:, None], b, out)
        out = np.where(((d1 <= 0) & (d2 <= 0))[:, None], a, out)
    return out


def _quat(angle, axis):
    a = np.asarray(axis, np.float64)
    a = a / np.linalg.norm(a)
    h = 0.5 * angle
    return (float(a[0] * math.sin(h)), float(a[1] * math.sin(h)), float(a[2] * math.sin(h)), float(math.cos(h)))


def _qmul(a, b):
    ax, ay, az, aw = a
    bx, by, bz, bw = b
    q = np.array([aw * bx + ax * bw + ay * bz - az * by, aw * by - ax * bz + ay * bw + az * bx,
                  aw * bz + ax * by - ay * bx + az * bw, aw * bw - ax * bx - ay * by - az * bz])
    return tuple(float(v) for v in q / np.linalg.norm(q))


def chainmail_scene(rings=6, drops=3):
    """Mail of interlocked tori hung on flexible posts (Demos2.swift:318-430
    layout): an n x n grid of flat rings, joined by upright rings along x and
    y between neighbours; posts at the corner and mid-edge cells carry welded
    cross pegs just under the sheet; seeded spheres and cubes drop on top."""
    s = PhysicsScene("chainmail", SimSettings(iterations=15, betaLin=20000.0, lambdaMax=1500.0))
    add_ground(s, 0.6)
    R, r, sheetZ, pitch = 0.45, 0.15, 8.0, 1.15
    n = rings
    extent = (n - 1) * pitch / 2
    mid = (n - 1) // 2

    def add_post(x, y):
        segH, cross = 1.6, 0.22
        pegZ, topZ = sheetZ - 0.3, sheetZ + 1.0
        prev = s.addBody((cross, cross, segH), 0.0, 0.4, (x, y, segH / 2))
        z = segH
        while z < topZ:
            h = min(segH, topZ - z)
            seg = s.addBody((cross, cross, h), 4.0, 0.4, (x, y, z + h / 2))
            s.addJoint(SceneJoint(prev, seg, (0, 0, s.bodies[prev]["size"][2] / 2), (0, 0, -h / 2),
                                  stiffnessLin=math.inf, stiffnessAng=math.inf))
            prev = seg
            z += h
        host = prev
        for axis in range(2):
            size = (1.1, 0.14, 0.14) if axis == 0 else (0.14, 1.1, 0.14)
            peg = s.addBody(size, 2.0, 0.5, (x, y, pegZ))
            segCenterZ = s.bodies[host]["position"][2]
            off = (0.4, 0, 0) if axis == 0 else (0, 0.4, 0)
            s.addJoint(SceneJoint(host, peg, (0, 0, pegZ - segCenterZ), (0, 0, 0),
                                  stiffnessLin=math.inf, stiffnessAng=math.inf))
            s.addJoint(SceneJoint(host, peg, (off[0], off[1], pegZ - segCenterZ), off,
                                  stiffnessLin=math.inf, stiffnessAng=math.inf))

    posts = set()
    for j in range(n):
        for i in range(n):
            corner = i in (0, n - 1) and j in (0, n - 1)
            edgeMid = (i == mid and j in (0, n - 1)) or (j == mid and i in (0, n - 1))
            if corner or edgeMid:
                posts.add(j * n + i)
    for cell in sorted(posts):
        add_post((cell % n) * pitch - extent, (cell // n) * pitch - extent)
    qX = _quat(math.pi / 2, (1, 0, 0))
    qY = _qmul(_quat(math.pi / 2, (0, 0, 1)), qX)
    for j in range(n):
        for i in range(n):
            cx, cy = i * pitch - extent, j * pitch - extent
            s.addTorus(R, r, 1.0, 0.3, (cx, cy, sheetZ))
            if i + 1 < n:
                s.addTorus(R, r, 1.0, 0.3, (cx + pitch / 2, cy, sheetZ), rotation=qX)
            if j + 1 < n:
                s.addTorus(R, r, 1.0, 0.3, (cx, cy + pitch / 2, sheetZ), rotation=qY)
    rng = SplitMix64(11)
    for k in range(drops):
        sz = float(np.float32(0.9) + rng.nextFloat() * np.float32(0.7))
        pos = (float((rng.nextFloat() - np.float32(0.5)) * np.float32(n * pitch * 0.4)),
               float((rng.nextFloat() - np.float32(0.5)) * np.float32(n * pitch * 0.4)), sheetZ + 4 + k * 2)
        if k % 2 == 0:
            s.addSphere(sz, 1.0, 0.5, pos)
        else:
            s.addBody((sz, sz, sz), 1.0, 0.5, pos)
    return s


def treadmill_scene(boxes=8):
    """A plank belt driven by kinematic spinning drums (Demos2.swift:492-580
    layout): 26 planks hinged in a closed loop around two wheels, each wheel
    four static octagon slabs spinning at 1.2 rad/s about y, slick support and
    return beds, and seeded cargo cubes dropped on the upstream end."""
    s = PhysicsScene("treadmill")
    add_ground(s)
    wheelX, wheelZ, wrapR, beltW = 2.6, 2.6, 1.0, 3.2
    span = 2 * wheelX
    circumference = 2 * span + 2 * math.pi * wrapR
    plankCount = 26
    plankLen = circumference / plankCount

    def path_point(t):
        t = math.fmod(t, circumference)
        if t < 0:
            t += circumference
        if t < span:
            return (-wheelX + t, 0.0, wheelZ + wrapR), 0.0
        t -= span
        if t < math.pi * wrapR:
            a = math.pi / 2 - t / wrapR
            return (wheelX + wrapR * math.cos(a), 0.0, wheelZ + wrapR * math.sin(a)), math.pi / 2 - a
        t -= math.pi * wrapR
        if t < span:
            return (wheelX - t, 0.0, wheelZ - wrapR), math.pi
        t -= span
        a = -math.pi / 2 - t / wrapR
        return (-wheelX + wrapR * math.cos(a), 0.0, wheelZ + wrapR * math.sin(a)), math.pi / 2 - a

    planks = []
    for i in range(plankCount):
        pos, pitchA = path_point((i + 0.5) * plankLen)
        planks.append(s.addBody((plankLen * 0.98, beltW, 0.1), 0.6, 0.9, pos, rotation=_quat(pitchA, (0, 1, 0))))
    for i in range(plankCount):
        nx = (i + 1) % plankCount
        for yo in (-beltW * 0.4, beltW * 0.4):
            s.addJoint(SceneJoint(planks[i], planks[nx], (plankLen / 2, yo, 0), (-plankLen / 2, yo, 0)))
    for wx in (-wheelX, wheelX):
        for k in range(4):
            idx = s.addBody((2 * (wrapR - 0.06), beltW * 0.8, 0.3), 0.0, 1.2, (wx, 0, wheelZ),
                            rotation=_quat(k * math.pi / 4, (0, 1, 0)))
            s.addSpinner(idx, (0, 1, 0), 1.2)
    s.addBody((span - 2.6, beltW, 0.25), 0.0, 0.02, (0, 0, wheelZ + wrapR - 0.26))
    s.addBody((span - 2.6, beltW, 0.25), 0.0, 0.02, (0, 0, wheelZ - wrapR - 0.45))
    rng = SplitMix64(21)
    for k in range(boxes):
        e = float(np.float32(0.5) + rng.nextFloat() * np.float32(0.25))
        x = -wheelX + 0.7 + float(rng.nextFloat() - np.float32(0.5)) * 0.6
        y = float(rng.nextFloat() - np.float32(0.5)) * (beltW - 1)
        s.addBody((e, e, e), 0.8, 0.9, (x, y, wheelZ + wrapR + 1.0 + k * 0.8))
    return s
