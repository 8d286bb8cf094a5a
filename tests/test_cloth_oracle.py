"""Pins of the oracle's cloth path (CPU only): StVK membranes, Bergou hinges,
the Planar-DAT pair stream, element contacts and whole-step behaviour.  Each
check is independent of the oracle's own formula: finite differences, a
numpy null-space solve, closed forms, brute force, or invariants of the
contact model."""
import numpy as np
import pytest

import oracle
from avbd_amd import demos
from avbd_amd.scene import PhysicsScene, SimSettings


# -- StVK membrane (40_solver.metal:1088-1130) --------------------------------
def _rest_tri():
    x = np.array([[0.0, 0.0, 0.0], [0.01, 0.0, 0.0], [0.002, 0.009, 0.0]])
    d1, d2 = x[1] - x[0], x[2] - x[0]
    n = np.cross(d1, d2)
    e1 = d1 / np.linalg.norm(d1)
    e2 = np.cross(n / np.linalg.norm(n), e1)
    e2 /= np.linalg.norm(e2)
    Dm = np.array([[d1 @ e1, d2 @ e1], [d1 @ e2, d2 @ e2]])
    Dinv = np.linalg.inv(Dm)
    # column-major (a, b, c, d): F col 1 = d0 a + d1 b, col 2 = d0 c + d1 d
    dm = np.array([Dinv[0, 0], Dinv[1, 0], Dinv[0, 1], Dinv[1, 1]])
    area = 0.5 * np.linalg.norm(n)
    return x, dm, area


def _stvk_energy_numpy(x, dm, muA, lamA):
    # independent restatement: F = Ds Dm^-1 (3x2), E = (F^T F - I)/2
    Ds = np.stack([x[1] - x[0], x[2] - x[0]], axis=1)
    Dinv = np.array([[dm[0], dm[2]], [dm[1], dm[3]]])
    F = Ds @ Dinv
    E = 0.5 * (F.T @ F - np.eye(2))
    return muA * np.sum(E * E) + 0.5 * lamA * np.trace(E) ** 2


@pytest.mark.parametrize("slot", [0, 1, 2])
def test_membrane_gradient_is_energy_derivative(slot):
    x0, dm, area = _rest_tri()
    rng = np.random.default_rng(11 + slot)
    x = x0 + rng.normal(0, 8e-4, x0.shape)
    muA, lamA = 300.0 * area, 300.0 * area
    e, g, _ = oracle.membrane(x, dm, muA, lamA, slot)
    assert e == pytest.approx(_stvk_energy_numpy(x, dm, muA, lamA), rel=1e-10)
    h = 1e-9
    fd = np.zeros(3)
    for k in range(3):
        xp, xm = x.copy(), x.copy()
        xp[slot, k] += h
        xm[slot, k] -= h
        fd[k] = (_stvk_energy_numpy(xp, dm, muA, lamA) - _stvk_energy_numpy(xm, dm, muA, lamA)) / (2 * h)
    np.testing.assert_allclose(g, fd, rtol=1e-5, atol=1e-7 * np.abs(fd).max())


def test_membrane_rest_is_force_free_and_gauss_newton_exact_there():
    x0, dm, area = _rest_tri()
    # any rigid motion of the rest triangle: zero strain, zero force
    th = 0.7
    R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]]) @ \
        np.array([[1, 0, 0], [0, np.cos(0.3), -np.sin(0.3)], [0, np.sin(0.3), np.cos(0.3)]])
    x = x0 @ R.T + np.array([0.3, -0.2, 1.0])
    muA = lamA = 300.0 * area
    for slot in range(3):
        e, g, H = oracle.membrane(x, dm, muA, lamA, slot)
        assert abs(e) < 1e-18 and np.abs(g).max() < 1e-12
        # at zero stress the geometric term vanishes: the Gauss-Newton block is
        # the exact Hessian (finite differences of the numpy energy)
        h = 1e-6
        Hfd = np.zeros((3, 3))
        for i in range(3):
            for j in range(3):
                def E(di, dj):
                    y = x.copy()
                    y[slot, i] += di
                    y[slot, j] += dj
                    return _stvk_energy_numpy(y, dm, muA, lamA)
                Hfd[i, j] = (E(h, h) - E(h, -h) - E(-h, h) + E(-h, -h)) / (4 * h * h)
        np.testing.assert_allclose(H, Hfd, rtol=1e-4, atol=1e-6 * np.abs(Hfd).max())


# -- Bergou hinge (GPUSolver.swift:1576-1631, 40_solver.metal:1136-1150) -------
def _hinge_points():
    # shared edge v0-v1 along x, opposite vertices on either side, slightly non-symmetric
    return np.array([[0.0, 0.0, 0.0], [0.01, 0.0, 0.0], [0.004, 0.008, 0.0], [0.006, -0.007, 0.0]], np.float32)


def test_hinge_constants_null_space_and_cotangent_scale():
    x = _hinge_points()
    K, s = oracle.hinge(x, 1e-3, 2e-3)
    # independent: null space of [1; x; y] (numpy SVD), scaled so K[2] equals
    # cot(angle at v0) + cot(angle at v1) in the first triangle
    A = np.stack([np.ones(4), x[:, 0], x[:, 1]]).astype(np.float64)
    ns = np.linalg.svd(A)[2][-1]

    def cot(ap, a, b):
        u, v = x[a] - x[ap], x[b] - x[ap]
        return float(u[:2] @ v[:2]) / abs(float(u[0] * v[1] - u[1] * v[0]))

    want = cot(0, 1, 2) + cot(1, 0, 2)
    ref = ns * (want / ns[2])
    np.testing.assert_allclose(K, ref, rtol=2e-5, atol=1e-5)
    # textbook Bergou vector: K = (c03 + c04, c01 + c02, -(c01 + c03), -(c02 + c04)) with
    # c0j / c1j the cotangents at v0 / v1 of the triangle containing opposite vertex j
    c01, c11 = cot(0, 1, 2), cot(1, 0, 2)   # triangle (v0, v1, v2)
    c02, c12 = cot(0, 1, 3), cot(1, 0, 3)   # triangle (v0, v1, v3)
    bergou = np.array([-(c11 + c12), -(c01 + c02), c01 + c11, c02 + c12])
    np.testing.assert_allclose(K, bergou, rtol=2e-5, atol=1e-5)
    a1 = 0.5 * 0.01 * 0.008
    a2 = 0.5 * 0.01 * 0.007
    assert s == pytest.approx(3 * (1e-3 + 2e-3) / 2 / (a1 + a2), rel=1e-5)


def test_hinge_annihilates_flat_and_rigidly_moved_rest():
    x = _hinge_points().astype(np.float64)
    K, _ = oracle.hinge(x.astype(np.float32), 1e-3, 1e-3)
    K = K.astype(np.float64)
    assert np.abs(K @ x).max() < 1e-6 * np.abs(K).max() * 0.01
    rng = np.random.default_rng(3)
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    y = x @ q.T + rng.normal(size=3)
    assert np.linalg.norm(K @ y) < 1e-5 * np.abs(K).max() * 0.01
    # a folded hinge does not vanish
    z = x.copy()
    z[3, 2] = 0.003
    assert np.linalg.norm(K @ z) > 1e-2 * np.abs(K).max() * 0.003


# -- cloth topology of the reference grid (Demos7.swift:16-80) -----------------
def test_cloth_topology_counts():
    n = 12
    o = oracle.OracleSolver(demos.cloth_scene(n=n, size=0.3, height=0.3, radius=0.004, sphere_radius=0.1))
    ci = o.clothInfo()
    assert ci["num_tris"] == 2 * (n - 1) ** 2
    assert ci["num_edges"] == 2 * n * (n - 1) + (n - 1) ** 2
    assert ci["num_particles"] == n * n
    assert ci["num_membranes"] == 2 * (n - 1) ** 2
    assert ci["num_bends"] == ci["num_edges"] - 4 * (n - 1)   # every interior edge
    # capacities (GPUSolver.swift:423-438): 64 / particle + 128 / edge pairs
    assert ci["max_dat_pairs"] == 64 * n * n + 128 * ci["num_edges"]
    e = ci["edges"]
    assert np.all(e[:, 0] < e[:, 1]) and len({tuple(r) for r in e}) == len(e)


# -- Planar-DAT pair stream: brute force over all V-T and E-E pairs ------------
def _closest_tri(p, a, b, c):
    ab, ac, ap = b - a, c - a, p - a
    d1, d2 = ab @ ap, ac @ ap
    if d1 <= 0 and d2 <= 0:
        return a
    bp = p - b
    d3, d4 = ab @ bp, ac @ bp
    if d3 >= 0 and d4 <= d3:
        return b
    vc = d1 * d4 - d3 * d2
    if vc <= 0 and d1 >= 0 and d3 <= 0:
        return a + ab * (d1 / (d1 - d3))
    cp = p - c
    d5, d6 = ab @ cp, ac @ cp
    if d6 >= 0 and d5 <= d6:
        return c
    vb = d5 * d2 - d1 * d6
    if vb <= 0 and d2 >= 0 and d6 <= 0:
        return a + ac * (d2 / (d2 - d6))
    va = d3 * d6 - d5 * d4
    if va <= 0 and (d4 - d3) >= 0 and (d5 - d6) >= 0:
        return b + (c - b) * ((d4 - d3) / ((d4 - d3) + (d5 - d6)))
    den = va + vb + vc
    return a + ab * (vb / den) + ac * (vc / den)


def _seg_dist(p0, p1, q0, q1):
    # dense sampling bound refined by a projected search (exact enough for the band test)
    best = np.inf
    for s in np.linspace(0, 1, 41):
        a = p0 + (p1 - p0) * s
        d = q1 - q0
        t = np.clip(((a - q0) @ d) / max(d @ d, 1e-30), 0, 1)
        best = min(best, np.linalg.norm(a - (q0 + d * t)))
    return best


def _exact_seg_dist(p0, p1, q0, q1):
    from scipy.optimize import minimize
    f = lambda st: np.linalg.norm((p0 + (p1 - p0) * st[0]) - (q0 + (q1 - q0) * st[1]))
    r = minimize(f, x0=[0.5, 0.5], bounds=[(0, 1), (0, 1)], method="L-BFGS-B")
    return min(r.fun, _seg_dist(p0, p1, q0, q1))


def test_dat_pairs_equal_brute_force():
    n = 10
    s = demos.cloth_scene(n=n, size=0.09, height=0.3, radius=0.004, sphere_radius=0.1, jitter=0.0)
    o = oracle.OracleSolver(s)
    ci = o.clothInfo()
    # crumple the sheet: fold it over itself so V-T and E-E pairs appear
    pos = o.positions().astype(np.float64)
    ids = ci["particles"]
    P = pos[ids].copy()
    xs = P[:, 0]
    fold = xs > 0.0
    P[fold, 0] = -P[fold, 0] + 0.004
    P[fold, 2] = P[fold, 2] + 0.006 + 0.002 * np.sin(50 * P[fold, 1])
    pos[ids] = P
    pairs = o.datPairsAt(pos.astype(np.float32))
    got = {tuple(int(v) for v in r) for r in pairs}
    assert len(got) == len(pairs), "duplicate pairs"
    assert any(r[0] == 0 for r in pairs) and any(r[0] == 1 for r in pairs)

    x = pos.astype(np.float32)
    cs, rq = np.float32(ci["elem_cell_size"]), np.float32(ci["dat_query_radius"])
    tris = np.array([t[0] for t in s.tris], np.int64)
    edges = ci["edges"].astype(np.int64)
    nbr = {v: set() for v in range(len(x))}
    for a, b in edges:
        nbr[a].add(b)
        nbr[b].add(a)

    def cell(p):
        return np.floor(p / cs).astype(np.int64)

    def rng(lo, hi):
        c0 = cell(lo)
        return c0, np.minimum(cell(hi), c0 + 3)

    sure, band = set(), set()
    r2 = float(rq) * float(rq)
    for v in ids:
        cv = cell(x[v])
        for ti, t in enumerate(tris):
            if v in t or any(u in nbr[v] for u in t):
                continue
            c0, c1 = rng(x[t].min(0) - rq, x[t].max(0) + rq)
            if np.any(cv < c0) or np.any(cv > c1):
                continue
            q = _closest_tri(*(x64 := [x[v].astype(np.float64)] + [x[u].astype(np.float64) for u in t]))
            d2 = float(np.sum((x64[0] - q) ** 2))
            key = (0, int(v), ti)
            if d2 <= r2 * (1 - 1e-5):
                sure.add(key)
            elif d2 <= r2 * (1 + 1e-5):
                band.add(key)
    for ga, (a0, a1) in enumerate(edges):
        ac0, ac1 = rng(np.minimum(x[a0], x[a1]), np.maximum(x[a0], x[a1]))
        for gb in range(ga + 1, len(edges)):
            b0, b1 = edges[gb]
            if len({a0, a1, b0, b1}) < 4:
                continue
            if b0 in nbr[a0] or b1 in nbr[a0] or b0 in nbr[a1] or b1 in nbr[a1]:
                continue
            bc0, bc1 = rng(np.minimum(x[b0], x[b1]) - rq, np.maximum(x[b0], x[b1]) + rq)
            if np.any(np.maximum(ac0, bc0) > np.minimum(ac1, bc1)):
                continue
            lo_bound = np.linalg.norm(0.5 * (x[a0] + x[a1]) - 0.5 * (x[b0] + x[b1])) - \
                0.5 * (np.linalg.norm(x[a1] - x[a0]) + np.linalg.norm(x[b1] - x[b0]))
            if lo_bound > float(rq) * 1.001:
                continue
            d = _exact_seg_dist(*[x[u].astype(np.float64) for u in (a0, a1, b0, b1)])
            key = (1, ga, gb)
            if d * d <= r2 * (1 - 1e-5):
                sure.add(key)
            elif d * d <= r2 * (1 + 1e-4):
                band.add(key)
    assert sure <= got, f"missing {sorted(sure - got)[:5]}"
    assert got <= sure | band, f"extra {sorted(got - sure - band)[:5]}"
    # producer-major, partner ascending
    order = [(0 if r[0] == 0 else 1, int(np.where(ids == r[1])[0][0]) if r[0] == 0 else r[1], r[2]) for r in pairs]
    assert order == sorted(order)


# -- whole-step behaviour -------------------------------------------------------
def test_rest_sheet_translates_force_free():
    # a rest-shaped sheet moving at uniform velocity without gravity: springs,
    # membranes and hinges exert no force at rest, so the predictor is already
    # the solution and the sheet translates exactly (x = x0 + v t) -- the
    # speed is kept under the Planar-DAT displacement cap 0.5 * 0.85 * rq per
    # step, and a uniform motion approaches no division plane
    s = demos.cloth_scene(n=8, size=0.1, height=2.0, radius=0.002, jitter=0.0, particle_damping=0.0, viscosity=0.0,
                          ground=False, sphere_z=-50.0, iterations=5)
    s.settings.gravity = 0.0
    o = oracle.OracleSolver(s, precision=64)
    n = len(s.bodies) - 1
    p0, r0, _, _ = o.bodyStates()
    idx = np.arange(1, n + 1, dtype=np.uint32)
    v = np.tile(np.array([0.15, -0.1, 0.2], np.float32), (n, 1))
    assert np.linalg.norm(v[0]) * s.settings.dt < 0.5 * 0.85 * o.clothInfo()["dat_query_radius"]
    o.setBodyStates(idx, p0[1:], r0[1:], v, np.zeros_like(v))
    steps = 30
    o.step(steps)
    want = p0[1:].astype(np.float64) + v.astype(np.float64) * (steps * s.settings.dt)
    np.testing.assert_allclose(o.positions()[1:], want, atol=2e-6)


def _drape(n=16, steps=90, precision=32):
    s = demos.cloth_scene(n=n, size=0.4, height=0.25, radius=0.004, sphere_radius=0.1, sphere_z=0.0, iterations=20)
    o = oracle.OracleSolver(s, precision=precision)
    o.step(steps)
    return s, o


def test_drape_settles():
    # ClothTests.testDrapeSettles (ClothTests.swift:205-219): Demos.drape(res: 24)
    # for 720 frames; the sheet must come to rest without pumping energy, with
    # no deep self-penetration and an intact weave
    s = demos.drape_scene(24)
    o = oracle.OracleSolver(s)
    o.step(720)
    p, _, v, _ = o.bodyStates()
    m = np.array([b["density"] * 4 / 3 * np.pi * (b["size"][0] / 2) ** 3 if b["particle"] else 0.0
                  for b in s.bodies])
    ke = 0.5 * float(np.sum(m * np.sum(v.astype(np.float64) ** 2, axis=1)))
    gap, stretch = demos.cloth_metrics(s, p)
    assert ke < 0.8, f"draped cloth must dissipate (KE {ke})"
    assert gap > -0.05, f"no deep self-penetration (gap {gap})"
    assert stretch < 0.15, f"drape hangs without tearing the weave (stretch {stretch})"
    # the centre of the sheet rests on the sphere top (r = 0.55, contact radius + particle radius)
    c = np.argmin(np.linalg.norm(p[2:, :2], axis=1)) + 2
    assert abs(p[c, 2] - (1.0 + 0.55 + s.bodies[c]["size"][0] / 2)) < 0.03
    assert o.clothStats()["num_soft_rt"] > 0   # sphere-vs-triangle contacts are live


def test_element_contact_dual_invariants():
    s, o = _drape(steps=60)
    sc = o.softContacts()
    assert len(sc) > 0
    lam = sc["lambda"][:, :3]
    mu = sc["normal"][:, 3]
    pen = sc["penalty"][:, :3]
    assert np.all(lam[:, 0] <= 0)
    assert np.all(np.hypot(lam[:, 1], lam[:, 2]) <= mu * np.abs(lam[:, 0]) * (1 + 1e-5) + 1e-12)
    assert np.all(pen <= 1e10) and np.all(pen >= 1.0)
    assert np.all(lam[:, 0] >= -sc["C0"][:, 3] * (1 + 1e-6))   # mass-scaled normal cap


def _two_sheets(gap=0.009):
    s = PhysicsScene("two sheets", SimSettings(iterations=20, particleDamping=1.5, clothViscosity=0.25))
    s.addBody((200, 200, 2), 0.0, 0.6, (0, 0, -1))
    rng = demos.SplitMix64(5)
    n, size, r = 10, 0.12, 0.003
    h = size / (n - 1)
    for layer, z in enumerate((0.02, 0.02 + gap)):
        base = len(s.bodies)
        for i in range(n):
            for j in range(n):
                s.addParticle(r, 0.3 * h * h, (i * h - size / 2 + 0.3 * h * layer, j * h - size / 2,
                                               z + rng.uniform(-1e-4, 1e-4)), friction=0.6)
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


def test_stacked_sheets_do_not_pass_through():
    s = _two_sheets()
    o = oracle.OracleSolver(s)
    n = 100
    o.step(60)
    p = o.positions()
    lower, upper = p[1:1 + n], p[1 + n:1 + 2 * n]
    st = o.clothStats()
    assert st["dat_pairs"] > 0 and st["soft_candidates"] > st["num_soft_rt"]   # V-T / E-E contacts are live
    # the upper sheet stays above the lower one over their overlap
    assert np.median(upper[:, 2]) > np.median(lower[:, 2]) + 0.002
    assert lower[:, 2].min() > 0.0   # lower sheet rests on the ground (contact radius above z = 0)
