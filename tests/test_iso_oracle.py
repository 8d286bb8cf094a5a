"""CPU pins of the oracle's isotropic OGC limiter (25_cloth.metal:350-940,
40_solver.metal:697-717, 1361-1382; selection GPUSolver.swift:93-99) against
brute force over every triangle and edge, at the step-start pose the
trackers read:

* tracked sets: eligible (not own, not 1-ring, not a same-solid pair),
  ascending in distance, and -- for the elements whose staggered reseed
  falls on the frame ((id + frame) & 7 == 0 rescans the grid whatever the
  speed) -- holding every eligible element closer than the fourth, among
  all those within one cell (which the 27-cell scan is shown to reach);
* clearance bounds: never below the true minimum over the pairs that may
  lower them, and for reseeded vertices never above their own closest
  non-2-ring distance within a cell;
* truncation: after a frame every bounded surface particle lies within
  max(0.45 sqrt(bound), 0.2 r) of its anchor pose;
* the reference's stacked-blocks contract (SoftBodyTests.swift:98-123)
  through the tracked V-T contacts.
A planted mistake (the tracker's 1-ring exclusion dropped) is shown to fail
the tracked-set pin."""
import numpy as np
import pytest

import oracle
import reference_scenes as R
from avbd_amd.scene import PhysicsScene, SimSettings

SC_VT, SC_EE = 1, 3


def _closest_tri_d2(p, a, b, c):
    """Squared distance of points p (n, 3) to triangle (a, b, c), float64,
    by projection onto the plane when inside, else the closest edge."""
    n = np.cross(b - a, c - a)
    nn = np.dot(n, n)
    best = np.full(len(p), np.inf)
    if nn > 1e-30:
        w = p - a
        t = w @ n / nn
        q = p - np.outer(t, n)
        v0, v1, v2 = b - a, c - a, q - a
        d00, d01, d11 = v0 @ v0, v0 @ v1, v1 @ v1
        d20, d21 = v2 @ v0, v2 @ v1
        den = d00 * d11 - d01 * d01
        bv = (d11 * d20 - d01 * d21) / den
        bw = (d00 * d21 - d01 * d20) / den
        inside = (bv >= 0) & (bw >= 0) & (bv + bw <= 1)
        best = np.where(inside, t * t * nn, best)
    for u, v in ((a, b), (b, c), (a, c)):
        best = np.minimum(best, _point_seg_d2(p, u, v))
    return best


def _point_seg_d2(p, u, v):
    d = v - u
    dd = max(float(d @ d), 1e-300)
    t = np.clip((p - u) @ d / dd, 0.0, 1.0)
    q = u + np.outer(t, d)
    return np.sum((p - q) ** 2, axis=1)


def _seg_seg_d2(a0, a1, b0, b1):
    """Squared distance between two segments, float64: the minimum over the
    four endpoint-to-segment distances and the interior critical point."""
    best = min(_point_seg_d2(a0[None], b0, b1)[0], _point_seg_d2(a1[None], b0, b1)[0],
               _point_seg_d2(b0[None], a0, a1)[0], _point_seg_d2(b1[None], a0, a1)[0])
    d1, d2, r = a1 - a0, b1 - b0, a0 - b0
    a, e, b = d1 @ d1, d2 @ d2, d1 @ d2
    den = a * e - b * b
    if den > 1e-18 * max(a * e, 1e-300):
        c, f = d1 @ r, d2 @ r
        s = (b * f - c * e) / den
        t = (a * f - b * c) / den
        if 0 <= s <= 1 and 0 <= t <= 1:
            best = min(best, float(np.sum((a0 + s * d1 - b0 - t * d2) ** 2)))
    return best


class Topo:
    """1-rings, 2-rings, surface groups and authored-vertex flags of the
    collision surface, from the oracle's triangle and edge lists."""

    def __init__(self, o, scene):
        info = o.clothInfo()
        self.tris = o.collisionTris().astype(np.int64)
        self.edges = info["edges"].astype(np.int64)
        self.ne = info["num_contact_edges"]
        self.particles = info["particles"].astype(np.int64)
        self.cs = info["surface_contact_cell_size"]
        nb = len(scene.bodies)
        self.r1 = [set() for _ in range(nb)]
        for a, b in self.edges:
            self.r1[a].add(b)
            self.r1[b].add(a)
        self.r2 = [set() for _ in range(nb)]
        for v in range(nb):
            for u in self.r1[v]:
                self.r2[v] |= self.r1[u]
            self.r2[v] |= self.r1[v]
            self.r2[v].discard(v)
        parent = list(range(nb))

        def find(x):
            while parent[x] != x:
                parent[x] = parent[parent[x]]
                x = parent[x]
            return x

        for t in self.tris:
            for k in (1, 2):
                ra, rb = find(t[0]), find(t[k])
                if ra != rb:
                    parent[rb] = ra
        self.group = np.zeros(nb, np.int64)
        roots = {}
        for t in self.tris:
            roots.setdefault(find(t[0]), len(roots) + 1)
        for v in set(self.tris.ravel().tolist()):
            self.group[v] = roots[find(v)]
        self.cloth = np.zeros(nb, bool)
        for t in scene.tris:
            self.cloth[list(t[0])] = True
        self.radius = np.array([abs(b["size"][0]) / 2 for b in scene.bodies])

    def solid(self, v):
        return self.group[v] != 0 and not self.cloth[v]

    def same_solid_tri(self, v, t):
        ids = self.tris[t]
        return self.solid(v) and all(self.group[i] == self.group[v] and not self.cloth[i] for i in ids)

    def same_solid_edge(self, ea, eb):
        a, b = self.edges[ea], self.edges[eb]
        sa = self.group[a[0]] != 0 and self.group[a[1]] == self.group[a[0]] and not self.cloth[a].any()
        return sa and all(self.group[i] == self.group[a[0]] and not self.cloth[i] for i in b)


def _vt_d2(topo, p):
    a, b, c = (p[topo.tris[:, k]] for k in range(3))
    out = np.empty((len(topo.particles), len(topo.tris)))
    for j in range(len(topo.tris)):
        out[:, j] = _closest_tri_d2(p[topo.particles], a[j], b[j], c[j])
    return out


def _check_vt(topo, p, frame, vt, bound, planted=False, refreshed=False):
    d2 = _vt_d2(topo, p)
    cs2 = float(topo.cs) ** 2
    checked = 0
    for row, v in enumerate(topo.particles):
        tr = [int(t) for t in vt[v] if t != 0xFFFFFFFF]
        elig = []
        for t in range(len(topo.tris)):
            ids = topo.tris[t]
            if v in ids or topo.same_solid_tri(v, t):
                continue
            if not planted and any(i in topo.r1[v] for i in ids):
                continue
            elig.append(t)
        elig = set(elig)
        assert set(tr) <= elig, (v, tr)
        dd = [d2[row, t] for t in tr]
        assert all(dd[i] <= dd[i + 1] * (1 + 1e-4) + 1e-9 for i in range(len(dd) - 1)), (v, dd)
        if (int(v) + frame) & 7 == 0:
            checked += 1
            near = [t for t in elig if d2[row, t] < cs2]
            cut = dd[3] * (1 - 1e-4) - 1e-9 if len(tr) == 4 else np.inf
            missing = [t for t in near if d2[row, t] < cut and t not in tr]
            assert not missing, (v, tr, missing)
            own = [d2[row, t] for t in near if not any(i in topo.r2[v] for i in topo.tris[t])]
            if own and not refreshed:   # a refresh recomputes the bounds at the live pose
                assert bound[v] <= min(own) * (1 + 1e-4) + 1e-9, (v, bound[v], min(own))
    return checked


def _bound_floor(topo, p, d2vt):
    """The true minimum of every pair allowed to lower each vertex's bound."""
    nb = len(p)
    floor = np.full(nb, np.inf)
    cs2 = float(topo.cs) ** 2
    for row, u in enumerate(topo.particles):
        for t in range(len(topo.tris)):
            ids = topo.tris[t]
            if u in ids or topo.same_solid_tri(u, t) or any(i in topo.r2[u] for i in ids):
                continue
            d = d2vt[row, t]
            if d >= cs2:
                continue
            for w in (u, *ids):
                floor[w] = min(floor[w], d)
    for ea in range(topo.ne):
        for eb in range(topo.ne):
            a, b = topo.edges[ea], topo.edges[eb]
            if ea == eb or set(a) & set(b) or topo.same_solid_edge(ea, eb):
                continue
            d = _seg_seg_d2(p[a[0]], p[a[1]], p[b[0]], p[b[1]])
            if d >= cs2:
                continue
            for w in (*a, *b):
                floor[w] = min(floor[w], d)
    return floor


def _run_and_check(s, frames, start, planted=False):
    o = oracle.OracleSolver(s)
    assert o.clothInfo()["iso_on"] == 1
    topo = Topo(o, s)
    reseeded = 0
    for f in range(frames):
        p0 = o.bodyStates()[0].astype(np.float64)
        o.step()
        if f < start:
            continue
        st = o.isoState()
        reseeded += _check_vt(topo, p0, f, st["vt_track"], st["bound"], planted, st["refreshes"] > 0)
        yield o, topo, p0, st, f
    assert reseeded > 0


def test_mode_selection():
    s, _ = R.stacked_soft_blocks()
    info = oracle.OracleSolver(s).clothInfo()
    assert (info["iso_on"], info["dat_on"], info["pair_stream"]) == (1, 0, 0)
    assert info["max_soft"] == 4 * len(info["particles"]) + 4 * info["num_tris"]
    one = R.stacked_soft_blocks.__wrapped__() if hasattr(R.stacked_soft_blocks, "__wrapped__") else None
    assert one is None
    s = PhysicsScene("block", SimSettings(iterations=5))
    R._ground(s)
    R.add_soft_lattice(s, (0, 0, 1.0), 3, 3, 3, 0.34)
    info = oracle.OracleSolver(s).clothInfo()
    assert (info["iso_on"], info["dat_on"], info["pair_stream"]) == (0, 0, 0)   # one closed surface: identity
    info = oracle.OracleSolver(R.sheets_and_block()).clothInfo()
    assert (info["iso_on"], info["dat_on"], info["pair_stream"]) == (1, 0, 0)
    assert info["num_contact_edges"] > 0


def test_blocks_trackers_bounds_truncation():
    s, top = R.stacked_soft_blocks()
    n_trunc = 0
    for o, topo, p0, st, f in _run_and_check(s, 64, 48):
        d2 = _vt_d2(topo, p0)
        floor = _bound_floor(topo, p0, d2)
        b = st["bound"].astype(np.float64)
        fin = b < 1e37
        assert fin.sum() > 0
        if st["refreshes"] == 0:
            assert np.all(b[fin] >= floor[fin] * (1 - 1e-4) - 1e-9)
        assert np.all(np.isinf(floor[~fin]) | (floor[~fin] >= 0))
        p = o.bodyStates()[0].astype(np.float64)
        bv = np.maximum(0.45 * np.sqrt(b), 0.2 * topo.radius)
        moved = np.linalg.norm(p - st["anchor"], axis=1)
        surf = topo.particles[fin[topo.particles]]
        assert np.all(moved[surf] <= bv[surf] * (1 + 1e-5) + 1e-6)
        n_trunc += int(np.sum(moved[surf] >= bv[surf] * (1 - 1e-3)))
    # the contract of the reference's test: the dropped block rests on the other
    sc = o.softContacts()
    kinds = sc["lambda"][:, 3].view(np.uint32) >> 28
    vt = sc[kinds == SC_VT]
    assert len(vt) > 0
    topset = set(top)
    for ids in vt["ids"]:
        assert (int(ids[0]) in topset) != (int(ids[1]) in topset)
    o.step(200)
    pz = o.positions()[sorted(top)][:, 2]
    assert pz.min() > 0.95 and pz.mean() > 1.15


def test_sheets_vt_ee_trackers():
    s = R.sheets_and_block()
    seen_ee = 0
    for o, topo, p0, st, f in _run_and_check(s, 24, 14):
        ee = st["ee_track"]
        for e in range(topo.ne):
            tr = [int(x) for x in ee[e] if x != 0xFFFFFFFF]
            a = topo.edges[e]
            d = {x: _seg_seg_d2(p0[a[0]], p0[a[1]], p0[topo.edges[x][0]], p0[topo.edges[x][1]]) for x in tr}
            for x in tr:
                b = topo.edges[x]
                assert x != e and not (set(a) & set(b)) and x < topo.ne
                assert not any(i in topo.r1[a[0]] or i in topo.r1[a[1]] for i in b)
            dd = [d[x] for x in tr]
            # the tracker ranks by the f32 clamped closest-point pair, which
            # loses up to ~1e-3 of the squared distance on nearly parallel
            # edges (the two sheets' grids cross at right angles)
            assert all(dd[i] <= dd[i + 1] * (1 + 2e-3) + 1e-8 for i in range(len(dd) - 1)), (e, dd)
            if (e + f) & 7 == 0 and tr:
                seen_ee += 1
                mid = 0.5 * (p0[a[0]] + p0[a[1]])
                cut = dd[3] * (1 - 2e-3) - 1e-8 if len(tr) == 4 else np.inf
                for x in range(topo.ne):
                    b = topo.edges[x]
                    if x == e or set(a) & set(b) or x in tr:
                        continue
                    if any(i in topo.r1[a[0]] or i in topo.r1[a[1]] for i in b):
                        continue
                    if _point_seg_d2(mid[None], p0[b[0]], p0[b[1]])[0] >= float(topo.cs) ** 2:
                        continue
                    assert _seg_seg_d2(p0[a[0]], p0[a[1]], p0[b[0]], p0[b[1]]) >= cut, (e, tr, x)
    assert seen_ee > 0
    sc = o.softContacts()
    kinds = set((sc["lambda"][:, 3].view(np.uint32) >> 28).tolist())
    assert SC_VT in kinds


def test_planted_ring_exclusion_fails():
    """With the 1-ring exclusion dropped from the checker's eligibility, the
    tracked sets of the oracle are no longer the brute-force top four: the
    pin above distinguishes the rule."""
    s = R.sheets_and_block()
    with pytest.raises(AssertionError):
        for _ in _run_and_check(s, 20, 12, planted=True):
            pass
