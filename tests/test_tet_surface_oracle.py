"""CPU pins of the oracle's tet collision surfaces (GPUSolver.swift:384-390,
872-897, 1289-1334, 1476-1522): the boundary faces of a lattice block
against the closed form (6 (n-1)^2 squares, two triangles each), a closed
2-manifold (every edge on two faces, V - E + F = 2), outward winding on a
convex block, ascending order; the surface groups, self-collision, pair
stream and limiter selection (GPUSolver.swift:93-99) of one block, two
blocks and a block beside a cloth sheet; and the tracked V-T contacts that
keep a dropped block on another."""
import numpy as np

import oracle
import reference_scenes as R
from avbd_amd.demos import soft_block_scene
from avbd_amd.scene import PhysicsScene, SimSettings


def _block(n=4, spacing=0.34, center=(0, 0, 1.0)):
    s = PhysicsScene("block", SimSettings(iterations=5))
    R._ground(s)
    ids = R.add_soft_lattice(s, center, n, n, n, spacing)
    return s, ids


def _manifold_checks(tris, pos, center):
    faces = {tuple(sorted(t)) for t in tris.tolist()}
    assert len(faces) == len(tris)                                  # no face twice
    edges = {}
    for a, b, c in tris.tolist():
        for u, v in ((a, b), (b, c), (a, c)):
            edges[min(u, v), max(u, v)] = edges.get((min(u, v), max(u, v)), 0) + 1
    assert set(edges.values()) == {2}                               # closed 2-manifold
    verts = {v for t in tris.tolist() for v in t}
    assert len(verts) - len(edges) + len(tris) == 2                 # sphere topology
    p = pos[tris]
    n = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
    out = p.mean(axis=1) - np.asarray(center)
    assert np.all(np.sum(n * out, axis=1) > 0)                      # wound outward
    assert [tuple(t) for t in tris.tolist()] == sorted(tuple(t) for t in tris.tolist())
    return verts, edges


def test_lattice_block_boundary_faces():
    for n in (2, 3, 4):
        s, ids = _block(n)
        o = oracle.OracleSolver(s)
        tris = o.collisionTris()
        assert len(tris) == 6 * (n - 1) ** 2 * 2
        pos = np.array([b["position"] for b in s.bodies], np.float64)
        verts, edges = _manifold_checks(tris, pos, (0, 0, 1.0))
        lattice = np.array(ids).reshape(n, n, n)
        inner = set(lattice[1:-1, 1:-1, 1:-1].ravel().tolist())
        assert verts == set(ids) - inner
        info = o.clothInfo()
        assert info["num_tris"] == len(tris) and info["num_edges"] == len(edges)
        assert sorted(info["particles"].tolist()) == sorted(verts)
        assert info["num_contact_edges"] == 0 and info["num_membranes"] == 0
        assert info["pair_stream"] == 0 and info["dat_on"] == 0     # one surface, no self-collision


def test_kuhn_block_boundary_faces():
    # the 6-tet split of the benchmark block: same closed-form face count
    s = soft_block_scene(cells=5)
    o = oracle.OracleSolver(s)
    tris = o.collisionTris()
    assert len(tris) == 6 * 5 * 5 * 2
    pos = np.array([b["position"] for b in s.bodies], np.float64)
    c = pos[[i for i, b in enumerate(s.bodies) if b.get("particle")]].mean(axis=0)
    _manifold_checks(tris, pos, c)
    assert o.clothInfo()["pair_stream"] == 0 and o.clothInfo()["dat_on"] == 0


def test_surface_groups_and_limiter_selection():
    s, top = R.stacked_soft_blocks()
    o = oracle.OracleSolver(s)
    info = o.clothInfo()
    assert info["num_tris"] == 108 + 48 and info["num_contact_edges"] == 0
    assert (info["pair_stream"], info["dat_on"], info["iso_on"]) == (0, 0, 1)   # closed tet surfaces: isotropic
    # a cloth sheet beside a block: its triangles first, its edges the contact prefix
    s = PhysicsScene("sheet+block", SimSettings(iterations=5))
    R._ground(s)
    sheet = [s.addParticle(0.01, 0.01, (x * 0.1 + 2, y * 0.1, 1)) for y in range(3) for x in range(3)]
    for y in range(2):
        for x in range(2):
            a, b, c, d = sheet[3 * y + x], sheet[3 * y + x + 1], sheet[3 * y + x + 3], sheet[3 * y + x + 4]
            s.addTri((a, b, d), mu=100.0)
            s.addTri((a, d, c), mu=100.0)
    R.add_soft_lattice(s, (0, 0, 1.0), 2, 2, 2, 0.3)
    o = oracle.OracleSolver(s)
    info = o.clothInfo()
    tris = o.collisionTris()
    assert info["num_tris"] == 8 + 12 and info["num_contact_edges"] == 16 and info["num_membranes"] == 8
    assert np.array_equal(tris[:8], np.array([t[0] for t in s.tris], np.uint32))
    assert (info["pair_stream"], info["dat_on"], info["iso_on"]) == (0, 0, 1)   # tets select the isotropic mode


def test_dropped_block_rests_on_block_through_vt_contacts():
    s, top = R.stacked_soft_blocks()
    o = oracle.OracleSolver(s)
    o.step(120)
    sc = o.softContacts()
    vt = sc[(sc["lambda"][:, 3].view(np.uint32) >> 28) == 1]
    assert len(vt) > 0
    # every tracked V-T contact joins the two blocks (no same-block pair on
    # a non-self-colliding surface)
    top = set(top)
    for ids in vt["ids"].tolist():
        assert (ids[0] in top) != (ids[1] in top)
    o.step(280)
    p = o.positions()[sorted(top)]
    assert p[:, 2].min() > 0.95 and p[:, 2].mean() > 1.15
