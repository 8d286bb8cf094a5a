"""Bit-exact integer outputs at the benchmark sizes, against the oracle.

* The 256^2 cloth drape: topology, derived parameters and the static
  palette at create; then both backends are set to the same S-folded state
  of the full sheet (the two halves of every row folded over, 4 mm apart on
  a band of 64 rows) and stepped
  once: the V-T / E-E pair stream, the element-contact ids, keys and flags
  equal the oracle's, one frame later positions agree within 1e-4 m.
* The 4096 batched rigid scenes: the first frame's hash cell lists, rigid
  pair stream, manifold headers and feature keys, and the contact-aware
  dynamic palette equal the oracle's.
* The 1M-tet block (55^3 cells): the collision triangles the device
  extracts from the tet meshes at create (sort-based boundary faces), the
  surface edges and particles, and the static palette equal the oracle's."""
import numpy as np
import pytest

import oracle
from avbd_amd import GPUSolver, demos
from conftest import flush_print

pytestmark = pytest.mark.gpu


def _folded_state(scene, n=256, gap=0.004, band=(96, 160), far=0.03):
    """Every row folded at its middle: the second half (y beyond the centre)
    mirrored back over the first (particle (i, j) sits at x_i, y_j), `gap`
    above it for the rows of `band` and `far` (beyond the contact query)
    elsewhere: a double layer of the full sheet in contact over a band."""
    p = np.array([b["position"] for b in scene.bodies], np.float64)
    first = len(scene.bodies) - n * n
    grid = p[first:].reshape(n, n, 3).copy()
    half = n // 2
    ym = 0.5 * (grid[:, half - 1, 1] + grid[:, half, 1])
    lift = np.full(n, far)
    lift[band[0]:band[1]] = gap
    for j in range(half, n):
        grid[:, j, 1] = 2 * ym - grid[:, j, 1]
        grid[:, j, 2] = grid[:, j, 2] + lift
    p[first:] = grid.reshape(-1, 3)
    nb = len(scene.bodies)
    rot = np.tile(np.array([0, 0, 0, 1], np.float32), (nb, 1))
    zero = np.zeros((nb, 3), np.float32)
    return p.astype(np.float32), rot, zero, zero


def test_cloth_256_topology_palette_and_folded_contacts():
    flush_print("[gpu case] 256^2 cloth: topology, palette, folded-sheet pair stream and contacts bit-exact")
    s = demos.cloth_scene()
    with GPUSolver(s) as g:
        o = oracle.OracleSolver(s)
        cg, co = g.clothInfo(), o.clothInfo()
        for k in ("num_tris", "num_edges", "num_particles", "num_membranes", "num_bends", "max_soft",
                  "max_dat_pairs", "elem_cell_size", "elem_margin", "dat_query_radius",
                  "surface_contact_cell_size"):
            assert cg[k] == co[k], k
        assert np.array_equal(cg["edges"], co["edges"])
        assert np.array_equal(cg["particles"], co["particles"])
        assert np.array_equal(g.debugColors(), o.debugColors())
        p, r, v, w = _folded_state(s)
        idx = np.arange(len(s.bodies), dtype=np.uint32)
        g.setBodyStates(idx, p, r, v, w)
        o.setBodyStates(idx, p, r, v, w)
        g.step()
        o.step()
        pg, po = g.datPairs(), o.datPairs()
        flush_print(f"  pair stream {len(pg)} (oracle {len(po)}), element contacts {g.stats().num_soft}")
        assert len(po) > 10000
        assert np.array_equal(pg, po)
        sg, so = g.softContacts(), o.softContacts()
        assert len(sg) == len(so) and len(sg) > 10000
        assert np.array_equal(sg["ids"], so["ids"])
        for f in ("lambda", "penalty", "anchorA"):
            assert np.array_equal(sg[f][:, 3].view(np.uint32), so[f][:, 3].view(np.uint32)), f
        err = float(np.abs(g.positions() - o.positions()).max())
        flush_print(f"  max|x_gpu - x_oracle| after the folded frame = {err:.3e} m")
        assert err < 1e-4


def test_batched_4096_first_frame_bit_exact():
    flush_print("[gpu case] 4096 batched scenes: first-frame cells, pairs, manifolds, keys, palette bit-exact")
    s, _ = demos.batched_scene(4096)
    with GPUSolver(s) as g:
        o = oracle.OracleSolver(s)
        g.step()
        o.step()
        cg, co = g.cellLists(), o.cellLists()
        assert cg["hash_size"] == co["hash_size"] and cg["cell_size"] == co["cell_size"]
        assert np.array_equal(cg["count"], co["count"])
        assert np.array_equal(cg["start"], co["start"])
        assert np.array_equal(cg["bodies"], co["bodies"])
        assert np.array_equal(g.pairs(), o.pairs())
        gm = g.manifolds()
        h, _, _ = o.manifolds()
        assert np.array_equal(gm["header"][:, :3], h)
        keys = gm["c"]["rA"][:, :, 3].copy().view(np.uint32)
        keys[np.arange(8)[None, :] >= gm["header"][:, 2][:, None]] = 0xFFFFFFFF
        assert np.array_equal(keys, o.manifoldKeys())
        assert np.array_equal(g.debugColors(), o.debugColors())
        flush_print(f"  {len(s.bodies)} bodies, {len(h)} pairs, {int((h[:, 2] > 0).sum())} active manifolds, "
                    f"{g.stats().num_colors} colors")


def test_million_tet_block_surface_and_palette():
    flush_print("[gpu case] 55^3 block: device-extracted boundary faces, surface topology, palette bit-exact")
    s = demos.soft_block_scene(cells=55)
    with GPUSolver(s) as g:
        o = oracle.OracleSolver(s)
        tg, to = g.collisionTris(), o.collisionTris()
        flush_print(f"  {len(tg)} boundary faces (oracle {len(to)}); create phases {g.createProfile()}")
        assert len(to) == 6 * 55 * 55 * 2
        assert np.array_equal(tg, to)
        cg, co = g.clothInfo(), o.clothInfo()
        assert np.array_equal(cg["edges"], co["edges"])
        assert np.array_equal(cg["particles"], co["particles"])
        assert np.array_equal(g.debugColors(), o.debugColors())
