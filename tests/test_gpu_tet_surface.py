"""GPU parity of the tet collision surfaces (GPUSolver.swift:384-390,
872-897, 1289-1334, 1476-1522) through the C ABI against the oracle: the
collision triangles (authored, then tet boundary faces), edges, surface
particles and derived parameters bit-exact on one block, two blocks, the
benchmark block and a cloth sheet beside a block; the pair stream, element
contacts and one iteration from an identical state of one block resting on
another; a 60-step trajectory against the FP32 and FP64 oracle; and a ball
dropped onto a block, whose rigid-feature contacts now reach the block's
faces (SoftBodyTests.swift:70-83).  The reference's soft-body contracts
themselves run in test_reference_contracts.py."""
import numpy as np
import pytest

import oracle
import reference_scenes as R
from avbd_amd import GPUSolver
from avbd_amd.demos import soft_block_scene
from avbd_amd.scene import PhysicsScene, SimSettings
from conftest import flush_print

pytestmark = pytest.mark.gpu

INFO_KEYS = ("num_tris", "num_edges", "num_particles", "num_membranes", "num_bends", "max_soft", "max_dat_pairs",
             "elem_cell_size", "elem_margin", "dat_query_radius", "surface_contact_cell_size", "num_contact_edges",
             "pair_stream", "dat_on")


def _sheet_and_block():
    s = PhysicsScene("sheet+block", SimSettings(iterations=5))
    R._ground(s)
    sheet = [s.addParticle(0.01, 0.01, (x * 0.1 + 2, y * 0.1, 1)) for y in range(3) for x in range(3)]
    for y in range(2):
        for x in range(2):
            a, b, c, d = sheet[3 * y + x], sheet[3 * y + x + 1], sheet[3 * y + x + 3], sheet[3 * y + x + 4]
            s.addTri((a, b, d), mu=100.0)
            s.addTri((a, d, c), mu=100.0)
    R.add_soft_lattice(s, (0, 0, 1.0), 2, 2, 2, 0.3)
    return s


@pytest.mark.parametrize("name", ["block", "stacked", "bench-block", "sheet+block"])
def test_collision_surface_topology_bitexact(name):
    flush_print(f"[gpu case] collision surface topology: {name}")
    if name == "block":
        s = PhysicsScene("block", SimSettings(iterations=5))
        R._ground(s)
        R.add_soft_lattice(s, (0, 0, 1.0), 4, 4, 4, 0.34)
    elif name == "stacked":
        s, _ = R.stacked_soft_blocks()
    elif name == "bench-block":
        s = soft_block_scene(cells=20)
    else:
        s = _sheet_and_block()
    g, o = GPUSolver(s), oracle.OracleSolver(s)
    cg, co = g.clothInfo(), o.clothInfo()
    for k in INFO_KEYS:
        assert cg[k] == co[k], k
    assert np.array_equal(g.collisionTris(), o.collisionTris())
    assert np.array_equal(cg["edges"], co["edges"])
    assert np.array_equal(cg["particles"], co["particles"])
    assert np.array_equal(g.debugColors(), o.debugColors())
    flush_print(f"  {cg['num_tris']} triangles, {cg['num_edges']} edges, dat_on {cg['dat_on']}")


@pytest.mark.parametrize("lanes", [8, 32])
def test_stacked_blocks_one_iteration_parity(lanes):
    flush_print(f"[gpu case] stacked soft blocks: pairs, contacts, one iteration, lanes={lanes}")
    s, top = R.stacked_soft_blocks()
    src = oracle.OracleSolver(s)
    src.step(120)                      # the top block rests on the bottom one
    g, o = GPUSolver(s, lanesPerBody=lanes, sweepMode=1), oracle.OracleSolver(s)
    p, r, v, w = src.bodyStates()
    idx = np.arange(len(s.bodies), dtype=np.uint32)
    g.setBodyStates(idx, p, r, v, w)
    o.setBodyStates(idx, p, r, v, w)
    st = g.settings
    st.iterations = 1
    g.settings = st
    o.set_settings(st)
    g.step()
    o.step()
    pg, po = g.datPairs(), o.datPairs()
    sg, so = g.softContacts(), o.softContacts()
    flush_print(f"  pairs gpu {len(pg)} oracle {len(po)}, soft gpu {len(sg)} oracle {len(so)}")
    assert len(po) > 0 and np.array_equal(pg, po)
    assert len(sg) == len(so) > 0 and np.array_equal(sg["ids"], so["ids"])
    np.testing.assert_allclose(sg["normal"], so["normal"], atol=1e-6)
    assert np.array_equal(g.pairs(), o.pairs())
    err = float(np.abs(g.positions() - o.positions()).max())
    flush_print(f"  max|x_gpu - x_oracle| = {err:.3e} m")
    assert err <= 1e-5


def test_stacked_blocks_trajectory_60_steps():
    flush_print("[gpu case] stacked soft blocks: 60-step trajectory vs oracle FP32 / FP64")
    s, top = R.stacked_soft_blocks()
    g, o, o64 = GPUSolver(s), oracle.OracleSolver(s), oracle.OracleSolver(s, precision=64)
    for _ in range(60):
        g.step()
        o.step()
        o64.step()
    pg, po, p64 = g.positions(), o.positions(), o64.bodyStates64()[0]
    rms = float(np.sqrt(np.mean(np.sum((pg - po) ** 2, axis=1))))
    e64 = float(np.sqrt(np.mean(np.sum((po - p64) ** 2, axis=1))))
    flush_print(f"  RMS gpu-oracle32 {rms:.3e} m, oracle32-oracle64 {e64:.3e} m")
    assert np.all(np.isfinite(pg))
    assert rms < max(2e-3, 3.0 * e64)


def test_ball_on_block_feature_contacts():
    flush_print("[gpu case] ball on soft block: rigid-feature contacts on the block's faces")
    s, ball = R.ball_on_soft_block()
    g, o = GPUSolver(s), oracle.OracleSolver(s)
    first = {}
    for k in range(90):
        g.step()
        o.step()
        for name, n in (("gpu", g.stats().num_soft_rt), ("oracle", o.clothStats()["num_soft_rt"])):
            if n > 0 and name not in first:
                first[name] = k
    flush_print(f"  first rigid-feature contact frame {first}, ball z {g.positions()[ball, 2]:.3f}")
    assert "gpu" in first and first.get("gpu") == first.get("oracle")


def test_surface_reach_skip_keeps_pairs_and_counters():
    # two stacked blocks and a third far away: the queries skip the far
    # block's producers; the stream and the frame's counters equal the
    # oracle's (which queries every producer) on every step
    flush_print("[gpu case] stacked blocks + a distant block: per-surface reach skip")
    s, top = R.stacked_soft_blocks()
    R.add_soft_lattice(s, (6.0, 0, 0.7), 3, 3, 3, 0.34, mu=4000.0, lam=40000.0)
    g, o = GPUSolver(s), oracle.OracleSolver(s)
    for k in range(150):
        g.step()
        o.step()
        if k % 10 == 9 or k > 110:
            pg, po = g.datPairs(), o.datPairs()
            assert np.array_equal(pg, po), k
            sg, so = g.stats(), o.clothStats()
            assert (sg.dat_pairs, sg.dat_vt_pairs, sg.dat_nonfinite) == \
                (so["dat_pairs"], so["dat_vt_pairs"], so["nonfinite"]), k
    flush_print(f"  pairs at the end: {len(g.datPairs())}")
    assert len(g.datPairs()) > 0


def test_cloth_over_block_parity():
    # a membrane cloth, a soft block and rigid bodies in one solve
    # (Demos.clothcombo): collision surfaces bit-exact, the Planar-DAT
    # stream of the mixed scene equal to the oracle's while the cloth lands
    # on the block, 40-step RMS against FP32 / FP64
    flush_print("[gpu case] cloth over a soft block with rigid bodies: topology, pair stream, 40-step RMS")
    s = R.clothcombo(16)
    g, o, o64 = GPUSolver(s), oracle.OracleSolver(s), oracle.OracleSolver(s, precision=64)
    assert np.array_equal(g.collisionTris(), o.collisionTris())
    cg, co = g.clothInfo(), o.clothInfo()
    for k in INFO_KEYS:
        assert cg[k] == co[k], k
    streams = 0
    for k in range(40):
        g.step()
        o.step()
        o64.step()
        if k < 12:
            pg, po = g.datPairs(), o.datPairs()
            assert np.array_equal(pg, po), k
            streams += len(pg)
    pg, po, p64 = g.positions(), o.positions(), o64.bodyStates64()[0]
    rms = float(np.sqrt(np.mean(np.sum((pg - po) ** 2, axis=1))))
    e64 = float(np.sqrt(np.mean(np.sum((po - p64) ** 2, axis=1))))
    flush_print(f"  pairs over 12 steps {streams}, RMS gpu-oracle32 {rms:.3e} m, oracle32-oracle64 {e64:.3e} m")
    assert np.all(np.isfinite(pg)) and rms < max(2e-3, 3.0 * e64)
