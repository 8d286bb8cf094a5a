"""GPU parity of the cloth path: the HIP backend (through the C ABI) against
the oracle on cloth surfaces -- StVK membranes, Bergou hinges, soft and hard
structural springs, the Planar-DAT V-T / E-E pair stream, OGC element contacts
(V-T, E-E, rigid feature vs triangle) and the 1-ring viscosity.

Bars: topology, static coloring and pair streams bit-exact; one iteration
from an identical cold state within 1e-5 m; multi-step trajectories within a
stated RMS, with the FP64 oracle's error reported beside it; bitwise
self-determinism; the reference's own drape gate (ClothTests.swift:205-219).
"""
import numpy as np
import pytest

import oracle
from avbd_amd import GPUSolver, RuntimeFailure, demos
from avbd_amd.scene import PhysicsScene, SimSettings

from conftest import flush_print

pytestmark = pytest.mark.gpu


def _case(name):
    flush_print(f"[gpu case] {name}")


def _small_drape(n=24, iterations=20):
    return demos.cloth_scene(n=n, size=0.4, height=0.25, radius=0.004, sphere_radius=0.1, sphere_z=0.0,
                             iterations=iterations)


def _two_sheets(gap=0.009, n=10):
    s = PhysicsScene("two sheets", SimSettings(iterations=20, particleDamping=1.5, clothViscosity=0.25))
    s.addBody((200, 200, 2), 0.0, 0.6, (0, 0, -1))
    rng = demos.SplitMix64(5)
    size, r = 0.12, 0.003
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


def test_cloth_topology_and_coloring_bitexact():
    _case("cloth topology + static palette, 64^2 sheet")
    s = demos.cloth_scene(n=64, size=0.5, height=0.3, radius=0.002, sphere_radius=0.1)
    g = GPUSolver(s)
    o = oracle.OracleSolver(s)
    cg, co = g.clothInfo(), o.clothInfo()
    for k in ("num_tris", "num_edges", "num_particles", "num_membranes", "num_bends", "max_soft", "max_dat_pairs",
              "elem_cell_size", "elem_margin", "dat_query_radius", "surface_contact_cell_size"):
        assert cg[k] == co[k], k
    assert np.array_equal(cg["edges"], co["edges"])
    assert np.array_equal(cg["particles"], co["particles"])
    assert np.array_equal(g.debugColors(), o.debugColors())
    assert g.stats().num_colors == 2   # the structural grid is bipartite


def _cold_pair(scene, state, **kw):
    """Fresh backend and oracle of `scene` set to the same state (no warm
    starts anywhere)."""
    g = GPUSolver(scene, **kw)
    o = oracle.OracleSolver(scene)
    p, r, v, w = state
    idx = np.arange(len(scene.bodies), dtype=np.uint32)
    g.setBodyStates(idx, p, r, v, w)
    o.setBodyStates(idx, p, r, v, w)
    return g, o


@pytest.mark.parametrize("lanes", [4, 8, 16, 32])
def test_cloth_one_iteration_parity_from_cold_state(lanes):
    _case(f"cloth: pairs + contacts + one iteration from an identical cold state, lanes={lanes}")
    s = _two_sheets()
    src = oracle.OracleSolver(s)
    src.step(45)                       # the upper sheet has settled onto the lower one
    state = src.bodyStates()
    g, o = _cold_pair(s, state, lanesPerBody=lanes, sweepMode=1)
    st = g.settings
    st.iterations = 1
    g.settings = st
    o.set_settings(st)
    g.step()
    o.step()
    pg, po = g.datPairs(), o.datPairs()
    flush_print(f"  pairs gpu {len(pg)} oracle {len(po)}, soft gpu {g.stats().num_soft}")
    assert len(po) > 0
    assert np.array_equal(pg, po)
    sg, so = g.softContacts(), o.softContacts()
    assert len(sg) == len(so) and len(sg) > 0
    assert np.array_equal(sg["ids"], so["ids"])
    assert np.array_equal(sg["lambda"][:, 3].view(np.uint32), so["lambda"][:, 3].view(np.uint32))   # keys
    assert np.array_equal(sg["penalty"][:, 3].view(np.uint32), so["penalty"][:, 3].view(np.uint32))
    assert np.array_equal(sg["anchorA"][:, 3].view(np.uint32), so["anchorA"][:, 3].view(np.uint32))  # flags
    np.testing.assert_allclose(sg["normal"], so["normal"], atol=1e-6)
    np.testing.assert_allclose(sg["weights"], so["weights"], atol=1e-6)
    np.testing.assert_allclose(sg["C0"], so["C0"], rtol=1e-5, atol=1e-7)
    err = np.abs(g.positions() - o.positions()).max()
    flush_print(f"  max|x_gpu - x_oracle| = {err:.3e} m")
    assert err <= 1e-5


def test_cloth_drape_one_iteration_parity():
    _case("drape over sphere: rigid-feature contacts + one iteration from an identical cold state")
    s = _small_drape()
    src = oracle.OracleSolver(s)
    src.step(40)
    g, o = _cold_pair(s, src.bodyStates(), lanesPerBody=8, sweepMode=1)
    st = g.settings
    st.iterations = 1
    g.settings = st
    o.set_settings(st)
    g.step()
    o.step()
    assert g.stats().num_soft_rt == o.clothStats()["num_soft_rt"] > 0
    assert np.array_equal(g.datPairs(), o.datPairs())
    assert np.array_equal(g.pairs(), o.pairs())
    err = np.abs(g.positions() - o.positions()).max()
    flush_print(f"  max|x_gpu - x_oracle| = {err:.3e} m")
    assert err <= 1e-5


def test_cloth_trajectory_rms_60_steps():
    _case("cloth drape 24^2, 60-step trajectory vs oracle FP32 / FP64")
    s = _small_drape()
    g = GPUSolver(s)
    o = oracle.OracleSolver(s)
    o64 = oracle.OracleSolver(s, precision=64)
    for _ in range(60):
        g.step()
        o.step()
        o64.step()
    pg, po = g.positions(), o.positions()
    p64 = o64.bodyStates64()[0]
    rms = float(np.sqrt(np.mean(np.sum((pg - po) ** 2, axis=1))))
    e64g = float(np.sqrt(np.mean(np.sum((pg - p64) ** 2, axis=1))))
    e64o = float(np.sqrt(np.mean(np.sum((po - p64) ** 2, axis=1))))
    flush_print(f"  RMS gpu-oracle32 {rms:.3e} m, gpu-oracle64 {e64g:.3e} m, oracle32-oracle64 {e64o:.3e} m")
    assert np.all(np.isfinite(pg))
    # the sheet is 0.4 m wide with 4 mm particles; chaotic contact branching
    # bounds the agreement, so the bar is relative to the FP32-vs-FP64 spread
    assert rms < max(2e-3, 3.0 * e64o)


def test_cloth_bitwise_determinism():
    _case("cloth: bitwise self-determinism, 40 steps")
    s = _two_sheets()
    a = GPUSolver(s, lanesPerBody=8, sweepMode=1)
    b = GPUSolver(s, lanesPerBody=8, sweepMode=1)
    a.step(40)
    b.step(40)
    for x, y in zip(a.bodyStates(), b.bodyStates()):
        assert np.array_equal(x, y)
    assert np.array_equal(a.datPairs(), b.datPairs())


def test_cloth_graph_vs_eager_and_sweep_modes():
    _case("cloth: hipGraph replay vs eager, per-color vs persistent sweep")
    s = _two_sheets()
    a = GPUSolver(s, useGraph=True, lanesPerBody=8, sweepMode=1)
    b = GPUSolver(s, useGraph=False, lanesPerBody=8, sweepMode=1)
    a.step(20)
    b.step(20)
    assert np.array_equal(a.positions(), b.positions())
    c = GPUSolver(s, lanesPerBody=8, sweepMode=2)
    if c.stats().sweep_mode == 2:
        c.step(20)
        assert np.array_equal(a.positions(), c.positions())


def test_drape_settles_gpu():
    # ClothTests.testDrapeSettles (ClothTests.swift:205-219) on the backend
    _case("reference drape gate, res 24, 720 frames")
    s = demos.drape_scene(24)
    g = GPUSolver(s)
    g.step(720)
    p, _, v, _ = g.bodyStates()
    m = g.bodyMass().astype(np.float64)
    part = np.array([b["particle"] for b in s.bodies])
    ke = 0.5 * float(np.sum((m * np.sum(v.astype(np.float64) ** 2, axis=1))[part]))
    gap, stretch = demos.cloth_metrics(s, p)
    flush_print(f"  KE {ke:.3e} J, gap {gap:.4f} m, stretch {stretch:.4f}")
    assert ke < 0.8
    assert gap > -0.05
    assert stretch < 0.15


def test_soft_capacity_and_pair_capacity_latch():
    _case("cloth: soft-contact and pair-stream capacity failures latch")
    s = _two_sheets()
    g = GPUSolver(s, planarDATSoftCapacityForTesting=3)
    with pytest.raises(RuntimeFailure) as e:
        for _ in range(60):
            g.step()
        g.synchronize()
    assert e.value.kind == "softContactCapacity"
    g2 = GPUSolver(s, planarDATPairCapacityForTesting=2)
    with pytest.raises(RuntimeFailure) as e2:
        for _ in range(60):
            g2.step()
        g2.synchronize()
    assert e2.value.kind == "planarDATPairCapacity"


def test_cloth_256_runs():
    _case("256^2 cloth drape (the benchmark workload), 20 frames")
    s = demos.cloth_scene()
    g = GPUSolver(s)
    g.step(20)
    g.synchronize()
    p = g.positions()
    st = g.stats()
    flush_print(f"  colors {st.num_colors}, pairs {st.num_pairs}, soft {st.num_soft} (rt {st.num_soft_rt}), "
                f"dat {st.dat_pairs}, z range {p[2:, 2].min():.3f}..{p[2:, 2].max():.3f}")
    assert np.all(np.isfinite(p))
    assert st.num_colors == 2
