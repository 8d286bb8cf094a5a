"""GPU parity of the tet collision surfaces (GPUSolver.swift:384-390,
872-897, 1289-1334, 1476-1522) and of the isotropic OGC limiter they select
(GPUSolver.swift:93-99, 25_cloth.metal:350-940, 40_solver.metal:697-717,
1361-1382) through the C ABI against the oracle: the collision triangles
(authored, then tet boundary faces), edges, surface particles, mode and
derived parameters bit-exact on one block, two blocks, the benchmark block
and a cloth sheet beside a block; frames in lockstep from identical states
(two blocks, two sheets beside a block, three blocks one of them apart,
a cloth over a block among rigid bodies) with the tracked triangle and edge
sets, the element contacts' participants, keys and flags, and the
clearance bounds bit-exact; a 60-step trajectory against the FP32 and FP64
oracle; and a ball dropped onto a block, whose rigid-feature contacts reach
the block's faces (SoftBodyTests.swift:70-83).  The reference's soft-body
contracts themselves run in test_reference_contracts.py."""
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
             "pair_stream", "dat_on", "iso_on")


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


def _three_blocks():
    s, _ = R.stacked_soft_blocks()
    R.add_soft_lattice(s, (6.0, 0, 0.7), 3, 3, 3, 0.34, mu=4000.0, lam=40000.0)
    return s


def _lockstep(s, frames, lanes=8, check_from=0):
    """Both backends set to the oracle's state before every frame, then
    stepped once: everything the trackers read (poses, velocities, the
    previous tracked sets and contact sides) is identical, so their integer
    outputs must be.  The bounds are compared where neither backend
    refreshed them mid-frame (a refresh reads the iterate, which agrees
    within float rounding only).  The poses are held to the spread between
    the FP32 and FP64 oracle over the same frames: a stiff scene amplifies
    rounding within one frame."""
    g, o = GPUSolver(s, lanesPerBody=lanes, sweepMode=1), oracle.OracleSolver(s)
    o64 = oracle.OracleSolver(s, precision=64)
    assert g.clothInfo()["iso_on"] == 1
    idx = np.arange(len(s.bodies), dtype=np.uint32)
    worst, spread, nsoft, nvt, nee, bounded, trunc = 0.0, 0.0, 0, 0, 0, 0, [0, 0]
    for k in range(frames):
        p, r, v, w = o.bodyStates()
        g.setBodyStates(idx, p, r, v, w)
        o.setBodyStates(idx, p, r, v, w)
        o64.setBodyStates(idx, p, r, v, w)
        g.step()
        o.step()
        o64.step()
        if k < check_from:
            continue
        ig, io = g.isoState(), o.isoState()
        assert np.array_equal(ig["vt_track"], io["vt_track"]), k
        assert np.array_equal(ig["ee_track"], io["ee_track"]), k
        if ig["refreshes"] == 0 and io["refreshes"] == 0:
            assert np.array_equal(ig["bound"].view(np.uint32), io["bound"].view(np.uint32)), k
            bounded += int((io["bound"] < 1e37).sum())
        trunc[0] += ig["truncations"]
        trunc[1] += io["truncations"]
        sg, so = g.softContacts(), o.softContacts()
        assert len(sg) == len(so), k
        assert np.array_equal(sg["ids"], so["ids"]), k
        for f in ("lambda", "penalty", "anchorA"):
            assert np.array_equal(sg[f][:, 3].view(np.uint32), so[f][:, 3].view(np.uint32)), (k, f)
        kinds = so["lambda"][:, 3].view(np.uint32) >> 28
        nsoft += len(so)
        nvt += int((kinds == 1).sum())
        nee += int((kinds == 3).sum())
        worst = max(worst, float(np.abs(g.positions() - o.positions()).max()))
        spread = max(spread, float(np.abs(o.positions() - o64.bodyStates64()[0]).max()))
    return worst, spread, nsoft, nvt, nee, bounded, trunc


@pytest.mark.parametrize("name", ["stacked", "sheets", "three-blocks", "clothcombo"])
def test_iso_lockstep_bitexact(name):
    flush_print(f"[gpu case] isotropic limiter in lockstep: {name}")
    if name == "stacked":
        s, _ = R.stacked_soft_blocks()
        frames, start = 140, 0
    elif name == "sheets":
        s, frames, start = R.sheets_and_block(), 30, 0
    elif name == "three-blocks":
        s, frames, start = _three_blocks(), 140, 0
    else:
        s, frames, start = R.clothcombo(16), 40, 0
    worst, spread, nsoft, nvt, nee, bounded, trunc = _lockstep(s, frames, check_from=start)
    flush_print(f"  {frames} frames: contacts {nsoft} (V-T {nvt}, E-E {nee}), bounded vertex-frames {bounded}, "
                f"truncations gpu {trunc[0]} oracle {trunc[1]}, worst max|x_gpu - x_oracle| {worst:.3e} m, "
                f"oracle FP32 - FP64 {spread:.3e} m")
    assert nvt > 0 and bounded > 0
    if name in ("sheets", "clothcombo"):
        assert nee > 0
    assert worst <= max(1e-4, 3.0 * spread)
    assert abs(trunc[0] - trunc[1]) <= max(3, 0.05 * max(trunc))


def test_iso_truncation_holds_on_gpu():
    """After every frame each bounded surface particle lies within
    max(0.45 sqrt(bound), 0.2 r) of the pose its bound holds around."""
    flush_print("[gpu case] isotropic truncation invariant, two sheets beside a block")
    s = R.sheets_and_block()
    g = GPUSolver(s)
    parts = g.clothInfo()["particles"]
    rad = np.array([abs(b["size"][0]) / 2 for b in s.bodies])
    held = 0
    for k in range(40):
        g.step()
        st = g.isoState()
        b = st["bound"].astype(np.float64)[parts]
        fin = b < 1e37
        bv = np.maximum(0.45 * np.sqrt(np.where(fin, b, 0.0)), 0.2 * rad[parts])
        moved = np.linalg.norm(g.positions()[parts].astype(np.float64) - st["anchor"][parts], axis=1)
        assert np.all(moved[fin] <= bv[fin] * (1 + 1e-5) + 1e-6), k
        held += st["truncations"]
    flush_print(f"  truncations over 40 frames: {held}")


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


def test_cloth_over_block_parity():
    # a membrane cloth, a soft block and rigid bodies in one solve
    # (Demos.clothcombo): collision surfaces bit-exact, the first frame's
    # tracked sets (isotropic mode) equal to the oracle's -- later frames
    # start from poses that agree within rounding, where near-equal distances
    # may rank differently (the lockstep test holds them bit-exact from
    # identical states) -- and the 40-step RMS against FP32 / FP64
    flush_print("[gpu case] cloth over a soft block with rigid bodies: topology, tracked sets, 40-step RMS")
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
        if k == 0:
            ig, io = g.isoState(), o.isoState()
            assert np.array_equal(ig["vt_track"], io["vt_track"])
            assert np.array_equal(ig["ee_track"], io["ee_track"])
        if k < 12:
            streams += int((o.isoState()["vt_track"] != 0xFFFFFFFF).sum())
    pg, po, p64 = g.positions(), o.positions(), o64.bodyStates64()[0]
    rms = float(np.sqrt(np.mean(np.sum((pg - po) ** 2, axis=1))))
    e64 = float(np.sqrt(np.mean(np.sum((po - p64) ** 2, axis=1))))
    flush_print(f"  tracked triangles over 12 steps {streams}, RMS gpu-oracle32 {rms:.3e} m, oracle32-oracle64 {e64:.3e} m")
    assert np.all(np.isfinite(pg)) and rms < max(2e-3, 3.0 * e64)
