"""GPU parity of the narrow phase and the two CAS persistence maps, through
the C ABI.

Narrow phase (30_narrowphase.metal:189-247, 337-1322): from identical
states taken along an oracle trajectory, the backend's pair stream, manifold
headers (bodies, contact counts) and every contact's feature key equal the
oracle's bit for bit, and normals agree within 1e-5, on the 10-cube stack,
a box pile (face-face, face-edge and edge-edge contacts), the mixed capsule /
sphere / box pile, a convex-hull wedge and 64 of the batched uneven scenes.

Maps (20_broadphase.metal:547-608, 25_cloth.metal:79-137): every collider
pair of an active manifold is found at its own manifold index, pairs of
inactive manifolds and pairs never inserted return -1, and the lookup equals
the oracle's map on the same frame; every element-contact key is found at a
record carrying that key and keys never inserted return -1."""
import numpy as np
import pytest

import oracle
import capsule_fixtures as F
from avbd_amd import GPUSolver, demos
from conftest import flush_print
from test_gpu_cloth import _two_sheets

pytestmark = pytest.mark.gpu

SCENES = {
    "stack": lambda: demos.stack_scene(),
    "box_pile": F.box_pile,
    "capsule_pile": F.pile,
    "hull_wedge": lambda: F.hull_wedge()[0],
    "batched_64": lambda: demos.batched_scene(64)[0],
}


def _same_state(scene, state):
    g = GPUSolver(scene)
    o = oracle.OracleSolver(scene)
    p, r, v, w = state
    idx = np.arange(len(scene.bodies), dtype=np.uint32)
    g.setBodyStates(idx, p, r, v, w)
    o.setBodyStates(idx, p, r, v, w)
    return g, o


def _gpu_keys(gm):
    keys = gm["c"]["rA"][:, :, 3].copy().view(np.uint32)
    counts = gm["header"][:, 2]
    keys[np.arange(8)[None, :] >= counts[:, None]] = 0xFFFFFFFF
    return keys


@pytest.mark.parametrize("name", list(SCENES))
def test_counts_and_feature_keys_bit_exact(name):
    flush_print(f"[gpu case] narrow phase counts / feature keys bit-exact: {name}")
    s = SCENES[name]()
    src = oracle.OracleSolver(s)
    done, contacts, kinds = 0, 0, set()
    for frame in (1, 25, 60, 100):
        src.step(frame - done)
        done = frame
        g, o = _same_state(s, src.bodyStates())
        g.step()
        o.step()
        assert np.array_equal(g.pairs(), o.pairs()), frame
        gm = g.manifolds()
        h, nrm, _ = o.manifolds()
        assert np.array_equal(gm["header"][:, :3], h), frame
        keys = o.manifoldKeys()
        assert np.array_equal(_gpu_keys(gm), keys), frame
        act = h[:, 2] > 0
        np.testing.assert_allclose(gm["basisN"][act, :3], nrm[act], atol=1e-5)
        contacts += int(h[:, 2].sum())
        kinds |= {int(k) >> 24 for k in keys[keys != 0xFFFFFFFF]}
        g.close()
    flush_print(f"  {contacts} contacts over 4 frames, feature kinds {sorted(kinds)}")
    assert contacts > 0
    if name == "box_pile":
        assert {0, 1, 2} <= kinds        # reference faces of A and of B, and edge-edge contacts


def test_pair_map_lookup():
    flush_print("[gpu case] pair persistence map: found at own index, absent -> -1, equals oracle")
    s = F.box_pile()
    src = oracle.OracleSolver(s)
    src.step(60)
    g, o = _same_state(s, src.bodyStates())
    g.step()
    o.step()
    gm = g.manifolds()
    pairs = np.ascontiguousarray(gm["colliderPair"][:, :2])
    act = gm["header"][:, 2] > 0
    assert act.sum() > 5 and (~act).sum() > 0
    found = g.mapFind(pairs)
    assert np.array_equal(found[act], np.nonzero(act)[0])
    assert np.all(found[~act] == -1)
    swapped = pairs[:, ::-1].copy()
    present = {tuple(p) for p in pairs.tolist()}
    swapped = swapped[[tuple(p) not in present for p in swapped.tolist()]]
    far = np.array([[len(s.bodies) + 5, 0], [0, 0], [7, 7]], np.uint32)
    absent = np.concatenate([swapped, far]).astype(np.uint32)
    assert np.all(g.mapFind(absent) == -1)
    assert np.array_equal(found, o.mapFind(pairs))
    assert np.array_equal(g.mapFind(absent), o.mapFind(absent))
    # the invariants hold on the backend's own later frames
    for _ in range(30):
        g.step()
        gm = g.manifolds()
        pairs = np.ascontiguousarray(gm["colliderPair"][:, :2])
        act = gm["header"][:, 2] > 0
        found = g.mapFind(pairs)
        assert np.array_equal(found[act], np.nonzero(act)[0])
        assert np.all(found[~act] == -1)
    flush_print(f"  last frame: {len(pairs)} pairs, {int(act.sum())} in the map")


def test_soft_map_lookup():
    flush_print("[gpu case] element-contact persistence map: every key found, absent -> -1")
    s = _two_sheets()
    g = GPUSolver(s)
    checked = 0
    for frame in range(60):
        g.step()
        rec = g.softContacts()
        if len(rec) == 0:
            continue
        keys = np.stack([rec["lambda"][:, 3].view(np.uint32), rec["penalty"][:, 3].view(np.uint32)], 1)
        found = g.softMapFind(keys)
        assert np.all(found >= 0) and np.all(found < len(rec)), frame
        assert np.array_equal(keys[found], keys), frame
        if len({tuple(k) for k in keys.tolist()}) == len(keys):
            assert np.array_equal(found, np.arange(len(rec))), frame
        present = {tuple(k) for k in keys.tolist()}
        cand = np.concatenate([keys[:, ::-1], keys + np.array([0, 1], np.uint32),
                               np.array([[0xFFFFFFF0, 3]], np.uint32)]).astype(np.uint32)
        absent = cand[[tuple(k) not in present for k in cand.tolist()]]
        assert np.all(g.softMapFind(absent) == -1), frame
        checked += 1
    flush_print(f"  {checked} frames checked, last frame {len(rec)} element contacts")
    assert checked > 10
