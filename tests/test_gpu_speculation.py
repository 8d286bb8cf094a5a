"""GPU tests of the speculation snapshots (GPUSolver.swift:2948-3057) through
the C ABI: the reference's replay contract
(HumanoidBoxCarryTests.testSpeculationSnapshotReplaysNextDynamicsExactly,
Tests/AVBDTests/HumanoidBoxCarryTests.swift:136-215: capture, step with a
schedule of actuator commands, restore, replay the schedule -> the same
states within 1e-6) on batched stacks and torque-driven chains, held here to
bit equality; a diverging branch rewound; in-place re-capture; frame index
and stats restored; odd and even frame distances; the rigid-only
precondition."""
import numpy as np
import pytest

from avbd_amd import GPUSolver, demos
from avbd_amd._capi import AVBDError
from conftest import flush_print

pytestmark = pytest.mark.gpu


def _case(name):
    flush_print(f"[gpu case] {name}")


def _scene():
    s, _ = demos.batched_scene(num_scenes=24, pitch=3.0, iterations=10, seed_base=11 << 32)
    return s


def _torques(g, step, scale=1.0):
    nj = len(g.scene.joints)
    j = np.arange(nj, dtype=np.uint32)
    t = scale * (2.0 + np.abs(np.sin(np.arange(nj) * 0.7 + step * 0.9)) * 3.0)
    g.setMotorTorques(j, t.astype(np.float32))


def _run(g, steps, start, scale=1.0):
    for k in range(start, start + steps):
        _torques(g, k, scale)
        g.step()
    g.synchronize()
    return [a.copy() for a in g.bodyStates()]


@pytest.mark.parametrize("use_graph", [True, False])
@pytest.mark.parametrize("ahead", [3, 4])
def test_snapshot_replays_next_dynamics_exactly(use_graph, ahead):
    _case(f"speculation snapshot replay: 24 batched scenes, graph={use_graph}, {ahead} steps ahead")
    g = GPUSolver(_scene(), useGraph=use_graph)
    _run(g, 5, 0)
    snap = g.captureRigidSpeculationSnapshot()
    frame0 = g.stats().frame_index
    want = _run(g, ahead, 5)
    want_stats = g.stats()
    g.restoreRigidSpeculationSnapshot(snap)
    assert g.stats().frame_index == frame0
    got = _run(g, ahead, 5)
    for a, b in zip(want, got):
        assert np.array_equal(a, b)
    st = g.stats()
    assert st.frame_index == want_stats.frame_index
    assert st.num_pairs == want_stats.num_pairs and st.num_active_manifolds == want_stats.num_active_manifolds


def test_diverging_branch_rewinds():
    _case("speculation snapshot: a branch with other torques is rewound, the original replays")
    g = GPUSolver(_scene())
    _run(g, 6, 0)
    snap = g.captureRigidSpeculationSnapshot()
    want = _run(g, 5, 6)
    g.restoreRigidSpeculationSnapshot(snap)
    other = _run(g, 3, 6, scale=2.5)           # different actuator commands
    assert not np.array_equal(other[0], want[0])
    g.restoreRigidSpeculationSnapshot(snap)
    got = _run(g, 5, 6)
    for a, b in zip(want, got):
        assert np.array_equal(a, b)


def test_recapture_in_place_and_reuse():
    _case("speculation snapshot: in-place recapture, the same snapshot restored twice")
    g = GPUSolver(_scene())
    _run(g, 2, 0)
    snap = g.captureRigidSpeculationSnapshot()
    _run(g, 3, 2)
    assert g.captureRigidSpeculationSnapshot(into=snap) is snap   # now at frame 5
    want = _run(g, 4, 5)
    for _ in range(2):
        g.restoreRigidSpeculationSnapshot(snap)
        got = _run(g, 4, 5)
        for a, b in zip(want, got):
            assert np.array_equal(a, b)


def test_snapshot_is_bound_to_rigid_scenes_and_its_solver():
    _case("speculation snapshot: soft scenes unsupported, foreign snapshot rejected")
    soft = GPUSolver(demos.soft_block_scene(cells=3))
    with pytest.raises(AVBDError):
        soft.captureRigidSpeculationSnapshot()
    a, b = GPUSolver(_scene()), GPUSolver(_scene())
    snap = a.captureRigidSpeculationSnapshot()
    with pytest.raises(AVBDError):
        b.restoreRigidSpeculationSnapshot(snap)


def _manifold_bytes(g):
    recs = g.manifolds()
    return np.asarray(recs).tobytes() if not isinstance(recs, bytes) else recs


@pytest.mark.parametrize("odd", [True, False])
def test_rejected_restore_leaves_the_solver_unchanged(odd):
    """A restore that is refused (a warm start attached since the capture
    changes the snapshot's size) leaves the manifold buffer roles, the
    manifolds and the next frames exactly as a solver that never tried it,
    for odd and even frame distances."""
    from avbd_amd import WarmStartMLP
    _case(f"speculation snapshot: rejected restore is side-effect free (odd distance={odd})")
    m = WarmStartMLP(seed=3, outputs=6)
    a, b = GPUSolver(_scene()), GPUSolver(_scene())
    for g in (a, b):
        _run(g, 3, 0)
    snap = a.captureRigidSpeculationSnapshot()
    n = 3 if odd else 4
    for g in (a, b):
        _run(g, n, 3)
        g.setWarmStart(rigid=m, learned=False, log=True)
    with pytest.raises(AVBDError):
        a.restoreRigidSpeculationSnapshot(snap)
    assert _manifold_bytes(a) == _manifold_bytes(b)
    sa, sb = _run(a, 3, 3 + n), _run(b, 3, 3 + n)
    for x, y in zip(sa, sb):
        assert np.array_equal(x, y)
    assert _manifold_bytes(a) == _manifold_bytes(b)
    snap.close()
    a.close()
    b.close()
    m.close()
