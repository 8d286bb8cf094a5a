"""GPU parity: the HIP backend (through the C ABI) against the oracle.

Bars: integer outputs (colorings, hash cell lists, pair streams, manifold
activity) bit-exact; FP32 state after one iteration from an identical state
within 1e-5 relative (positions are O(1 m), so 1e-5 m absolute); 60-step
trajectories within an RMS of 2e-3 m; bitwise self-determinism.
"""
import numpy as np
import pytest

import oracle
from oracle import mlp as omlp
from avbd_amd import AVBDError, GPUSolver, RuntimeFailure, demos
from avbd_amd.scene import PhysicsScene, SceneJoint, SimSettings

from conftest import flush_print

pytestmark = pytest.mark.gpu


def _case(name):
    flush_print(f"[gpu case] {name}")


def _block(cells=20, iterations=20):
    return demos.soft_block_scene(cells=cells, iterations=iterations)


def test_static_coloring_bitexact_c2():
    _case("static coloring, 21^3 soft block")
    s = _block()
    g = GPUSolver(s)
    o = oracle.OracleSolver(s)
    cg, co = g.debugColors(), o.debugColors()
    assert np.array_equal(cg, co)
    st = g.stats()
    assert st.num_colors == int(co[1:].max()) + 1
    assert sum(st.color_counts) == s.numDynamic


def _cloud(n=300, seed=11):
    rng = demos.SplitMix64(seed)
    s = demos.PhysicsScene("cloud", SimSettings(iterations=4))
    demos.add_ground(s)
    for i in range(n):
        p = (rng.uniform(-3, 3), rng.uniform(-3, 3), rng.uniform(0.2, 4))
        if i % 3 == 0:
            s.addSphere(rng.uniform(0.2, 0.6), 1.0, 0.5, p)
        else:
            s.addBody((rng.uniform(0.2, 0.8), rng.uniform(0.2, 0.8), rng.uniform(0.2, 0.8)), 1.0, 0.5, p)
    return s


def _granular(n=1500, seed=12):
    """Overlapping free particles (r = 1 cm) with small boxes and spheres:
    hash cells (0.1 m, the radius floor) hold dozens of colliders, so pairs
    come from the candidate search grid."""
    rng = demos.SplitMix64(seed)
    s = demos.PhysicsScene("granular", SimSettings(iterations=2))
    demos.add_ground(s)
    for i in range(n):
        p = (rng.uniform(-0.12, 0.12), rng.uniform(-0.12, 0.12), rng.uniform(0.0, 0.4))
        s.addParticle(0.01, 0.001, p)
    for i in range(50):
        p = (rng.uniform(-0.12, 0.12), rng.uniform(-0.12, 0.12), rng.uniform(0.0, 0.4))
        if i % 2:
            s.addSphere(0.03, 500.0, 0.5, p)
        else:
            s.addBody((0.02, 0.02, 0.02), 500.0, 0.5, p)
    return s


def _block_with_rigid(seed=13):
    """A 9^3 soft block with small hashed rigid spheres and cubes resting on
    and beside it: some tet particles have a rigid collider in a neighbour
    cell (pair search on the wave path), the rest only the ground."""
    s = _block(8, 5)
    rng = demos.SplitMix64(seed)
    for i in range(12):
        p = (rng.uniform(-0.6, 0.6), rng.uniform(-0.6, 0.6), rng.uniform(1.45, 1.6))
        if i % 2:
            s.addSphere(0.1, 500.0, 0.5, p)
        else:
            s.addBody((0.08, 0.08, 0.08), 500.0, 0.5, p)
    return s


def _granular_large(n=40000, seed=14):
    """40000 free particles (r = 1 cm) in a 0.6 m cube, about six overlaps
    each, with 60 small rigid bodies: enough producers for the one-lane
    search-grid walk, and a tail of producers with more than eight accepted
    partners that it hands to the one-wave walk."""
    rng = demos.SplitMix64(seed)
    s = demos.PhysicsScene("granular large", SimSettings(iterations=2))
    demos.add_ground(s)
    for i in range(n):
        s.addParticle(0.01, 0.001, (rng.uniform(-0.3, 0.3), rng.uniform(-0.3, 0.3), rng.uniform(0.0, 0.6)))
    for i in range(60):
        p = (rng.uniform(-0.3, 0.3), rng.uniform(-0.3, 0.3), rng.uniform(0.0, 0.6))
        if i % 2:
            s.addSphere(0.03, 500.0, 0.5, p)
        else:
            s.addBody((0.02, 0.02, 0.02), 500.0, 0.5, p)
    return s


def test_hash_cells_and_pairs_bitexact():
    for name, s in [("stack 10", demos.stack_scene(10)), ("soft block 9^3", _block(8, 5)),
                    ("soft block 9^3 + small rigid bodies", _block_with_rigid()),
                    ("soft block 21^3", _block(20, 2)), ("rigid cloud 300", _cloud()),
                    ("granular 1500 + 50 small rigid (search grid)", _granular()),
                    ("soft block 41^3 (search grid, 64-collider buckets)", _block(40, 1)),
                    ("granular 40000 + 60 small rigid (search grid walked one lane per producer)", _granular_large())]:
        _case(f"hash + pairs, {name}")
        g = GPUSolver(s)
        o = oracle.OracleSolver(s)
        g.step()
        o.step()
        g.synchronize()
        cg, co = g.cellLists(), o.cellLists()
        assert cg["hash_size"] == co["hash_size"] and cg["cell_size"] == co["cell_size"]
        assert np.array_equal(cg["count"], co["count"])
        assert np.array_equal(cg["start"], co["start"])
        assert np.array_equal(cg["bodies"], co["bodies"])
        assert np.array_equal(g.pairs(), o.pairs())
        # manifold activity of the first frame (identical start state)
        mg = g.manifolds()
        ho, _, _ = o.manifolds()
        assert np.array_equal(mg["header"][:, :3], ho)


def _set_same_state(g, o, n):
    p, r, v, w = g.bodyStates()
    idx = np.arange(n, dtype=np.uint32)
    g.setBodyStates(idx, p, r, v, w)
    o.setBodyStates(idx, p, r, v, w)
    return p


@pytest.mark.parametrize("lanes", [4, 8, 16, 32])
def test_one_iteration_parity_c2(lanes):
    _case(f"one iteration from identical state, 21^3 block, lanes={lanes}")
    s = _block()
    g = GPUSolver(s, lanesPerBody=lanes, sweepMode=1)
    o = oracle.OracleSolver(s)
    g.step(12)  # reach contact with the ground
    p0 = _set_same_state(g, o, len(s.bodies))
    st = g.settings
    st.iterations = 1
    g.settings = st
    o.set_settings(st)
    g.step()
    o.step()
    pg, po = g.positions(), o.positions()
    assert np.array_equal(g.pairs(), o.pairs())
    err = np.abs(pg - po).max()
    dx = np.abs(po - p0).max()
    flush_print(f"  max|x_gpu-x_oracle| = {err:.3e} m, max|dx| = {dx:.3e} m")
    assert err <= 1e-5 * max(1.0, float(np.abs(po).max()))


def test_trajectory_rms_c2_60_steps():
    _case("60-step trajectory, 21^3 block")
    s = _block()
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
    assert rms < 2e-3


@pytest.mark.parametrize("name", ["stack", "block"])
def test_warm_started_one_iteration_frames(name):
    """Warm starts carried frame to frame on both backends (no state reset):
    every frame runs one iteration from the previous frame's duals and
    penalties (40_solver.metal:546-655, 30_narrowphase.metal:1265-1321), and
    the positions agree within 1e-5 m after every frame; the first contact's
    multiplier and penalty agree where both backends keep the manifold."""
    _case(f"warm-started one-iteration frames, {name}")
    s = demos.stack_scene() if name == "stack" else _block(12, 1)
    st = s.settings
    st.iterations = 1
    g = GPUSolver(s)
    o = oracle.OracleSolver(s)
    g.settings = st
    o.set_settings(st)
    worst = 0.0
    for k in range(40):
        g.step()
        o.step()
        err = float(np.abs(g.positions() - o.positions()).max())
        worst = max(worst, err)
        assert err <= 1e-5, (k, err)
    gm = g.manifolds()
    h, _, lp = o.manifolds()
    assert np.array_equal(gm["header"][:, :3], h)
    act = h[:, 2] > 0
    assert act.sum() > 0
    lam_g, pen_g = gm["c"]["lambda"][act, 0, 0], gm["c"]["penalty"][act, 0, 0]
    # a multiplier is penalty x constraint value, so small ones carry the
    # absolute error of the largest penalty times the 1e-5 m agreement of the
    # positions: 1e-3 relative plus 1e-5 of the largest multiplier
    np.testing.assert_allclose(lam_g, lp[act, 0], rtol=1e-3, atol=1e-5 * float(np.abs(lp[act, 0]).max()))
    np.testing.assert_allclose(pen_g, lp[act, 1], rtol=1e-3)
    flush_print(f"  40 frames, worst max|x_gpu - x_oracle| = {worst:.3e} m, {int(act.sum())} manifolds compared")


def test_bitwise_determinism_c2():
    _case("bitwise self-determinism, 21^3 block, 30 steps")
    s = _block()
    a = GPUSolver(s, lanesPerBody=16, sweepMode=1)
    b = GPUSolver(s, lanesPerBody=16, sweepMode=1)
    a.step(30)
    b.step(30)
    for x, y in zip(a.bodyStates(), b.bodyStates()):
        assert np.array_equal(x, y)


def test_sweep_modes_agree():
    _case("per-color launches vs persistent cooperative sweep")
    s = _block(12, 10)
    a = GPUSolver(s, lanesPerBody=8, sweepMode=1)
    b = GPUSolver(s, lanesPerBody=8, sweepMode=2)
    if b.stats().sweep_mode != 2:
        pytest.skip("cooperative launch unavailable")
    a.step(10)
    b.step(10)
    assert np.array_equal(a.positions(), b.positions())


def test_graph_vs_eager_identical():
    _case("hipGraph replay vs eager launches")
    s = _block(8, 10)
    # same lane split (the autotuner may pick different ones per instance)
    a = GPUSolver(s, useGraph=True, lanesPerBody=8, sweepMode=1)
    b = GPUSolver(s, useGraph=False, lanesPerBody=8, sweepMode=1)
    a.step(10)
    b.step(10)
    assert np.array_equal(a.positions(), b.positions())


def test_stack_parity_and_thresholds():
    _case("rigid stack of 10, 240 steps")
    s = demos.stack_scene(10)
    g = GPUSolver(s)
    o = oracle.OracleSolver(s)
    for _ in range(240):
        g.step()
        o.step()
    pg, po = g.positions(), o.positions()
    flush_print(f"  heights gpu {np.round(pg[1:, 2], 3)}")
    # CPU-GPU agreement of the reference's own parity test (GPUSolverTests.swift:831-846)
    assert np.abs(pg - po).max() < 0.05
    # each box sinks by about the 1 cm contact margin: z_i = 0.5 + i - 0.01 (i + 1)
    for i in range(10):
        assert abs(pg[1 + i, 2] - (0.5 + i - 0.01 * (i + 1))) < 0.03
        assert np.hypot(pg[1 + i, 0], pg[1 + i, 1]) < 0.15
    assert g.stats().color_conflicts == 0


def test_five_stack_reference_thresholds():
    _case("5-stack, 240 steps, reference thresholds")
    # CPUSolverTests.swift:80-98: z = 0.5 + i +- 0.08, lateral drift < 0.15
    s = demos.stack_scene(5, friction=0.7, jitter=0.0)
    g = GPUSolver(s)
    g.step(240)
    p = g.positions()
    for i in range(5):
        assert abs(p[1 + i, 2] - (0.5 + i)) < 0.08
        assert np.hypot(p[1 + i, 0], p[1 + i, 1]) < 0.15


def test_box_rests_on_ground():
    _case("single box rests")
    # CPUSolverTests.swift:35-47: z = 0.5 +- 0.03 after 120 steps
    s = demos.stack_scene(1, jitter=0.0)
    g = GPUSolver(s)
    g.step(120)
    assert abs(g.positions()[1, 2] - 0.5) < 0.03


def test_free_fall_implicit_euler():
    _case("free fall")
    s = demos.PhysicsScene("fall", SimSettings(iterations=4))
    s.addBody((1, 1, 1), 1.0, 0.5, (0, 0, 10), velocity=(1.0, 0.0, 2.0))
    g = GPUSolver(s)
    dt = s.settings.dt
    n = 60
    g.step(n)
    p = g.positions()[0].astype(np.float64)
    # implicit Euler: v_k = v0 + k g dt ; x_n = x0 + sum_k v_k dt
    vz = 2.0 + (-10.0) * dt * np.arange(1, n + 1)
    z = 10.0 + np.sum(vz) * dt
    assert abs(p[2] - z) < 1e-3
    assert abs(p[0] - n * dt) < 1e-4


def test_pair_capacity_latches():
    _case("rigid pair capacity fail-closed")
    s = demos.stack_scene(10)
    g = GPUSolver(s, pairCapacityForTesting=4)
    g.step()
    with pytest.raises(RuntimeFailure) as e:
        g.synchronize()
    assert e.value.kind == "rigidPairCapacity" and e.value.capacity == 4
    with pytest.raises(RuntimeFailure):
        g.step()


def test_denied_stage_latches():
    _case("denied stage latches a sticky failure")
    s = demos.stack_scene(3)
    g = GPUSolver(s)
    g.step()
    g.denyStageForTesting("narrowphase", frame=2)
    with pytest.raises(RuntimeFailure) as e:
        g.step()
    assert e.value.kind == "commandEncoderCreation" and e.value.frame == 2
    with pytest.raises(RuntimeFailure):
        g.synchronize()


def test_injected_completion_failure_latches():
    _case("completion failure injected for testing: latched, sticky, earliest wins")
    s = demos.stack_scene(3)
    g = GPUSolver(s)
    g.step(2)
    g.failCompletionForTesting("commandExecution", frame=4)
    with pytest.raises(RuntimeFailure) as e:
        for _ in range(6):
            g.step()
        g.synchronize()
    assert e.value.kind == "commandExecution" and e.value.frame == 4
    with pytest.raises(RuntimeFailure) as e2:
        g.step()
    assert e2.value.frame == 4          # the earliest failure stays


def test_set_body_states_cold_starts_contacts():
    _case("setBodyStates clears incident warm starts")
    s = demos.stack_scene(3)
    g = GPUSolver(s)
    g.step(30)
    m = g.manifolds()
    assert (m["header"][:, 2] > 0).any()
    p, r, v, w = g.bodyStates()
    g.setBodyStates(np.arange(len(p)), p, r, v, w)
    m2 = g.manifolds()
    assert np.all(m2["c"]["lambda"] == 0)


# -- rigid + joints: batched uneven scenes ------------------------------------------
def test_batched_rigid_frame_parity():
    _case("batched 32 scenes: first frame pairs/colors bit-exact and state within 1e-5")
    # The torque-driven chains are chaotic (the oracle's own FP32 and FP64
    # runs part by decimetres within 10 frames), so the motorized batch is
    # compared on its first frame from the identical scene state.
    s, ranges = demos.batched_scene(32)
    g = GPUSolver(s)
    o = oracle.OracleSolver(s)
    g.step()
    o.step()
    assert np.array_equal(g.pairs(), o.pairs())
    assert np.array_equal(g.debugColors(), o.debugColors())
    st = g.stats()
    assert st.uses_dynamic_coloring == 1 and st.color_conflicts == 0
    (pg, rg, _, _), (po, ro, _, _) = g.bodyStates(), o.bodyStates()
    err = max(np.abs(pg - po).max(), np.abs(rg - ro).max())
    flush_print(f"  frame-1 max state difference {err:.3e}")
    assert err < 1e-5 * max(1.0, float(np.abs(po).max()))


def test_batched_passive_trajectory_agreement():
    _case("batched 32 scenes, passive chains: 30-step error vs the FP64 oracle")
    # Toppling stacks and swinging chains amplify rounding: the oracle's own
    # FP32 and FP64 runs part by decimetres in 30 frames.  The GPU result must
    # be as close to the FP64 rerun as the FP32 oracle is.
    s, ranges = demos.batched_scene(32, torque_sigma=0.0)
    g = GPUSolver(s)
    o = oracle.OracleSolver(s)
    o64 = oracle.OracleSolver(s, precision=64)
    for _ in range(30):
        g.step()
        o.step()
        o64.step()
    p64 = o64.bodyStates64()[0]

    def rms(a, b):
        return float(np.sqrt(np.mean(np.sum((a - b) ** 2, axis=1))))

    e_g, e_o = rms(g.positions(), p64), rms(o.positions(), p64)
    flush_print(f"  30-step RMS vs FP64: gpu {e_g:.3e} m, oracle32 {e_o:.3e} m")
    assert np.all(np.isfinite(g.positions()))
    assert e_g <= 3.0 * e_o + 1e-3


def test_replica_isolation_bitwise():
    _case("overlapping replicas in different collision groups evolve bitwise identically")
    # GPUSolverTests.swift:437-468
    s = demos.PhysicsScene("replicas", SimSettings(iterations=10))
    demos.add_ground(s)
    for grp in (1, 2):
        for i in range(4):
            s.addBody((1, 1, 1), 1.0, 0.5, (0.01 * i, 0.0, 0.5 + i), collisionGroup=grp)
    g = GPUSolver(s, lanesPerBody=8, sweepMode=2)
    g.step(120)
    p, r, v, w = g.bodyStates()
    assert np.array_equal(p[1:5], p[5:9]) and np.array_equal(r[1:5], r[5:9])


def test_hinge_pendulum_and_chain_stretch():
    _case("hinge chain holds its links")
    # ConvergenceTests.swift:13-31: chain stretch < 5 %
    s = demos.PhysicsScene("chain", SimSettings(iterations=10))
    demos.add_ground(s)
    prev, L = -1, 0.3
    top = (0.0, 0.0, 8.0)
    for i in range(20):
        dens = 1000.0 if i == 19 else 1.0   # heavy end link: 1000x mass ratio
        b = s.addBody((0.1, 0.1, L), dens, 0.5, (0.0, 0.0, top[2] - L * (i + 0.5)), collisionEnabled=False)
        rA = (0.0, 0.0, -L / 2) if prev >= 0 else top
        s.addJoint(demos.SceneJoint(prev, b, rA, (0.0, 0.0, L / 2), hingeAxis=(1.0, 0.0, 0.0),
                                    stiffnessLin=float("inf"), stiffnessAng=float("inf")))
        prev = b
    # start swinging: tilt by giving the heavy link a sideways velocity
    g = GPUSolver(s)
    st = g.settings
    g.step(120)
    p = g.positions()
    ends = np.vstack([[0.0, 0.0, 8.0], p[1:]])
    length = np.linalg.norm(p[-1] - np.array([0.0, 0.0, 8.0]))
    assert np.all(np.isfinite(p))
    assert length < 20 * L * 1.05
    assert g.maxConstraintError() < 0.05 * L


def test_batched_4096_runs_and_isolates():
    _case("batched 4096 scenes: 30 steps finite, no coloring conflicts")
    s, ranges = demos.batched_scene(4096)
    g = GPUSolver(s)
    g.step(30)
    st = g.stats()
    p = g.positions()
    flush_print(f"  bodies {len(p)}, colors {st.num_colors}, pairs {st.num_pairs}, sweep {st.sweep_mode}, "
                f"lanes {st.lanes_per_body}")
    assert np.all(np.isfinite(p))
    assert st.color_conflicts == 0
    assert p[1:, 2].min() > -0.05


def _motor_chain(gravity=-10.0):
    """Three hinged links: joint 0 unpowered (world hinge), joint 1 a velocity
    motor, joint 2 an implicit position PD motor; a fourth body on a fragile
    ball joint (fracture 1 N, linear multiplier counted)."""
    from avbd_amd.scene import SceneJoint
    s = demos.PhysicsScene("motors", SimSettings(iterations=10, gravity=gravity))
    demos.add_ground(s)
    a = s.addBody((0.1, 0.1, 0.3), 500.0, 0.5, (0, 0, 2.85))
    b = s.addBody((0.1, 0.1, 0.3), 500.0, 0.5, (0, 0, 2.55))
    c = s.addBody((0.1, 0.1, 0.3), 500.0, 0.5, (0, 0, 2.25))
    d = s.addBody((0.1, 0.1, 0.1), 500.0, 0.5, (0, 0, 2.05))
    inf = float("inf")
    s.addJoint(SceneJoint(-1, a, (0, 0, 3.0), (0, 0, 0.15), stiffnessLin=inf, stiffnessAng=inf, hingeAxis=(1, 0, 0)))
    s.addJoint(SceneJoint(a, b, (0, 0, -0.15), (0, 0, 0.15), stiffnessLin=inf, stiffnessAng=inf, hingeAxis=(1, 0, 0),
                          motorTarget=2.0, motorTorque=5.0, motorDamping=0.5, motorMode=2))
    s.addJoint(SceneJoint(b, c, (0, 0, -0.15), (0, 0, 0.15), stiffnessLin=inf, stiffnessAng=inf, hingeAxis=(0, 1, 0),
                          motorTarget=0.5, motorTorque=50.0, motorStiffness=500.0, motorDamping=20.0, motorMode=0))
    s.addJoint(SceneJoint(c, d, (0, 0, -0.15), (0, 0, 0.05), stiffnessLin=inf, fracture=1.0, fractureLinear=True))
    return s


def _twist(qA, rest, qB, axis):
    r = omlp.quat_mul(omlp.quat_mul(qA, rest) * np.array([-1, -1, -1, 1.0]), qB)
    if r[3] < 0:
        r = -r
    return 2 * np.arctan2(np.dot(r[:3], axis), r[3])


def test_motor_setters_states_and_preconditions():
    _case("motor velocities / targets / torques, motor states, world anchors, poses, joint diagnostics")
    # without gravity the velocity motor alone sets joint 1's relative rate
    g = GPUSolver(_motor_chain(gravity=0.0))
    g.step(90)
    _, q, _, w = g.bodyStates()
    st = g.motorStates([0, 1, 2])
    # motor states vs a numpy evaluation from the body states (rest rotation = identity here)
    axes = [np.array([1.0, 0, 0]), np.array([1.0, 0, 0]), np.array([0, 1.0, 0])]
    pairs = [(None, 1), (1, 2), (2, 3)]
    ident = np.array([0, 0, 0, 1.0])
    for k, (pa, pb) in enumerate(pairs):
        qA = ident if pa is None else q[pa].astype(np.float64)
        qB = q[pb].astype(np.float64)
        ang = _twist(qA, ident, qB, axes[k])
        wA = np.zeros(3) if pa is None else w[pa]
        axw = np.array(omlp.quat_mul(omlp.quat_mul(qB, np.array([*axes[k], 0.0])), qB * np.array([-1, -1, -1, 1.0])))[:3]
        vel = float(np.dot(w[pb] - wA, axw))
        assert abs(st[k][0] - ang) < 1e-4, (k, st[k], ang)
        assert abs(st[k][1] - vel) < 1e-3 * max(1.0, abs(vel)), (k, st[k], vel)
    flush_print(f"  joint 1 rate after 90 frames at target 2 rad/s: {st[1][1]:.3f}")
    # the damped motor (kd 0.5 N m s against the chain's inertia) approaches its command
    assert 1.0 < st[1][1] < 2.3, st[1]
    g.setMotorVelocities([1], [-1.0])
    g.step(90)
    r = g.motorStates([1])[0][1]
    flush_print(f"  joint 1 rate after 90 more frames at target -1 rad/s: {r:.3f}")
    assert -1.3 < r < -0.4
    # preconditions (argument errors, nothing written)
    for call in (lambda: g.setMotorVelocities([2], [1.0]),       # not a velocity motor
                 lambda: g.setMotorTargets([1], [0.3]),          # not a position motor
                 lambda: g.setMotorTorques([1], [-1.0]),         # negative effort
                 lambda: g.setMotorTorques([0], [1.0]),          # no authored motor
                 lambda: g.setMotorTargets([9], [0.0])):         # out of range
        with pytest.raises(AVBDError):
            call()
    g.setMotorTargets([0, 2], [0.7, -0.4])                       # joint 0 has no motor: ignored
    g.setMotorTorques([1, 2], [10.0, 60.0])
    # with gravity the fragile joint breaks under the body's weight
    h = GPUSolver(_motor_chain())
    h.step(30)
    assert h.debugBrokenJoints() == 1
    ml, ma = h.debugMaxLambda()
    assert ml > 0 and np.isfinite(ma)
    # world anchors move the chain; poses zero the velocities
    h.setJointWorldAnchors([0], [[0.5, 0.0, 3.0]])
    h.step(120)
    p, _, v, _ = h.bodyStates()
    assert p[1][0] > 0.3
    h.setBodyPoses([4], [[2.0, 2.0, 1.0]], [[0, 0, 0, 1]])
    p, _, v, _ = h.bodyStates()
    assert np.allclose(p[4], [2, 2, 1]) and np.all(v[4] == 0)


def _rigid_capacity_scene(shared_statics):
    # rigidCapacityScene (GPUSolverTests.swift:251-270)
    s = PhysicsScene("rigid-capacity", SimSettings(gravity=0.0, iterations=1))
    for _ in range(shared_statics):
        s.addBody((0.1, 0.1, 0.1), 0.0, 0.0, (0, 0, 0), shape=1)
    for g in range(1, 65):
        s.addBody((0.1, 0.1, 0.1), 1.0, 0.0, (0, 0, 0), shape=1, collisionGroup=g)
    return s


def test_reference_capacity_and_coloring_failures():
    _case("exact pair capacity 4096 valid / 4160 poisons; unresolved coloring; static palette exhaustion")
    g = GPUSolver(_rigid_capacity_scene(64), maxPairsPerBody=0)
    g.step()
    g.synchronize()
    st = g.stats()
    assert st.pair_candidates == 4096 and st.num_pairs == 4096 and g.runtimeFailure is None
    g = GPUSolver(_rigid_capacity_scene(65), maxPairsPerBody=0)
    g.step()
    with pytest.raises(RuntimeFailure) as e:
        g.synchronize()
    assert (e.value.kind, e.value.frame, e.value.required, e.value.capacity) == ("rigidPairCapacity", 1, 4160, 4096)
    assert g.stats().pair_candidates == 4160 and g.stats().num_pairs == 4096
    # GPUSolverTests.testUnresolvedDynamicColoringPoisonsTheSolve (:305-326)
    s = PhysicsScene("color-capacity", SimSettings(gravity=0.0, iterations=1))
    for _ in range(65):
        s.addBody((0.1, 0.1, 0.1), 1.0, 0.0, (0, 0, 0), shape=1)
    g = GPUSolver(s)
    g.step()
    with pytest.raises(RuntimeFailure) as e:
        g.synchronize()
    assert e.value.kind == "unresolvedColoring" and e.value.frame == 1 and e.value.required > 0
    # GPUSolverTests.testStaticColorPaletteExhaustionFailsDuringInitialization (:342-364)
    s = PhysicsScene("static-color-capacity", SimSettings(gravity=0.0))
    ps = [s.addParticle(0.01, 0.01, (float(i), 0.0, 1.0), friction=0.0) for i in range(65)]
    s.addTri((ps[0], ps[1], ps[2]))
    for a in range(65):
        for b in range(a + 1, 65):
            s.addJoint(SceneJoint(ps[a], ps[b], (0, 0, 0), (0, 0, 0), stiffnessLin=1.0))
    with pytest.raises(RuntimeFailure) as e:
        GPUSolver(s)
    assert (e.value.kind, e.value.required, e.value.capacity) == ("staticColorCapacity", 65, 64)
