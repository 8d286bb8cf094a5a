"""The reference's own CPU AVBD path (CPUSolver semantics, oracle/cpu_path.cpp)
against the reference's CPU solver tests (Tests/PhysicsAVBDTests/
CPUSolverTests.swift:35-98) with their thresholds, and the reference's
CPU-vs-GPU contract (GPUSolverTests.swift:831-846: a 3-cube stack, every
body within 0.05 m after 120 frames) between the CPU path and the GPU-kernel
semantics restated in the oracle.  The backend side of that contract runs in
tests/test_gpu_cpu_path.py.

The CPUSolverTests build `CPUSolver()` directly, whose defaults differ from
the scene defaults in betaLin (5000 instead of 1e4, CPUSolver.swift:144); the
scenes below set it."""
import math

import numpy as np

import oracle
import reference_scenes as R
from avbd_amd.scene import PhysicsScene, SceneJoint, SimSettings


def _cpu_defaults(name):
    return PhysicsScene(name, SimSettings(betaLin=5000.0))


def test_box_rests_on_ground():
    s = _cpu_defaults("box on ground")
    s.addBody((100, 100, 1), 0.0, 0.5, (0, 0, -0.5))
    b = s.addBody((1, 1, 1), 1.0, 0.5, (0, 0, 0.6))
    c = oracle.CpuPathSolver(s)
    c.step(120)
    p, _, v, _ = c.bodyStates64()
    assert abs(p[b, 2] - 0.5) < 0.03
    assert np.linalg.norm(v[b]) < 0.05


def test_free_fall_matches_implicit_euler():
    s = _cpu_defaults("free fall")
    b = s.addBody((1, 1, 1), 1.0, 0.5, (0, 0, 100))
    c = oracle.CpuPathSolver(s)
    c.step(30)
    v, z = 0.0, 100.0
    for _ in range(30):
        v += -10.0 / 60.0
        z += v / 60.0
    assert abs(c.positions()[b, 2] - z) < 0.01


def test_pendulum_joint_error():
    s = _cpu_defaults("pendulum")
    bob = s.addBody((0.5, 0.5, 0.5), 1.0, 0.5, (1, 0, 0), collisionEnabled=False)
    s.addJoint(SceneJoint(-1, bob, (0, 0, 0), (-1, 0, 0), stiffnessLin=math.inf, stiffnessAng=0.0))
    c = oracle.CpuPathSolver(s)
    c.step(120)
    assert c.maxConstraintError() < 0.01


def test_small_stack_stability():
    s = _cpu_defaults("small stack")
    s.addBody((100, 100, 1), 0.0, 0.7, (0, 0, -0.5))
    tops = [s.addBody((1, 1, 1), 1.0, 0.7, (0, 0, 0.5 + i)) for i in range(5)]
    c = oracle.CpuPathSolver(s)
    c.step(300)
    p = c.positions()
    for i, b in enumerate(tops):
        assert abs(p[b, 2] - (0.5 + i)) < 0.08
        assert math.hypot(p[b, 0], p[b, 1]) < 0.15


def test_cpu_path_tracks_gpu_semantics_on_the_stack():
    # GPUSolverTests.testCPUGPUParity with the GPU side restated by the oracle
    s = R.stack(3)
    c = oracle.CpuPathSolver(s)
    o = oracle.OracleSolver(s)
    c.step(120)
    o.step(120)
    d = np.linalg.norm(c.positions() - o.positions().astype(np.float64), axis=1)
    assert d.max() < 0.05, d


def test_forks_are_visible():
    """The CPU path is a different algorithm at the forks, not a copy of the
    GPU semantics: on a 10-cube stack (plain LDL^T, no trust region, lambda_max
    cap, index-order sweep) the two trajectories differ, while both keep the
    stack standing."""
    s = R.stack(10)
    c = oracle.CpuPathSolver(s)
    o = oracle.OracleSolver(s)
    c.step(240)
    o.step(240)
    pc, po = c.positions(), o.positions().astype(np.float64)
    assert np.abs(pc - po).max() > 1e-6
    assert np.all(np.abs(pc[1:, 2] - (0.5 + np.arange(10))) < 0.15)
