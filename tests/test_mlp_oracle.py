"""CPU pins of the learned warm start's restatement (oracle/mlp.py): the
FP64 network gradient against central finite differences and a closed form,
bf16 rounding known answers, the emulated-rounding gradient within the bf16
bound of FP64, and the feature restatement on hand-computed cases."""
import numpy as np

from oracle import mlp as o
from avbd_amd.scene import PhysicsScene, SimSettings


def _params(seed, scale=0.3):
    rng = np.random.default_rng(seed)
    return np.concatenate([rng.standard_normal(int(np.prod(s))) * scale for _, s in o.LAYOUT])


def test_layout_size_matches_abi():
    assert o.NUM_PARAMS == 20099 == o.num_params(3)      # AVBD_MLP_PARAMS in avbd.h
    assert o.num_params(6) == 20486                       # AVBD_MLP_PARAMS_RIGID
    for nout in (3, 6):
        p = np.arange(o.num_params(nout), dtype=np.float64)
        assert o.outputs_of(p) == nout
        assert np.array_equal(o.join(o.split(p)), p)


def test_six_output_gradient_matches_finite_differences():
    rng = np.random.default_rng(5)
    p = np.concatenate([rng.standard_normal(int(np.prod(s))) * 0.3 for _, s in o.layout(6)])
    x, t = rng.standard_normal((19, 24)), rng.standard_normal((19, 6))
    _, g = o.loss_grad(p, x, t)
    for i in rng.choice(len(p), 40, replace=False):
        pp, pm = p.copy(), p.copy()
        pp[i] += 1e-6
        pm[i] -= 1e-6
        fd = (o.loss_grad(pp, x, t)[0] - o.loss_grad(pm, x, t)[0]) / 2e-6
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), i


def test_fp64_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = _params(0)
    x, t = rng.standard_normal((37, 24)), rng.standard_normal((37, 3))
    _, g = o.loss_grad(p, x, t)
    # every layer's block sampled
    offs = np.cumsum([0] + [int(np.prod(s)) for _, s in o.LAYOUT])
    idx = np.concatenate([rng.choice(np.arange(a, b), min(8, b - a), replace=False) for a, b in zip(offs, offs[1:])])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (o.loss_grad(pp, x, t)[0] - o.loss_grad(pm, x, t)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), i


def test_linear_regime_closed_form():
    # biases large enough that no ReLU clips: the network is the affine map
    # y = x A + c, and dL/db3 = 2 sum(y - t) / (3 n), dL/dW1 = x^T (dY W3^T W2^T) / ...
    rng = np.random.default_rng(2)
    q = o.split(_params(3, 0.05))  # 3 outputs
    q["b1"][:] = 50.0
    q["b2"][:] = 1000.0
    p = o.join(q)
    x, t = rng.standard_normal((20, 24)), rng.standard_normal((20, 3))
    A = q["W1"] @ q["W2"] @ q["W3"]
    c = (q["b1"] @ q["W2"] + q["b2"]) @ q["W3"] + q["b3"]
    y, _ = o.forward(p, x)
    np.testing.assert_allclose(y, x @ A + c, rtol=1e-12, atol=1e-10)
    loss, g = o.loss_grad(p, x, t)
    e = x @ A + c - t
    assert abs(loss - np.sum(e * e) / 60.0) <= 1e-10 * loss
    G = o.split(g)
    np.testing.assert_allclose(G["b3"], 2 * e.sum(0) / 60.0, rtol=1e-12)
    dY = 2 * e / 60.0
    np.testing.assert_allclose(G["W1"], x.T @ (dY @ (q["W2"] @ q["W3"]).T), rtol=1e-10, atol=1e-12)


def test_bf16_rounding_known_answers():
    v = o.bf16(np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, 1 + 2 ** -8 + 2 ** -20, -2.0 - 2 ** -7, 0.0]))
    assert list(v) == [1.0, 1.0, 1.015625, 1.0078125, -2.0, 0.0]


def test_emulated_rounding_within_bf16_bound():
    rng = np.random.default_rng(4)
    p = _params(5, 0.15)
    x, t = rng.standard_normal((300, 24)), rng.standard_normal((300, 3))
    y64, _ = o.forward(p, x)
    ye, _ = o.forward(p, x, emulate=True)
    assert np.abs(ye - y64).max() <= 3e-2 * np.abs(y64).max()
    _, g64 = o.loss_grad(p, x, t)
    _, ge = o.loss_grad(p, x, t, emulate=True)
    assert np.linalg.norm(ge - g64) <= 3e-2 * np.linalg.norm(g64)


def test_features_hand_case():
    # one particle at x = (0, 0, 1), v = (1, 0, -2), v_prev = (1, 0, 0),
    # ring {1, 2} with velocities (0, 0, 0) and (2, 0, 0), correction two steps
    # back (0.5, -1, 0)
    dt, L = np.float32(0.01), np.float32(1e-3)
    pos = np.array([[0, 0, 1], [0, 0, 0], [0, 0, 0]], np.float32)
    vel = np.array([[1, 0, -2], [0, 0, 0], [2, 0, 0]], np.float32)
    vp = np.array([[1, 0, 0], [0, 0, 0], [0, 0, 0]], np.float32)
    xg = pos + vel * dt
    corr = np.zeros((3, 3), np.float32)
    corr[2] = [3, 0, 4]
    cprev = np.array([[0.5, -1, 0], [0, 0, 0], [0, 0, 0]], np.float32)
    f = o.features_of([0], pos, vel, vp, cprev, xg, corr, [np.array([1, 2])], [3], L, dt)[0]
    np.testing.assert_allclose(f[0:3], [10, 0, -20], rtol=1e-6)
    np.testing.assert_allclose(f[3:6], [0, 0, -20], rtol=1e-6)
    np.testing.assert_allclose(f[6:9], [0.5, -1, 0])
    np.testing.assert_allclose(f[9:12], [10, 0, -20], rtol=1e-5)
    np.testing.assert_allclose(f[12:15], 0)
    np.testing.assert_allclose(f[15:18], [0, 0, 20], rtol=1e-6, atol=1e-5)   # mean v (1,0,0) minus own
    np.testing.assert_allclose(f[18:21], [1.5, 0, 2], rtol=1e-6)
    assert f[21] == np.float32(0.375) and f[22] == np.float32(0.125) and f[23] == 0


def test_rigid_features_hand_case():
    # body 0 of radius 0.5 spinning about z at 2 rad/s, joint partner 1 with correction (2, 0, 0)
    dt, L = np.float32(0.01), np.float32(1e-3)
    z3 = np.zeros((2, 3), np.float32)
    vel = np.array([[0, 1, 0], [0, 0, 0]], np.float32)
    w = np.array([[0, 0, 2], [0, 0, 0]], np.float32)
    corr = np.array([[0, 0, 0], [2, 0, 0]], np.float32)
    f = o.features_rigid_of([0], z3, vel, vel, w, z3 + vel * dt, corr, z3, np.array([0.5, 0.5], np.float32),
                            [np.array([1])], [2], L, dt)[0]
    np.testing.assert_allclose(f[0:3], [0, 10, 0], rtol=1e-6)
    np.testing.assert_allclose(f[3:6], [0, 0, 10], rtol=1e-6)        # w dt r / L
    np.testing.assert_allclose(f[9:12], [0, 10, 0], rtol=1e-6)
    np.testing.assert_allclose(f[18:21], [2, 0, 0], rtol=1e-6)
    assert f[21] == np.float32(0.25) and f[22] == np.float32(0.25) and f[23] == 0


def test_ring_of_tet_and_hinge():
    s = PhysicsScene("ring", SimSettings())
    for i in range(6):
        s.addParticle(0.01, 1.0, (i, 0, 0))
    s.addTet((0, 1, 2, 3), 1.0, 1.0)
    s.addSpring(3, 4, (0, 0, 0), (0, 0, 0), 10.0)
    r = o.ring_of(s, [0, 3, 4, 5])
    assert [list(a) for a in r] == [[1, 2, 3], [0, 1, 2, 4], [3], []]
    c = PhysicsScene("hinge", SimSettings())
    for i in range(4):
        c.addParticle(0.01, 1.0, (i, i % 2, 0))
    c.addTri((0, 1, 2), 1.0, 1.0, 1e-3)
    c.addTri((1, 3, 2), 1.0, 1.0, 1e-3)
    r = o.ring_of(c, [0, 3])
    assert [list(a) for a in r] == [[1, 2, 3], [0, 1, 2]]   # opposite vertices linked by the hinge
