"""GPU tests of the learned warm start: the BF16 MFMA network kernels against
the CPU restatement (oracle/mlp.py), GPU training, the feature kernel against
its restatement, the residual fallback, and graph capture of the stage.

Bars: forward within 5e-3 (relative to max |y|) of the bf16-emulating oracle
and 5e-2 of FP64; per-layer gradient relative error <= 2e-2 against the
emulation and <= 0.1 against FP64; features within 5e-4 (units of
|g| dt^2: a last-bit difference of the positions); a rejected learned start
leaves the trajectory bitwise equal to the plain solver's.
"""
import numpy as np
import pytest

from oracle import mlp as omlp
from avbd_amd import GPUSolver, WarmStartMLP, demos

from conftest import flush_print

pytestmark = pytest.mark.gpu


def _case(name):
    flush_print(f"[gpu case] {name}")


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


def _data(n, seed):
    rng = np.random.default_rng(seed)
    x = (rng.standard_normal((n, 24)) * 2.0).astype(np.float32)
    t = (rng.standard_normal((n, 3))).astype(np.float32)
    return x, t


@pytest.mark.parametrize("outputs,n", [(3, 1), (3, 15), (3, 1000), (3, 4133), (6, 15), (6, 2000)])
def test_mlp_forward_matches_oracle(outputs, n):
    _case(f"MLP forward vs oracle, {outputs} outputs, n={n}")
    m = WarmStartMLP(seed=3, outputs=outputs)
    p = m.params
    p[len(p) - outputs:] = np.linspace(-0.3, 0.3, outputs)  # nonzero output biases
    p[24 * 128:24 * 128 + 128] = np.linspace(-0.5, 0.5, 128)
    m.params = p
    x, _ = _data(n, 1)
    y = m.forward(x)
    ye, _ = omlp.forward(p, x, emulate=True)
    y64, _ = omlp.forward(p, x, emulate=False)
    scale = max(np.abs(y64).max(), 1e-3)
    e_em = np.abs(y - ye).max() / scale
    e_64 = np.abs(y - y64).max() / scale
    flush_print(f"  max|y|={scale:.3f}  err vs bf16 emulation {e_em:.2e}, vs FP64 {e_64:.2e}")
    assert e_em <= 5e-3
    assert e_64 <= 5e-2


@pytest.mark.parametrize("outputs,n", [(3, 37), (3, 2048), (3, 5000), (6, 37), (6, 3000)])
def test_mlp_gradient_matches_oracle(outputs, n):
    _case(f"MLP loss gradient vs oracle, {outputs} outputs, n={n}")
    m = WarmStartMLP(seed=5, outputs=outputs)
    p = m.params
    p[24 * 128:24 * 128 + 128] = 0.05                    # most units active
    m.params = p
    x, t = _data(n, 2)
    t = np.random.default_rng(3).standard_normal((n, outputs)).astype(np.float32)
    loss, g = m.lossGrad(x, t)
    le, ge = omlp.loss_grad(p, x, t, emulate=True)
    l64, g64 = omlp.loss_grad(p, x, t, emulate=False)
    flush_print(f"  loss gpu {loss:.6f} emulated {le:.6f} fp64 {l64:.6f}")
    assert abs(loss - le) <= 1e-3 * abs(le)
    off = 0
    for name, shape in omlp.layout(outputs):
        k = int(np.prod(shape))
        a, b, c = g[off:off + k], ge[off:off + k], g64[off:off + k]
        off += k
        r_em, r_64 = _rel(a, b), _rel(a, c)
        flush_print(f"  {name}: rel err vs emulation {r_em:.2e}, vs FP64 {r_64:.2e}")
        assert r_em <= 2e-2, name
        # against unrounded FP64 the bf16 operands (inputs, weights, activations,
        # back-propagated errors) leave a few percent; a 37-row batch does not average them
        assert r_64 <= 0.1, name


def test_mlp_training_reduces_loss():
    _case("MLP training (Adam, shuffled batches) on a smooth regression")
    rng = np.random.default_rng(7)
    x = rng.standard_normal((20000, 24)).astype(np.float32)
    t = np.stack([np.tanh(x[:, 0]) + 0.5 * x[:, 1] * x[:, 2], 0.3 * x[:, 3] - x[:, 4] ** 2 * 0.2,
                  np.sin(x[:, 5])], 1).astype(np.float32)
    m = WarmStartMLP(seed=1)
    losses = m.train(x, t, epochs=25, batch=1024, lr=2e-3, seed=3)
    flush_print(f"  epoch losses {losses[0]:.4f} -> {losses[-1]:.4f} (var t {t.var():.4f})")
    assert np.all(np.isfinite(losses))
    assert losses[-1] < 0.15 * losses[0]
    # the trained forward agrees with the oracle at the trained parameters
    y = m.forward(x[:500])
    ye, _ = omlp.forward(m.params, x[:500], emulate=True)
    assert np.abs(y - ye).max() <= 5e-3 * max(np.abs(ye).max(), 1e-3)


def _free_block(cells=3, seed=4):
    s = demos.soft_block_scene(cells=cells, seed=seed, bottom=6.0, vel_sigma=0.5)
    return s


def test_features_match_restatement():
    _case("warm-start features + converged corrections vs restatement, two steps of a free block")
    s = _free_block()
    g = GPUSolver(s, useGraph=False)
    g.setWarmStart(None, log=True)
    nb = len(s.bodies)
    rng = np.random.default_rng(3)
    p0, r0, _, w0 = g.bodyStates()
    v0 = rng.normal(0, 0.4, (nb, 3)).astype(np.float32)
    v0[0] = 0
    g.setBodyStates(np.arange(nb, dtype=np.uint32), p0, r0, v0, w0)
    st = g.settings
    dt, grav = np.float32(st.dt), np.float32(st.gravity)
    L = np.float32(max(np.float32(abs(grav)) * dt * dt, np.float32(1e-6)))
    mass = g.bodyMass()
    ids = np.array([i for i, b in enumerate(s.bodies) if b["particle"] and mass[i] > 0])
    ring = omlp.ring_of(s, ids)
    gz = np.array([0, 0, grav], np.float32)
    corr = np.zeros((nb, 3), np.float32)
    corr_prev = np.zeros((nb, 3), np.float32)
    x_n, v_n, v_prev = p0.copy(), v0.copy(), v0.copy()
    for step in range(2):
        g.step()
        X, T, idx = g.warmStartSamples()
        assert np.array_equal(idx, ids)
        acc = (v_n - v_prev) / dt
        w = np.clip(acc[:, 2] * np.float32(np.sign(grav)) / np.float32(abs(grav)), 0, 1).astype(np.float32)
        x_guess = x_n + v_n * dt + gz * ((w[:, None] * dt) * dt)
        want = omlp.features_of(ids, x_n, v_n, v_prev, corr_prev, x_guess, corr, ring, np.zeros(len(ids)), L, dt)
        err = np.abs(X - want).max()
        x1 = g.positions()
        tw = (x1[ids] - x_guess[ids]) / L
        terr = np.abs(T - tw).max()
        flush_print(f"  step {step}: features max err {err:.2e}, targets max err {terr:.2e}")
        # a last-bit difference of a ~6 m position is 2e-4 in units of L
        assert err <= 5e-4
        assert terr <= 1e-3
        if step == 0:
            # free flight closed form: guess offset v dt (within the rounding of
            # a ~6 m position, 2e-4 in units of L); no correction history yet
            np.testing.assert_allclose(X[:, 9:12], v0[ids] * dt / L, rtol=0, atol=5e-4)
            assert np.all(X[:, 6:9] == 0) and np.all(X[:, 12:15] == 0)
        v_prev = v_n
        _, _, v_n, _ = g.bodyStates()
        x_n = x1
        corr_prev = corr
        corr = np.zeros((nb, 3), np.float32)
        corr[ids] = T


def _bad_net():
    m = WarmStartMLP(seed=2)
    p = m.params
    p[omlp.NUM_PARAMS - 3:] = [40.0, -30.0, 25.0]   # a start tens of |g| dt^2 off
    m.params = p
    return m


def test_fallback_keeps_plain_trajectory():
    _case("learned start with a higher residual falls back: bitwise the plain trajectory")
    s = demos.soft_block_scene(cells=4, seed=6)
    kw = dict(lanesPerBody=8, sweepMode=1)   # one kernel configuration for every solver compared
    a = GPUSolver(s, **kw)
    b = GPUSolver(s, **kw)
    m = _bad_net()
    b.setWarmStart(m, learned=True)
    for _ in range(12):
        a.step()
        b.step()
        st = b.stats()
        assert st.ws_accepted == 0
        assert st.ws_residual_learned > st.ws_residual_plain
    for x, y in zip(a.bodyStates(), b.bodyStates()):
        assert np.array_equal(x, y)
    c = GPUSolver(s, **kw)
    c.setWarmStart(m, learned=True, force=True)
    d = GPUSolver(s, **kw)
    c.step(3)
    d.step(3)
    st = c.stats()
    assert st.ws_accepted == 1
    assert not np.array_equal(c.positions(), d.positions())


def test_learned_stage_graph_matches_eager():
    _case("learned stage under hipGraph replay vs eager launches")
    s = demos.soft_block_scene(cells=4, seed=8)
    m = WarmStartMLP(seed=4)
    a = GPUSolver(s, useGraph=True, lanesPerBody=8, sweepMode=1)
    b = GPUSolver(s, useGraph=False, lanesPerBody=8, sweepMode=1)
    for g in (a, b):
        g.setWarmStart(m, learned=True, force=True, log=True, smooth=2)
    for k in range(8):
        a.step()
        b.step()
        a.synchronize()
        assert np.array_equal(a.positions(), b.positions()), k
    xa, ta, _ = a.warmStartSamples()
    xb, tb, _ = b.warmStartSamples()
    assert np.array_equal(xa, xb) and np.array_equal(ta, tb)


def test_trained_network_gains_over_plain_warm_start():
    _case("learned warm start trained on the GPU vs the plain guess, 9^3 block")
    from avbd_amd.learned import collect_samples
    cells = 8
    X, T, label_its = collect_samples([demos.soft_block_scene(cells=cells, seed=20 + k) for k in range(3)], 60)
    m = WarmStartMLP(seed=1)
    losses = m.train(X, T, epochs=40, batch=4096, lr=1e-3, seed=2)
    flush_print(f"  {len(X)} pairs (label sweeps {label_its:.0f}), loss {losses[0]:.3f} -> {losses[-1]:.3f}, "
                f"target variance {T.var():.3f}")
    assert losses[-1] < 0.3 * T.var()
    s = demos.soft_block_scene(cells=cells, seed=99)     # a scene the network never saw

    def residuals(learned):
        g = GPUSolver(s)
        if learned:
            g.setWarmStart(m)
        out = []
        for _ in range(60):
            g.step()
            out.append(g.residual())
        return np.array(out)

    def iterations(learned):
        g = GPUSolver(s)
        if learned:
            g.setWarmStart(m)
        return np.array([g.stepToResidual(1e-3, 200, 5)[0] for _ in range(60)])

    rp, rl = residuals(False), residuals(True)
    ip, il = iterations(False), iterations(True)
    flush_print(f"  residual after 20 sweeps: median plain {np.median(rp):.3e} learned {np.median(rl):.3e}; "
                f"sweeps to 1e-3: plain {ip.mean():.1f} learned {il.mean():.1f}")
    assert np.median(rl) < 0.6 * np.median(rp)
    assert il.mean() < ip.mean()


def _chain(links=6, seed=9):
    """A free chain of boxes joined by ball joints, hung from the world 10 m
    above the ground (no contacts)."""
    from avbd_amd.scene import SceneJoint
    rng = demos.SplitMix64(seed)
    s = demos.PhysicsScene("chain", demos.SimSettings(iterations=10))
    demos.add_ground(s)
    prev, top = -1, (0.0, 0.0, 10.0)
    for i in range(links):
        L = rng.uniform(0.1, 0.4)
        c = (top[0] + 0.05 * i, top[1], top[2] - L / 2)
        b = s.addBody((0.08, 0.08, L), 500.0, 0.5, c)
        rA = (0.0, 0.0, -prevL / 2) if prev >= 0 else top
        s.addJoint(SceneJoint(prev, b, rA, (0.0, 0.0, L / 2)))
        prev, prevL, top = b, L, (c[0], c[1], c[2] - L / 2)
    return s


def _q_mul(a, b):
    return np.array(omlp.quat_mul(a.astype(np.float64), b.astype(np.float64)), np.float32)


def test_rigid_features_match_restatement():
    _case("rigid warm-start features + corrections vs restatement, free chain")
    s = _chain()
    g = GPUSolver(s, useGraph=False)
    g.setWarmStart(None, log=True)
    nb = len(s.bodies)
    rng = np.random.default_rng(4)
    p0, r0, _, _ = g.bodyStates()
    v0 = rng.normal(0, 0.3, (nb, 3)).astype(np.float32)
    w0 = rng.normal(0, 1.0, (nb, 3)).astype(np.float32)
    v0[0] = 0
    w0[0] = 0
    g.setBodyStates(np.arange(nb, dtype=np.uint32), p0, r0, v0, w0)
    st = g.settings
    dt, grav = np.float32(st.dt), np.float32(st.gravity)
    L = np.float32(max(np.float32(abs(grav)) * dt * dt, np.float32(1e-6)))
    ids = np.arange(1, nb)
    radius = np.array([np.linalg.norm(np.array(b["size"]) / 2) for b in s.bodies], np.float32)
    ring = omlp.joint_ring_of(s, ids)
    X, T, idx = None, None, None
    g.step()
    X, T, idx = g.warmStartSamples(rigid=True)
    assert np.array_equal(idx, ids)
    z = np.zeros((nb, 3), np.float32)
    x_guess = p0 + v0 * dt
    want = omlp.features_rigid_of(ids, p0, v0, v0, w0, x_guess, z, z, radius, ring, np.zeros(len(ids)), L, dt)
    err = np.abs(X - want).max()
    p1, r1, _, _ = g.bodyStates()
    tl = (p1[ids] - x_guess[ids]) / L
    terr = np.abs(T[:, :3] - tl).max()
    # angular target: rotation from the guess q (+) w dt to the solution, times r / L
    ta = []
    for k, b in enumerate(ids):
        q = r0[b]
        dq = _q_mul(np.array([*(w0[b] * dt), 0], np.float32), q)
        qg = q + 0.5 * dq
        qg = qg / np.linalg.norm(qg)
        rel = _q_mul(r1[b], np.array([-qg[0], -qg[1], -qg[2], qg[3]], np.float32))
        if rel[3] < 0:
            rel = -rel
        ta.append(2 * rel[:3] * radius[b] / L)
    aerr = np.abs(T[:, 3:] - np.array(ta)).max()
    flush_print(f"  features max err {err:.2e}, linear targets {terr:.2e}, angular targets {aerr:.2e}")
    assert err <= 5e-4
    assert terr <= 1e-3
    assert aerr <= 2e-2
    assert np.abs(T).max() > 1e-3     # the joints moved the solution off the guess


def test_rigid_fallback_keeps_plain_trajectory():
    _case("rigid learned start with a higher residual falls back: bitwise the plain trajectory")
    s = demos.stack_scene(6)
    kw = dict(lanesPerBody=8, sweepMode=1)
    a = GPUSolver(s, **kw)
    b = GPUSolver(s, **kw)
    m = WarmStartMLP(seed=2, outputs=6)
    p = m.params
    p[len(p) - 6:] = [30.0, -20.0, 25.0, 5.0, 5.0, -5.0]
    m.params = p
    b.setWarmStart(rigid=m)
    for _ in range(10):
        a.step()
        b.step()
        assert b.stats().ws_accepted == 0
    for x, y in zip(a.bodyStates(), b.bodyStates()):
        assert np.array_equal(x, y)
    c = GPUSolver(s, **kw)
    c.setWarmStart(rigid=m, force=True)
    d = GPUSolver(s, **kw)
    c.step(2)
    d.step(2)
    assert c.stats().ws_accepted == 1
    assert not np.array_equal(c.bodyStates()[1], d.bodyStates()[1])


def test_trained_rigid_network_gains_over_plain_warm_start():
    _case("rigid learned warm start trained on the GPU vs the plain guess, 32 batched rigid scenes")
    from avbd_amd.learned import collect_samples
    train = [demos.batched_scene(32, seed_base=(8 << 32) + 1000 * k)[0] for k in range(2)]
    X, T, _ = collect_samples(train, 40, tol=1e-4, max_iterations=100, check_every=10, rigid=True)
    m = WarmStartMLP(seed=1, outputs=6)
    losses = m.train(X, T, epochs=40, batch=4096, lr=1e-3, seed=2)
    scene, _ = demos.batched_scene(32)      # draws the network never saw

    def residuals(learned):
        g = GPUSolver(scene)
        if learned:
            g.setWarmStart(rigid=m)
        out = []
        for _ in range(40):
            g.step()
            out.append(g.residual())
        return np.array(out)

    rp, rl = residuals(False), residuals(True)
    flush_print(f"  {len(X)} pairs, loss {losses[0]:.1f} -> {losses[-1]:.1f}; residual after 10 sweeps: "
                f"median plain {np.median(rp):.3e} learned {np.median(rl):.3e}")
    assert np.median(rl) < 0.6 * np.median(rp)


def test_learned_start_at_scale_block_55_and_cloth_256():
    """The learned start on the benchmark's at-scale scenes: the 55^3-cell
    block (175,616 vertices, 998,250 tets) and the 256^2 cloth drape, with a
    network trained on small blocks.  On the block the fixed-iteration
    residual drops against the plain start; on the cloth, where the network
    has no skill, the residual fallback keeps the plain result."""
    _case("learned start at scale: 55^3 block and 256^2 cloth vs the plain start")
    from avbd_amd.learned import collect_samples
    X, T, _ = collect_samples([demos.soft_block_scene(cells=8, seed=20 + k) for k in range(3)], 60)
    m = WarmStartMLP(seed=1)
    m.train(X, T, epochs=40, batch=4096, lr=1e-3, seed=2)

    def medians(scene, frames):
        out = []
        for learned in (False, True):
            with GPUSolver(scene) as g:
                if learned:
                    g.setWarmStart(m)
                r = []
                for _ in range(frames):
                    g.step()
                    r.append(g.residual())
                out.append(float(np.median(r)))
        return out

    bp, bl = medians(demos.soft_block_scene(cells=55), 12)
    cp, cl = medians(demos.cloth_scene(), 12)
    flush_print(f"  55^3 block residual after 20 sweeps: plain {bp:.3e} learned {bl:.3e}; "
                f"256^2 cloth after 30 sweeps: plain {cp:.3e} learned {cl:.3e}")
    assert np.isfinite([bp, bl, cp, cl]).all()
    assert bl < 0.8 * bp
    assert cl <= 1.02 * cp
