"""oracle.mlp -- TEST INFRASTRUCTURE ONLY.

CPU restatement of the learned warm start's network and of its per-particle
features (the backend's own addition; the reference has no learned warm
start, so nothing in the reference pins it).  Used by tests/ as the checker
of the GPU forward / backward kernels and of the feature kernel.

Parity status:
  forward / loss_grad (FP64)   pinned: central finite differences of the loss
                               (tests/test_mlp_oracle.py), loss by direct
                               evaluation, ReLU-free linear case vs the
                               closed-form least-squares gradient
  forward / loss_grad (bf16)   emulates the GPU's rounding points on top of
                               the FP64 restatement; checked against it within
                               the bf16 bound
  features                     restated from the definition in avbd.h; pinned
                               by a hand-computed case and closed forms (free
                               flight: guess offset v dt) in tests
"""
import numpy as np

INPUTS, HIDDEN, OUTPUTS = 24, 128, 3


def layout(nout=OUTPUTS):
    return (("W1", (INPUTS, HIDDEN)), ("b1", (HIDDEN,)), ("W2", (HIDDEN, HIDDEN)), ("b2", (HIDDEN,)),
            ("W3", (HIDDEN, nout)), ("b3", (nout,)))


def num_params(nout=OUTPUTS):
    return sum(int(np.prod(s)) for _, s in layout(nout))


LAYOUT = layout()
NUM_PARAMS = num_params()


def outputs_of(p):
    """Output count of a flat parameter array (its length is 19712 + 129 nout)."""
    n = (len(p) - num_params(0)) // (HIDDEN + 1)
    assert num_params(n) == len(p), len(p)
    return n


def split(p):
    out, o = {}, 0
    for name, shape in layout(outputs_of(p)):
        n = int(np.prod(shape))
        out[name] = np.asarray(p[o:o + n], dtype=np.float64).reshape(shape)
        o += n
    return out


def join(d):
    nout = np.asarray(d["b3"]).size
    return np.concatenate([np.asarray(d[name], dtype=np.float64).ravel() for name, _ in layout(nout)])


def bf16(a):
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


def _id(a):
    return np.asarray(a, dtype=np.float64)


def forward(p, x, emulate=False):
    """y [n,3] and the activations; emulate=True rounds inputs, weights and
    hidden activations to bf16 where the GPU kernels do."""
    q = split(p)
    r = bf16 if emulate else _id
    x = r(x)
    W1, W2, W3 = r(q["W1"]), r(q["W2"]), r(q["W3"])
    a1 = x @ W1 + q["b1"]
    h1 = r(np.maximum(a1, 0.0))
    a2 = h1 @ W2 + q["b2"]
    h2 = r(np.maximum(a2, 0.0))
    y = h2 @ W3 + q["b3"]
    return y, (x, h1, h2, W1, W2, W3)


def loss_grad(p, x, t, emulate=False):
    """loss = sum |y - t|^2 / (nout n) and its gradient (flat, the parameter layout)."""
    n = len(x)
    r = bf16 if emulate else _id
    y, (xr, h1, h2, W1, W2, W3) = forward(p, x, emulate)
    e = y - np.asarray(t, dtype=np.float64)
    k = float(y.shape[1] * n)
    loss = float(np.sum(e * e) / k)
    if emulate:  # the GPU forms dY in f32
        dy = (2.0 * e.astype(np.float32) * np.float32(1.0 / k)).astype(np.float64)
    else:
        dy = 2.0 * e / k
    dyr = r(dy)
    g = {"W3": h2.T @ dyr, "b3": dy.sum(0)}
    dz2 = (dyr @ W3.T) * (h2 > 0)
    dz2r = r(dz2)
    g["W2"] = h1.T @ dz2r
    g["b2"] = dz2.sum(0)
    dz1 = (dz2r @ W2.T) * (h1 > 0)
    dz1r = r(dz1)
    g["W1"] = xr.T @ dz1r
    g["b1"] = dz1.sum(0)
    return loss, join(g)


def features_of(ids, pos_start, vel, vel_prev, corr_prev, x_guess, corr, ring, contacts, L, dt):
    """Features of particles `ids` (body ids).  Every per-body array is
    [num_bodies, 3] float32: pos_start x_n, vel v_n, vel_prev v_{n-1},
    corr_prev the corrections of two steps back, x_guess the adaptive guess,
    corr the previous step's corrections; ring[k] the ascending 1-ring body
    ids of ids[k]; contacts[k] the particle's contact-entry count."""
    f32 = np.float32
    invL = f32(1.0) / f32(L)
    vs = f32(dt) * invL
    out = np.zeros((len(ids), INPUTS), np.float32)
    for k, b in enumerate(ids):
        xv, vv, cv = pos_start[b], vel[b], corr[b]
        rc = len(ring[k])
        vsum = np.zeros(3, np.float32)
        csum = np.zeros(3, np.float32)
        for j in ring[k]:
            vsum = vsum + vel[j]
            csum = csum + corr[j]
        f = np.zeros(INPUTS, np.float32)
        f[0:3] = vv * vs
        f[3:6] = (vv - vel_prev[b]) * vs
        f[6:9] = corr_prev[b]
        f[9:12] = (x_guess[b] - xv) * invL
        f[12:15] = cv
        if rc:
            inv = f32(1.0) / f32(rc)
            f[15:18] = (vsum * inv - vv) * vs
            f[18:21] = csum * inv - cv
        f[21] = f32(min(float(contacts[k]), 8.0)) * f32(0.125)
        f[22] = f32(rc) * f32(0.0625)
        f[23] = np.sqrt(np.float32(cv[0] * cv[0] + cv[1] * cv[1] + cv[2] * cv[2]))
        out[k] = f
    return out


def quat_mul(a, b):
    ax, ay, az, aw = a
    bx, by, bz, bw = b
    return np.array([aw * bx + ax * bw + ay * bz - az * by, aw * by - ax * bz + ay * bw + az * bx,
                     aw * bz + ax * by - ay * bx + az * bw, aw * bw - ax * bx - ay * by - az * bz])


def features_rigid_of(ids, pos_start, vel, vel_prev, ang_vel, x_guess, corr, corr_ang, radius, ring, contacts, L,
                      dt):
    """Rigid-body features of bodies `ids` (avbd.h): per-body [num_bodies, 3]
    float32 arrays, radius [num_bodies] (bounding radius, floored at 1 mm),
    ring[k] the ascending joint partners of ids[k]."""
    f32 = np.float32
    invL = f32(1.0) / f32(L)
    vs = f32(dt) * invL
    out = np.zeros((len(ids), INPUTS), np.float32)
    for k, b in enumerate(ids):
        ws = f32(dt) * f32(max(radius[b], f32(1e-3))) * invL
        cl, ca = corr[b], corr_ang[b]
        rc = len(ring[k])
        csum = np.zeros(3, np.float32)
        for j in ring[k]:
            csum = csum + corr[j]
        f = np.zeros(INPUTS, np.float32)
        f[0:3] = vel[b] * vs
        f[3:6] = ang_vel[b] * ws
        f[6:9] = (vel[b] - vel_prev[b]) * vs
        f[9:12] = (x_guess[b] - pos_start[b]) * invL
        f[12:15] = cl
        f[15:18] = ca
        if rc:
            f[18:21] = csum * (f32(1.0) / f32(rc)) - cl
        f[21] = f32(min(float(contacts[k]), 8.0)) * f32(0.125)
        f[22] = f32(min(float(rc), 4.0)) * f32(0.25)
        f[23] = np.sqrt(np.float32(cl[0] * cl[0] + cl[1] * cl[1] + cl[2] * cl[2]))
        out[k] = f
    return out


def joint_ring_of(scene, ids):
    """Joint partners of each rigid body (bodies of active joints), ascending."""
    nb = [set() for _ in range(len(scene.bodies))]
    for j in scene.joints:
        a, b = j.bodyA, j.bodyB
        inert = j.stiffnessLin == 0 and j.stiffnessAng == 0 and j.motorTorque == 0
        if a < 0 or inert:
            continue
        nb[a].add(b)
        nb[b].add(a)
    return [np.array(sorted(nb[b]), dtype=np.int64) for b in ids]


def ring_of(scene, ids):
    """Topological 1-ring of each particle: bodies sharing a tet, triangle
    element or spring with it, ascending (the bend stencil of a triangle pair
    adds the opposite vertices of adjacent triangles)."""
    nb = [set() for _ in range(len(scene.bodies))]

    def link(group):
        for a in group:
            for b in group:
                if a != b:
                    nb[a].add(b)
    for t in scene.tets:
        link(list(t["ids"]) if isinstance(t, dict) else list(t[0]))
    for s in scene.springs:
        link([s["a"], s["b"]] if isinstance(s, dict) else [s[0], s[1]])
    tris = [tuple(t["ids"]) if isinstance(t, dict) else tuple(t[0]) for t in getattr(scene, "tris", [])]
    for t in tris:
        link(list(t))
    # hinges: triangle pairs sharing an edge link their opposite vertices
    edge_opp = {}
    for t in tris:
        for k in range(3):
            a, b, c = t[k], t[(k + 1) % 3], t[(k + 2) % 3]
            edge_opp.setdefault((min(a, b), max(a, b)), []).append(c)
    for (a, b), opp in edge_opp.items():
        if len(opp) == 2:
            link([a, b, opp[0], opp[1]])
    return [np.array(sorted(nb[b]), dtype=np.int64) for b in ids]
