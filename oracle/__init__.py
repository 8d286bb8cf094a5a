"""oracle -- TEST INFRASTRUCTURE ONLY.

ctypes wrapper of oracle/liboracle.so, the serial CPU restatement of the
reference AVBD step.  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may use it, and only as the checker / baseline.

Parity status per oracle function (what pins it; see tests/test_oracle.py):
  splitmix64            pinned: published SplitMix64 known answers
  snh_tet               pinned: finite-difference gradient of the SNH energy,
                        rest-state closed form, Hessian vs FD of the gradient
  solve6x6              pinned: numpy.linalg.solve on random SPD systems
  cell hash / lists     pinned: brute-force recompute of floor(p/cell) + hash
  pair stream           pinned: O(n^2) sphere-bound brute force
  static coloring       pinned: validity + serial greedy on small graphs
  dynamic coloring      pinned: validity (no same-color contact edge)
  predict / finalize    pinned: free fall equals implicit Euler (closed form)
  contact / dual        pinned: box rest height 0.5 +- 0.03, stack thresholds
                        (CPUSolverTests.swift:35-98), dual sign / cone invariants
  joints                pinned: pendulum length invariant (CPUSolverTests.swift:67-77)
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OSError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
    lib = C.CDLL(LIB_PATH)
    vp, u32p, fp, dp, i32p = C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_float), C.POINTER(C.c_double), \
        C.POINTER(C.c_int32)
    sig = {
        "oracle_create": (C.c_int, [vp, C.c_int, C.c_int, C.POINTER(vp)]),
        "oracle_destroy": (None, [vp]),
        "oracle_step": (C.c_int, [vp]),
        "oracle_set_settings": (C.c_int, [vp, vp]),
        "oracle_set_threads": (C.c_int, [vp, C.c_int]),
        "oracle_step_to_residual": (C.c_int, [vp, C.c_double, C.c_int, C.c_int, C.POINTER(C.c_int), dp]),
        "oracle_get_body_states": (C.c_int, [vp, C.c_uint32, fp, fp, fp, fp]),
        "oracle_get_body_states_f64": (C.c_int, [vp, C.c_uint32, dp, dp, dp, dp]),
        "oracle_set_body_states": (C.c_int, [vp, u32p, C.c_uint32, fp, fp, fp, fp]),
        "oracle_get_colors": (C.c_int, [vp, C.c_uint32, u32p]),
        "oracle_get_cell_lists": (C.c_int, [vp, u32p, u32p, fp, u32p, u32p, u32p]),
        "oracle_get_pairs": (C.c_int, [vp, C.c_uint32, u32p, u32p]),
        "oracle_get_manifolds": (C.c_int, [vp, C.c_uint32, u32p, u32p, fp, fp]),
        "oracle_residual": (C.c_int, [vp, dp]),
        "oracle_stats": (C.c_int, [vp, i32p, i32p, i32p, i32p, i32p]),
        "oracle_map_find": (C.c_int, [vp, u32p, C.c_uint32, i32p]),
        "oracle_get_manifold_keys": (C.c_int, [vp, C.c_uint32, u32p, u32p]),
        "oracle_get_manifold_records": (C.c_int, [vp, C.c_uint32, u32p, C.c_void_p]),
        "oracle_get_iso_state": (C.c_int, [vp, u32p, u32p, fp, fp, u32p]),
        "oracle_splitmix64": (None, [C.c_uint64, C.c_uint32, C.POINTER(C.c_uint64), fp]),
        "oracle_snh_tet": (None, [dp, dp, C.c_double, C.c_double, C.c_double, C.c_int, dp, dp, dp]),
        "oracle_solve6x6": (None, [dp, dp, dp, C.c_int]),
        "oracle_cell_hash": (C.c_uint32, [C.c_float, C.c_float, C.c_float, C.c_float, C.c_uint32, C.c_uint32]),
        "oracle_get_cloth_info": (C.c_int, [vp, vp, u32p, u32p]),
        "oracle_get_collision_tris": (C.c_int, [vp, C.c_uint32, u32p, u32p]),
        "oracle_get_motor_targets": (C.c_int, [vp, C.c_uint32, C.POINTER(C.c_float)]),
        "oracle_get_dat_pairs": (C.c_int, [vp, C.c_uint32, u32p, u32p]),
        "oracle_get_soft_contacts": (C.c_int, [vp, C.c_uint32, u32p, vp]),
        "oracle_cloth_stats": (C.c_int, [vp, i32p, i32p, i32p, i32p, i32p, i32p]),
        "oracle_dat_stats": (C.c_int, [vp, i32p]),
        "oracle_set_dat_test_hooks": (C.c_int, [vp, C.c_float, C.c_int32, C.c_int32]),
        "oracle_set_joint_world_anchors": (C.c_int, [vp, C.POINTER(C.c_uint32), C.c_uint32, C.POINTER(C.c_float)]),
        "oracle_dat_pairs_at": (C.c_int, [vp, fp, C.c_uint32, u32p, u32p]),
        "oracle_membrane": (None, [dp, dp, C.c_double, C.c_double, C.c_int, dp, dp, dp]),
        "oracle_hinge": (C.c_int, [fp, C.c_float, C.c_float, fp, fp]),
        "oracle_cpu_create": (C.c_int, [vp, C.c_int, C.POINTER(vp)]),
        "oracle_cpu_destroy": (None, [vp]),
        "oracle_cpu_step": (C.c_int, [vp, C.c_int]),
        "oracle_cpu_get_body_states_f64": (C.c_int, [vp, C.c_uint32, dp, dp, dp, dp]),
        "oracle_cpu_num_manifolds": (C.c_int, [vp]),
        "oracle_cpu_max_constraint_error": (C.c_double, [vp]),
    }
    for k, (r, a) in sig.items():
        f = getattr(lib, k)
        f.restype = r
        f.argtypes = a
    _lib = lib
    return lib


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


class CpuPathSolver:
    """The reference's own CPU AVBD path (makeCPUSolver semantics: plain
    LDL^T, no trust region, O(n^2) broad phase, index-order Gauss-Seidel,
    lambda_max contact cap, diagonal hinge block, bilateral springs), FP32 or
    FP64.  Rigid boxes and spheres, joints and springs."""

    def __init__(self, scene, precision=32):
        self._lib = load()
        desc, self._keep = scene.to_desc()
        h = C.c_void_p()
        rc = self._lib.oracle_cpu_create(C.byref(desc), int(precision), C.byref(h))
        if rc != 0:
            raise ValueError(f"scene outside the CPU path ({rc})")
        self._h = h
        self.bodyCount = int(desc.num_bodies)

    def __del__(self):
        if getattr(self, "_h", None):
            self._lib.oracle_cpu_destroy(self._h)
            self._h = None

    def step(self, n=1):
        self._lib.oracle_cpu_step(self._h, int(n))

    def bodyStates64(self):
        n = self.bodyCount
        p, r, v, w = (np.zeros((n, 3)), np.zeros((n, 4)), np.zeros((n, 3)), np.zeros((n, 3)))
        self._lib.oracle_cpu_get_body_states_f64(self._h, n, _p(p, C.c_double), _p(r, C.c_double),
                                                 _p(v, C.c_double), _p(w, C.c_double))
        return p, r, v, w

    def positions(self):
        return self.bodyStates64()[0]

    def numManifolds(self):
        return self._lib.oracle_cpu_num_manifolds(self._h)

    def maxConstraintError(self):
        return self._lib.oracle_cpu_max_constraint_error(self._h)


class OracleSolver:
    """Same surface as avbd_amd.GPUSolver, CPU serial, FP32 or FP64."""

    def __init__(self, scene, precision=32, maxPairsPerBody=16):
        self._lib = load()
        self.scene = scene
        desc, self._keep = scene.to_desc()
        h = C.c_void_p()
        rc = self._lib.oracle_create(C.byref(desc), int(precision), int(maxPairsPerBody), C.byref(h))
        if rc != 0:
            raise RuntimeError(f"oracle_create failed ({rc})")
        self._h = h
        self.bodyCount = int(desc.num_bodies)
        self.precision = precision

    def __del__(self):
        if getattr(self, "_h", None):
            self._lib.oracle_destroy(self._h)
            self._h = None

    def step(self, n=1):
        for _ in range(n):
            self._lib.oracle_step(self._h)

    def stepToResidual(self, tol=1e-4, maxIterations=200, checkEvery=5):
        it = C.c_int()
        r = C.c_double()
        self._lib.oracle_step_to_residual(self._h, float(tol), int(maxIterations), int(checkEvery), C.byref(it),
                                          C.byref(r))
        return it.value, r.value

    def set_threads(self, n):
        """Split each color's primal over n threads (results identical to n = 1)."""
        self._lib.oracle_set_threads(self._h, int(n))

    def set_settings(self, settings):
        s = settings.to_c()
        self._lib.oracle_set_settings(self._h, C.byref(s))

    def bodyStates(self):
        n = self.bodyCount
        p, r, v, w = (np.zeros((n, 3), np.float32), np.zeros((n, 4), np.float32), np.zeros((n, 3), np.float32),
                      np.zeros((n, 3), np.float32))
        self._lib.oracle_get_body_states(self._h, n, _p(p, C.c_float), _p(r, C.c_float), _p(v, C.c_float),
                                         _p(w, C.c_float))
        return p, r, v, w

    def bodyStates64(self):
        n = self.bodyCount
        p, r, v, w = (np.zeros((n, 3)), np.zeros((n, 4)), np.zeros((n, 3)), np.zeros((n, 3)))
        self._lib.oracle_get_body_states_f64(self._h, n, _p(p, C.c_double), _p(r, C.c_double), _p(v, C.c_double),
                                             _p(w, C.c_double))
        return p, r, v, w

    def positions(self):
        return self.bodyStates()[0]

    def setBodyStates(self, idx, pos, rot, vel, angvel):
        idx = np.ascontiguousarray(idx, np.uint32)
        a = [np.ascontiguousarray(x, np.float32) for x in (pos, rot, vel, angvel)]
        self._lib.oracle_set_body_states(self._h, _p(idx, C.c_uint32), len(idx), *[_p(x, C.c_float) for x in a])

    def debugColors(self):
        c = np.zeros(self.bodyCount, np.uint32)
        self._lib.oracle_get_colors(self._h, self.bodyCount, _p(c, C.c_uint32))
        return c

    def cellLists(self):
        hs, nh, cs = C.c_uint32(), C.c_uint32(), C.c_float()
        self._lib.oracle_get_cell_lists(self._h, C.byref(hs), C.byref(nh), C.byref(cs), None, None, None)
        start, count = np.zeros(hs.value, np.uint32), np.zeros(hs.value, np.uint32)
        bodies = np.zeros(max(1, nh.value), np.uint32)
        self._lib.oracle_get_cell_lists(self._h, None, None, None, _p(start, C.c_uint32), _p(count, C.c_uint32),
                                        _p(bodies, C.c_uint32))
        return dict(hash_size=hs.value, cell_size=cs.value, start=start, count=count, bodies=bodies[:nh.value])

    def pairs(self):
        n = C.c_uint32()
        self._lib.oracle_get_pairs(self._h, 0, C.byref(n), None)
        out = np.zeros((max(1, n.value), 2), np.uint32)
        self._lib.oracle_get_pairs(self._h, n.value, C.byref(n), _p(out, C.c_uint32))
        return out[:n.value]

    def manifolds(self):
        n = C.c_uint32()
        self._lib.oracle_get_manifolds(self._h, 0, C.byref(n), None, None, None)
        h = np.zeros((max(1, n.value), 3), np.uint32)
        nrm = np.zeros((max(1, n.value), 3), np.float32)
        lp = np.zeros((max(1, n.value), 2), np.float32)
        self._lib.oracle_get_manifolds(self._h, n.value, C.byref(n), _p(h, C.c_uint32), _p(nrm, C.c_float),
                                       _p(lp, C.c_float))
        return h[:n.value], nrm[:n.value], lp[:n.value]

    # the 704-byte manifold record: header (a, b, count, flags), collider
    # pair, (normal, dynamic friction), (t1, static friction), 8 contacts
    MANIFOLD_DTYPE = np.dtype([("header", np.uint32, 4), ("colliderPair", np.uint32, 4), ("basisN", np.float32, 4),
                               ("basisT1", np.float32, 4),
                               ("c", [("rA", np.float32, 4), ("rB", np.float32, 4), ("C0", np.float32, 4),
                                      ("lambda", np.float32, 4), ("penalty", np.float32, 4)], 8)])

    def manifoldRecords(self):
        """The last frame's manifolds in the backend's record layout
        (rA.w holds the feature key bits, rB.w the stick flag)."""
        n = C.c_uint32()
        self._lib.oracle_get_manifold_records(self._h, 0, C.byref(n), None)
        out = np.zeros(max(1, n.value), self.MANIFOLD_DTYPE)
        self._lib.oracle_get_manifold_records(self._h, n.value, C.byref(n), out.ctypes.data_as(C.c_void_p))
        return out[:n.value]

    def manifoldKeys(self):
        """Feature keys of the last frame's contacts: (n, 8), 0xFFFFFFFF past each count."""
        n = C.c_uint32()
        self._lib.oracle_get_manifold_keys(self._h, 0, C.byref(n), None)
        out = np.zeros((max(1, n.value), 8), np.uint32)
        self._lib.oracle_get_manifold_keys(self._h, n.value, C.byref(n), _p(out, C.c_uint32))
        return out[:n.value]

    def isoState(self):
        """Isotropic limiter state after the last frame: tracked triangles
        (nb, 4), tracked edges (num_contact_edges, 4), squared clearance
        bounds (nb,), anchor poses (nb, 3), the frame's truncations and
        refreshes; None when the mode is off."""
        info = self.clothInfo()
        nb, ne = self.bodyCount, info["num_contact_edges"]
        vt = np.zeros((max(1, nb), 4), np.uint32)
        ee = np.zeros((max(1, ne), 4), np.uint32)
        bound = np.zeros(max(1, nb), np.float32)
        anchor = np.zeros((max(1, nb), 3), np.float32)
        counts = np.zeros(2, np.uint32)
        if self._lib.oracle_get_iso_state(self._h, _p(vt, C.c_uint32), _p(ee, C.c_uint32), _p(bound, C.c_float),
                                          _p(anchor, C.c_float), _p(counts, C.c_uint32)) != 0:
            return None
        return {"vt_track": vt[:nb], "ee_track": ee[:ne], "bound": bound[:nb], "anchor": anchor[:nb],
                "truncations": int(counts[0]), "refreshes": int(counts[1])}

    def mapFind(self, pairs):
        q = np.ascontiguousarray(pairs, np.uint32).reshape(-1, 2)
        out = np.zeros(len(q), np.int32)
        self._lib.oracle_map_find(self._h, _p(q, C.c_uint32), len(q), _p(out, C.c_int32))
        return out

    def motorTargets(self, n):
        t = np.zeros(max(1, n), np.float32)
        self._lib.oracle_get_motor_targets(self._h, n, _p(t, C.c_float))
        return t[:n]

    def collisionTris(self):
        n = C.c_uint32(0)
        self._lib.oracle_get_collision_tris(self._h, 0, C.byref(n), None)
        t = np.zeros((max(1, n.value), 3), np.uint32)
        self._lib.oracle_get_collision_tris(self._h, n.value, C.byref(n), _p(t, C.c_uint32))
        return t[:n.value]

    def clothInfo(self):
        import sys
        sys.path.insert(0, os.path.join(os.path.dirname(_HERE), "avbd-metal_amd"))
        from avbd_amd import _capi
        info = _capi.ClothInfo()
        self._lib.oracle_get_cloth_info(self._h, C.byref(info), None, None)
        edges = np.zeros((max(1, info.num_edges), 2), np.uint32)
        parts = np.zeros(max(1, info.num_particles), np.uint32)
        self._lib.oracle_get_cloth_info(self._h, None, _p(edges, C.c_uint32), _p(parts, C.c_uint32))
        out = {f: getattr(info, f) for f, _ in _capi.ClothInfo._fields_}
        out["edges"] = edges[:info.num_edges]
        out["particles"] = parts[:info.num_particles]
        return out

    def datPairs(self):
        n = C.c_uint32()
        self._lib.oracle_get_dat_pairs(self._h, 0, C.byref(n), None)
        out = np.zeros((max(1, n.value), 3), np.uint32)
        self._lib.oracle_get_dat_pairs(self._h, n.value, C.byref(n), _p(out, C.c_uint32))
        return out[:n.value]

    SOFT_DTYPE = np.dtype([("ids", np.uint32, 4), ("normal", np.float32, 4), ("anchorA", np.float32, 4),
                           ("weights", np.float32, 4), ("C0", np.float32, 4), ("lambda", np.float32, 4),
                           ("penalty", np.float32, 4)])

    def softContacts(self):
        n = C.c_uint32()
        self._lib.oracle_get_soft_contacts(self._h, 0, C.byref(n), None)
        out = np.zeros(max(1, n.value), self.SOFT_DTYPE)
        self._lib.oracle_get_soft_contacts(self._h, n.value, C.byref(n), out.ctypes.data_as(C.c_void_p))
        return out[:n.value]

    def datPairsAt(self, positions):
        """Pair stream the element grid yields at the given positions (state unchanged)."""
        p = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
        n = C.c_uint32()
        self._lib.oracle_dat_pairs_at(self._h, _p(p, C.c_float), 0, C.byref(n), None)
        out = np.zeros((max(1, n.value), 3), np.uint32)
        self._lib.oracle_dat_pairs_at(self._h, _p(p, C.c_float), n.value, C.byref(n), _p(out, C.c_uint32))
        return out[:n.value]

    def clothStats(self):
        v = [C.c_int32() for _ in range(6)]
        self._lib.oracle_cloth_stats(self._h, *[C.byref(x) for x in v])
        return dict(zip(["dat_pairs", "dat_vt_pairs", "soft_candidates", "num_soft_rt", "overflow", "nonfinite"],
                        [x.value for x in v]))

    def setJointWorldAnchors(self, joints, points):
        j = np.ascontiguousarray(joints, np.uint32)
        p = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
        self._lib.oracle_set_joint_world_anchors(self._h, _p(j, C.c_uint32), len(j), _p(p, C.c_float))

    def setDatTestHooks(self, queryRadius=0.0, pairCapacity=0, softCapacity=0):
        """Planar-DAT query radius override and pair / soft capacity caps (0 keeps)."""
        self._lib.oracle_set_dat_test_hooks(self._h, float(queryRadius), int(pairCapacity), int(softCapacity))

    def datStats(self):
        v = (C.c_int32 * 6)()
        self._lib.oracle_dat_stats(self._h, v)
        return dict(zip(["dat_truncations", "dat_invalid_anchors", "dat_vt_degeneracies", "dat_ee_degeneracies",
                         "dat_nonfinite", "overflow"], list(v)))

    def residual(self):
        v = C.c_double()
        self._lib.oracle_residual(self._h, C.byref(v))
        return v.value

    def stats(self):
        v = [C.c_int32() for _ in range(5)]
        self._lib.oracle_stats(self._h, *[C.byref(x) for x in v])
        return dict(num_pairs=v[0].value, active_manifolds=v[1].value, num_colors=v[2].value,
                    pair_candidates=v[3].value, color_conflicts=v[4].value)


def splitmix64(seed, n):
    lib = load()
    u = np.zeros(n, np.uint64)
    f = np.zeros(n, np.float32)
    lib.oracle_splitmix64(seed, n, _p(u, C.c_uint64), _p(f, C.c_float))
    return u, f


def snh_tet(x, dminv, mu, lam, vol, slot):
    lib = load()
    x = np.ascontiguousarray(x, np.float64).reshape(12)
    dm = np.ascontiguousarray(dminv, np.float64).reshape(9)
    e = C.c_double()
    g = np.zeros(3)
    h = np.zeros(9)
    lib.oracle_snh_tet(_p(x, C.c_double), _p(dm, C.c_double), mu, lam, vol, int(slot), C.byref(e),
                       _p(g, C.c_double), _p(h, C.c_double))
    return e.value, g, h.reshape(3, 3)


def solve6x6(A, b, precision=64):
    lib = load()
    A = np.ascontiguousarray(A, np.float64).reshape(36)
    b = np.ascontiguousarray(b, np.float64).reshape(6)
    x = np.zeros(6)
    lib.oracle_solve6x6(_p(A, C.c_double), _p(b, C.c_double), _p(x, C.c_double), int(precision))
    return x


def cell_hash(p, cell_size, group, hash_size):
    return load().oracle_cell_hash(float(p[0]), float(p[1]), float(p[2]), float(cell_size), int(group),
                                   int(hash_size))


def membrane(x, dm, muA, lamA, slot):
    """StVK membrane energy, gradient on `slot` and its Gauss-Newton block (double)."""
    lib = load()
    x = np.ascontiguousarray(x, np.float64).reshape(9)
    dm = np.ascontiguousarray(dm, np.float64).reshape(4)
    e = C.c_double()
    g = np.zeros(3)
    h = np.zeros(9)
    lib.oracle_membrane(_p(x, C.c_double), _p(dm, C.c_double), float(muA), float(lamA), int(slot), C.byref(e),
                        _p(g, C.c_double), _p(h, C.c_double))
    return e.value, g, h.reshape(3, 3)


def hinge(x, k1, k2):
    """Hinge constants (K[4], s) the upload derives for two triangles sharing
    the edge x[0]-x[1], opposite vertices x[2], x[3]; None when skipped."""
    lib = load()
    x = np.ascontiguousarray(x, np.float32).reshape(12)
    K = np.zeros(4, np.float32)
    s = C.c_float()
    ok = lib.oracle_hinge(_p(x, C.c_float), float(k1), float(k2), _p(K, C.c_float), C.byref(s))
    return (K, s.value) if ok else None
