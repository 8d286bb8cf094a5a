"""GPUSolver mirror over the C ABI (Sources/PhysicsAVBD/GPU/GPUSolver.swift).

Method names follow the reference's Swift surface (step, submitStep,
synchronize, bodyPosition, bodyStates, setBodyStates, debugColors,
maxConstraintError, lastNumPairs ...) so tests read like the reference's own.
"""
import ctypes as C

import numpy as np

from . import _capi
from .scene import SimSettings


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _u32p(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


class GPUSolver:
    def __init__(self, scene, maxPairsPerBody=16, lanesPerBody=0, sweepMode=0, useGraph=True,
                 pairCapacityForTesting=0, planarDATPairCapacityForTesting=0, planarDATSoftCapacityForTesting=0):
        self._lib = _capi.load()
        self.scene = scene
        desc, keep = scene.to_desc()
        opts = _capi.Options()
        self._lib.avbd_options_default(C.byref(opts))
        opts.max_pairs_per_body = int(maxPairsPerBody)
        opts.lanes_per_body = int(lanesPerBody)
        opts.sweep_mode = int(sweepMode)
        opts.use_graph = 1 if useGraph else 0
        opts.pair_capacity_override = int(pairCapacityForTesting)
        opts.dat_pair_capacity_override = int(planarDATPairCapacityForTesting)
        opts.soft_capacity_override = int(planarDATSoftCapacityForTesting)
        h = C.c_void_p()
        st = self._lib.avbd_create(C.byref(desc), C.byref(opts), C.byref(h))
        self._h = h
        self.bodyCount = int(desc.num_bodies)
        if st == _capi.ERR_RUNTIME_FAILURE and h:
            f = self._failure()
            self._lib.avbd_destroy(h)
            self._h = None
            raise f
        if st != _capi.OK:
            self._h = None
            raise _capi.AVBDError(st, "avbd_create")
        _capi.track("solver", self)

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def close(self):
        if getattr(self, "_h", None):
            self._lib.avbd_destroy(self._h)
            self._h = None

    # -- failures ---------------------------------------------------------------
    def _failure(self):
        f = _capi.Failure()
        self._lib.avbd_last_failure(self._h, C.byref(f))
        return _capi.RuntimeFailure(f.kind, f.frame, f.required, f.capacity, f.message.decode())

    def _check(self, st, what):
        if st == _capi.OK:
            return
        if st == _capi.ERR_RUNTIME_FAILURE:
            raise self._failure()
        raise _capi.AVBDError(st, what)

    @property
    def runtimeFailure(self):
        f = _capi.Failure()
        self._lib.avbd_last_failure(self._h, C.byref(f))
        return None if f.kind == 0 else self._failure()

    # -- stepping -----------------------------------------------------------------
    def submitStep(self):
        self._check(self._lib.avbd_submit_step(self._h), "submitStep")

    def step(self, n=1):
        for _ in range(n):
            self._check(self._lib.avbd_step(self._h), "step")

    def stepToResidual(self, tol=1e-4, maxIterations=200, checkEvery=5):
        """One frame iterated until residual <= tol (at most maxIterations
        sweeps, checked every checkEvery); returns (iterations, residual)."""
        it = C.c_int32()
        r = C.c_float()
        self._check(self._lib.avbd_step_to_residual(self._h, float(tol), int(maxIterations), int(checkEvery),
                                                     C.byref(it), C.byref(r)), "stepToResidual")
        return it.value, r.value

    def synchronize(self):
        self._check(self._lib.avbd_synchronize(self._h), "synchronize")

    @property
    def settings(self):
        s = _capi.Settings()
        self._lib.avbd_get_settings(self._h, C.byref(s))
        return SimSettings.from_c(s)

    @settings.setter
    def settings(self, value):
        s = value.to_c()
        self._check(self._lib.avbd_set_settings(self._h, C.byref(s)), "settings")

    # -- state --------------------------------------------------------------------
    def bodyStates(self, idx=None):
        """Returns (pos[n,3], rot[n,4], vel[n,3], angvel[n,3]) as float32 arrays."""
        if idx is None:
            n = self.bodyCount
            ip = None
        else:
            idx = np.ascontiguousarray(idx, dtype=np.uint32)
            n = len(idx)
            ip = _u32p(idx)
        p = np.zeros((n, 3), np.float32)
        r = np.zeros((n, 4), np.float32)
        v = np.zeros((n, 3), np.float32)
        w = np.zeros((n, 3), np.float32)
        self._check(self._lib.avbd_get_body_states(self._h, ip, n, _fp(p), _fp(r), _fp(v), _fp(w)), "bodyStates")
        return p, r, v, w

    def positions(self):
        return self.bodyStates()[0]

    def bodyPosition(self, i):
        return self.bodyStates([i])[0][0]

    def bodyRotation(self, i):
        return self.bodyStates([i])[1][0]

    def bodyVelocity(self, i):
        return self.bodyStates([i])[2][0]

    def bodyAngularVelocity(self, i):
        return self.bodyStates([i])[3][0]

    def bodyMass(self, i=None):
        m = np.zeros(self.bodyCount, np.float32)
        inertia = np.zeros((self.bodyCount, 3), np.float32)
        self._check(self._lib.avbd_get_body_mass(self._h, self.bodyCount, _fp(m), _fp(inertia)), "bodyMass")
        return m if i is None else m[i]

    def bodyDiagonalInertia(self, i=None):
        """bodyDiagonalInertia (GPUSolver.swift:4461): principal moments."""
        m = np.zeros(self.bodyCount, np.float32)
        inertia = np.zeros((self.bodyCount, 3), np.float32)
        self._check(self._lib.avbd_get_body_mass(self._h, self.bodyCount, _fp(m), _fp(inertia)),
                    "bodyDiagonalInertia")
        return inertia if i is None else inertia[i]

    def setBodyStates(self, idx, pos, rot, vel, angvel):
        idx = np.ascontiguousarray(idx, dtype=np.uint32)
        pos = np.ascontiguousarray(pos, dtype=np.float32)
        rot = np.ascontiguousarray(rot, dtype=np.float32)
        vel = np.ascontiguousarray(vel, dtype=np.float32)
        angvel = np.ascontiguousarray(angvel, dtype=np.float32)
        self._check(self._lib.avbd_set_body_states(self._h, _u32p(idx), len(idx), _fp(pos), _fp(rot), _fp(vel),
                                                   _fp(angvel)), "setBodyStates")

    def setMotorTargets(self, joints, targets):
        j = np.ascontiguousarray(joints, dtype=np.uint32)
        t = np.ascontiguousarray(targets, dtype=np.float32)
        self._check(self._lib.avbd_set_motor_targets(self._h, _u32p(j), len(j), _fp(t)), "setMotorTargets")

    def setMotorVelocities(self, joints, radiansPerSecond):
        j = np.ascontiguousarray(joints, dtype=np.uint32)
        v = np.ascontiguousarray(radiansPerSecond, dtype=np.float32)
        self._check(self._lib.avbd_set_motor_velocities(self._h, _u32p(j), len(j), _fp(v)), "setMotorVelocities")

    def setJointWorldAnchors(self, joints, points):
        j = np.ascontiguousarray(joints, dtype=np.uint32)
        p = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
        self._check(self._lib.avbd_set_joint_world_anchors(self._h, _u32p(j), len(j), _fp(p)), "setJointWorldAnchors")

    def setBodyPoses(self, idx, pos, rot):
        i = np.ascontiguousarray(idx, dtype=np.uint32)
        p = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 3)
        r = np.ascontiguousarray(rot, dtype=np.float32).reshape(-1, 4)
        self._check(self._lib.avbd_set_body_poses(self._h, _u32p(i), len(i), _fp(p), _fp(r)), "setBodyPoses")

    def motorStates(self, joints):
        """[(angle, velocity)] of hinge joints (GPUSolver.motorStates)."""
        j = np.ascontiguousarray(joints, dtype=np.uint32)
        a = np.zeros(len(j), np.float32)
        v = np.zeros(len(j), np.float32)
        self._check(self._lib.avbd_get_motor_states(self._h, _u32p(j), len(j), _fp(a), _fp(v)), "motorStates")
        return list(zip(a.tolist(), v.tolist()))

    def motorAngles(self, joints):
        return [a for a, _ in self.motorStates(joints)]

    def debugBrokenJoints(self):
        n = C.c_int32()
        self._check(self._lib.avbd_get_joint_diagnostics(self._h, C.byref(n), None, None), "debugBrokenJoints")
        return n.value

    def debugMaxLambda(self):
        a, b = C.c_float(), C.c_float()
        self._check(self._lib.avbd_get_joint_diagnostics(self._h, None, C.byref(a), C.byref(b)), "debugMaxLambda")
        return a.value, b.value

    def setMotorTorques(self, joints, torques):
        j = np.ascontiguousarray(joints, dtype=np.uint32)
        t = np.ascontiguousarray(torques, dtype=np.float32)
        self._check(self._lib.avbd_set_motor_torques(self._h, _u32p(j), len(j), _fp(t)), "setMotorTorques")

    # -- stats / debug --------------------------------------------------------------
    def stats(self):
        s = _capi.Stats()
        self._check(self._lib.avbd_get_stats(self._h, C.byref(s)), "stats")
        return s

    @property
    def lastNumPairs(self):
        return self.stats().num_pairs

    @property
    def lastPairCandidates(self):
        return self.stats().pair_candidates

    @property
    def lastColorCounts(self):
        return list(self.stats().color_counts)

    @property
    def lastMaxColorUsed(self):
        return self.stats().max_color_used

    @property
    def frameIndex(self):
        return self.stats().frame_index

    def debugColors(self):
        c = np.zeros(self.bodyCount, np.uint32)
        self._check(self._lib.avbd_get_colors(self._h, self.bodyCount, _u32p(c)), "debugColors")
        return c

    def cellLists(self):
        hs, nh, cs = C.c_uint32(), C.c_uint32(), C.c_float()
        self._check(self._lib.avbd_get_cell_lists(self._h, C.byref(hs), C.byref(nh), C.byref(cs), None, None, None),
                    "cellLists")
        start = np.zeros(hs.value, np.uint32)
        count = np.zeros(hs.value, np.uint32)
        bodies = np.zeros(max(1, nh.value), np.uint32)
        self._check(self._lib.avbd_get_cell_lists(self._h, None, None, None, _u32p(start), _u32p(count),
                                                  _u32p(bodies)), "cellLists")
        return dict(hash_size=hs.value, cell_size=cs.value, start=start, count=count, bodies=bodies[:nh.value])

    def pairs(self):
        n = C.c_uint32()
        self._check(self._lib.avbd_get_pairs(self._h, 0, C.byref(n), None), "pairs")
        out = np.zeros((max(1, n.value), 2), np.uint32)
        self._check(self._lib.avbd_get_pairs(self._h, n.value, C.byref(n), _u32p(out)), "pairs")
        return out[:n.value]

    MANIFOLD_DTYPE = np.dtype([("header", np.uint32, 4), ("colliderPair", np.uint32, 4), ("basisN", np.float32, 4),
                               ("basisT1", np.float32, 4),
                               ("c", [("rA", np.float32, 4), ("rB", np.float32, 4), ("C0", np.float32, 4),
                                      ("lambda", np.float32, 4), ("penalty", np.float32, 4)], 8)])

    def manifolds(self):
        n = C.c_uint32()
        self._check(self._lib.avbd_get_manifolds(self._h, 0, C.byref(n), None), "manifolds")
        out = np.zeros(max(1, n.value), self.MANIFOLD_DTYPE)
        self._check(self._lib.avbd_get_manifolds(self._h, n.value, C.byref(n), out.ctypes.data_as(C.c_void_p)),
                    "manifolds")
        return out[:n.value]

    SOFT_DTYPE = np.dtype([("ids", np.uint32, 4), ("normal", np.float32, 4), ("anchorA", np.float32, 4),
                           ("weights", np.float32, 4), ("C0", np.float32, 4), ("lambda", np.float32, 4),
                           ("penalty", np.float32, 4)])

    def clothInfo(self):
        """Surface topology and parameters derived at create (GPUSolver.swift:1289-1690)."""
        info = _capi.ClothInfo()
        self._check(self._lib.avbd_get_cloth_info(self._h, C.byref(info), None, None), "cloth info")
        edges = np.zeros((max(1, info.num_edges), 2), np.uint32)
        parts = np.zeros(max(1, info.num_particles), np.uint32)
        self._check(self._lib.avbd_get_cloth_info(self._h, None, _u32p(edges), _u32p(parts)), "cloth info")
        out = {f: getattr(info, f) for f, _ in _capi.ClothInfo._fields_}
        out["edges"] = edges[:info.num_edges]
        out["particles"] = parts[:info.num_particles]
        return out

    def motorTargets(self, n):
        """motorTargetForTesting: the motor targets of joints [0, n)."""
        t = np.zeros(max(1, n), np.float32)
        self._check(self._lib.avbd_get_motor_targets(self._h, n, _fp(t)), "motor targets")
        return t[:n]

    def collisionTris(self):
        """Collision triangles: the authored ones, then the tet boundary faces (GPUSolver.swift:872-897)."""
        n = C.c_uint32(0)
        self._check(self._lib.avbd_get_collision_tris(self._h, 0, C.byref(n), None), "collision tris")
        t = np.zeros((max(1, n.value), 3), np.uint32)
        self._check(self._lib.avbd_get_collision_tris(self._h, n.value, C.byref(n), _u32p(t)), "collision tris")
        return t[:n.value]

    def datPairs(self):
        """Planar-DAT pair stream of the last frame: rows (kind 0 V-T / 1 E-E, owner, other)."""
        n = C.c_uint32()
        self._check(self._lib.avbd_get_dat_pairs(self._h, 0, C.byref(n), None), "dat pairs")
        out = np.zeros((max(1, n.value), 3), np.uint32)
        self._check(self._lib.avbd_get_dat_pairs(self._h, n.value, C.byref(n), _u32p(out)), "dat pairs")
        return out[:n.value]

    def softContacts(self):
        """Element contacts (V-T, rigid-T, E-E) of the last frame, SoftContactGPU records."""
        n = C.c_uint32()
        self._check(self._lib.avbd_get_soft_contacts(self._h, 0, C.byref(n), None), "soft contacts")
        out = np.zeros(max(1, n.value), self.SOFT_DTYPE)
        self._check(self._lib.avbd_get_soft_contacts(self._h, n.value, C.byref(n), out.ctypes.data_as(C.c_void_p)),
                    "soft contacts")
        return out[:n.value]

    def createProfile(self):
        """[(phase, ms)] of this solver's creation, in order."""
        out, i = [], 0
        name, ms = C.c_char_p(), C.c_float()
        while self._lib.avbd_get_create_profile(self._h, i, C.byref(name), C.byref(ms)) == 0:
            out.append((name.value.decode(), ms.value))
            i += 1
        return out

    def isoState(self):
        """Isotropic limiter state after the last frame: tracked triangles
        (nb, 4), tracked edges (num_contact_edges, 4), squared clearance
        bounds (nb,), anchor poses (nb, 3), the frame's truncations and
        refreshes; None when the mode is off."""
        info = self.clothInfo()
        if not info["iso_on"]:
            return None
        nb, ne = self.bodyCount, info["num_contact_edges"]
        vt = np.zeros((max(1, nb), 4), np.uint32)
        ee = np.zeros((max(1, ne), 4), np.uint32)
        bound = np.zeros(max(1, nb), np.float32)
        anchor = np.zeros((max(1, nb), 3), np.float32)
        counts = np.zeros(2, np.uint32)
        self._check(self._lib.avbd_get_iso_state(self._h, _u32p(vt), _u32p(ee), _fp(bound), _fp(anchor), _u32p(counts)),
                    "iso state")
        return {"vt_track": vt[:nb], "ee_track": ee[:ne], "bound": bound[:nb], "anchor": anchor[:nb],
                "truncations": int(counts[0]), "refreshes": int(counts[1])}

    def mapFind(self, pairs):
        q = np.ascontiguousarray(pairs, dtype=np.uint32).reshape(-1, 2)
        out = np.zeros(len(q), np.int32)
        self._check(self._lib.avbd_map_find(self._h, _u32p(q), len(q), out.ctypes.data_as(C.POINTER(C.c_int32))),
                    "mapFind")
        return out

    def softMapFind(self, keys):
        """Element-contact persistence map: (keyA, keyB) rows -> index of the
        last frame's element contact with that key (softContacts order), or -1."""
        q = np.ascontiguousarray(keys, dtype=np.uint32).reshape(-1, 2)
        out = np.zeros(len(q), np.int32)
        self._check(self._lib.avbd_soft_map_find(self._h, _u32p(q), len(q),
                                                 out.ctypes.data_as(C.POINTER(C.c_int32))), "softMapFind")
        return out

    def maxConstraintError(self):
        v = C.c_float()
        self._check(self._lib.avbd_max_constraint_error(self._h, C.byref(v)), "maxConstraintError")
        return v.value

    def residual(self):
        v = C.c_float()
        self._check(self._lib.avbd_residual(self._h, C.byref(v)), "residual")
        return v.value

    def timePrimal(self, reps=20):
        us, by, la = C.c_float(), C.c_double(), C.c_int32()
        self._check(self._lib.avbd_time_primal(self._h, int(reps), C.byref(us), C.byref(by), C.byref(la)),
                    "timePrimal")
        return dict(avg_launch_us=us.value, sweep_bytes=by.value, launches_per_sweep=la.value)

    # -- learned warm start (no reference counterpart) ------------------------
    def setWarmStart(self, mlp=None, learned=True, log=False, force=False, smooth=2, rigid=None):
        """Attach WarmStartMLP networks: `mlp` (3 outputs) serves the dynamic
        particles, `rigid` (6 outputs) the dynamic rigid bodies.  learned=True
        replaces their adaptive guess with guess + |g| dt^2 * MLP(features)
        (particle predictions smoothed by `smooth` 1-ring passes), keeping the
        plain guess whenever the learned start's residual is higher
        (force=True keeps it unconditionally); log=True records each step's
        features and converged corrections (warmStartSamples).  No arguments
        detach."""
        nets = [n for n in (mlp, rigid) if n is not None]
        flags = (_capi.WS_LEARNED if (learned and nets) else 0) | (_capi.WS_LOG if log else 0)
        if flags & _capi.WS_LEARNED and force:
            flags |= _capi.WS_FORCE
        if flags & _capi.WS_LEARNED:
            flags |= (int(smooth) & 0xF) << _capi.WS_SMOOTH_SHIFT
        self._check(self._lib.avbd_set_warm_start(self._h, None, 0), "setWarmStart")   # detach previous
        if flags & _capi.WS_LEARNED:
            for n in nets:
                self._check(self._lib.avbd_set_warm_start(self._h, n._h, flags), "setWarmStart")
        elif flags:
            self._check(self._lib.avbd_set_warm_start(self._h, None, flags), "setWarmStart")
        self._warm_nets = nets if flags & _capi.WS_LEARNED else []   # keeps the networks alive while attached

    def warmStartSamples(self, rigid=False):
        """(features [n,24], corrections [n,3] or [n,6] for rigid=True, body ids [n]) of the last frame."""
        fn = self._lib.avbd_get_warm_start_samples_rigid if rigid else self._lib.avbd_get_warm_start_samples
        w = _capi.MLP_OUTPUTS_RIGID if rigid else _capi.MLP_OUTPUTS
        n = C.c_uint32()
        self._check(fn(self._h, 0, C.byref(n), None, None, None), "warmStartSamples")
        x = np.zeros((n.value, _capi.MLP_INPUTS), np.float32)
        t = np.zeros((n.value, w), np.float32)
        idx = np.zeros(n.value, np.uint32)
        self._check(fn(self._h, n.value, C.byref(n), _fp(x), _fp(t), _u32p(idx)), "warmStartSamples")
        return x, t, idx

    def setProfiling(self, on=True):
        self._check(self._lib.avbd_get_profile(self._h, 1 if on else 0, None, 0, None, None), "profile")

    def profile(self):
        names = C.create_string_buffer(1024)
        ms = np.zeros(32, np.float32)
        n = C.c_uint32()
        self._check(self._lib.avbd_get_profile(self._h, 1, names, 1024, _fp(ms), C.byref(n)), "profile")
        keys = [k for k in names.value.decode().split(",") if k]
        return dict(zip(keys, ms[:n.value].tolist()))

    def failCompletionForTesting(self, kind, frame=-1):
        """completionFailureForTesting: frame `frame` completes with failure
        `kind` (a RuntimeFailure kind name such as "commandExecution")."""
        codes = {v: k for k, v in _capi.FAILURE_NAMES.items()}
        self._check(self._lib.avbd_test_fail_completion(self._h, codes[kind], int(frame)), "failCompletionForTesting")

    # -- speculation snapshots (GPUSolver.swift:2948-3057) ---------------------
    def captureRigidSpeculationSnapshot(self, into=None):
        """Device-resident snapshot of the rigid scene's temporal state;
        `into` (a snapshot of this solver) is overwritten in place."""
        h = C.c_void_p(into._h.value if into is not None else None)
        self._check(self._lib.avbd_capture_speculation_snapshot(self._h, C.byref(h)),
                    "captureRigidSpeculationSnapshot")
        if into is not None:
            return into
        return RigidSpeculationSnapshot(self, h)

    def restoreRigidSpeculationSnapshot(self, snapshot):
        self._check(self._lib.avbd_restore_speculation_snapshot(self._h, snapshot._h),
                    "restoreRigidSpeculationSnapshot")

    def setDATQueryRadiusForTesting(self, r):
        """params.planarDATQueryRadius = r (PlanarDATTests.swift:930); r <= 0 restores."""
        self._check(self._lib.avbd_test_set_dat_query_radius(self._h, float(r)), "setDATQueryRadiusForTesting")

    def denyStageForTesting(self, stage, frame=-1):
        self._lib.avbd_test_deny_stage(self._h, stage.encode() if stage else None, int(frame))


class RigidSpeculationSnapshot:
    """Opaque snapshot handle; keeps its solver alive while it exists."""

    def __init__(self, solver, handle):
        self._solver = solver
        self._h = handle
        self._lib = solver._lib
        _capi.track("snapshot", self)

    def __del__(self):
        self.close()

    def close(self):
        if getattr(self, "_h", None):
            self._lib.avbd_destroy_speculation_snapshot(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def version():
    return _capi.load().avbd_version().decode()
