"""PhysicsScene / SimSettings mirror (Sources/SimCore/Scenes/Scene.swift).

Builds the packed SoA scene wire format of include/avbd.h.  Builder methods
keep the reference's names and defaults (addBody, addParticle, addTet,
addJoint, addSpring, addCollisionExclusion) so scene code reads the same.
Float arithmetic that the reference performs in Float (particle density,
sphere size normalisation) is done in float32 here too.
"""
import ctypes as C
import itertools
import math
import os
from dataclasses import dataclass, field

import numpy as np

from . import _capi

BOX, SPHERE, TORUS, CAPSULE = 0, 1, 2, 3
_F = np.float32


@dataclass
class SimSettings:
    """SimSettings, Scene.swift:485-545 (same defaults)."""
    dt: float = 1.0 / 60.0
    gravity: float = -10.0
    iterations: int = 10
    collisionMargin: float = 0.01
    frictionCombineMode: int = 0
    alpha: float = 0.99
    betaLin: float = 10000.0
    betaAng: float = 100.0
    gamma: float = 0.999
    lambdaMax: float = 1.0e6
    rodDecayPow: float = 0.0
    particleDamping: float = 0.0
    rigidLinearDamping: float = 0.0
    rigidAngularDamping: float = 0.0
    clothViscosity: float = 0.0

    def to_c(self):
        return _capi.Settings(self.dt, self.gravity, int(self.iterations), self.collisionMargin,
                              int(self.frictionCombineMode), self.alpha, self.betaLin, self.betaAng,
                              self.gamma, self.lambdaMax, self.rodDecayPow, self.particleDamping,
                              self.rigidLinearDamping, self.rigidAngularDamping, self.clothViscosity)

    @staticmethod
    def from_c(s):
        return SimSettings(s.dt, s.gravity, s.iterations, s.collision_margin, s.friction_combine_mode,
                           s.alpha, s.beta_lin, s.beta_ang, s.gamma, s.lambda_max, s.rod_decay_pow,
                           s.particle_damping, s.rigid_linear_damping, s.rigid_angular_damping,
                           s.cloth_viscosity)


@dataclass
class SceneJoint:
    """SceneJoint, Scene.swift:206-299 (motor modes: 0 implicit PD, 1 explicit PD, 2 velocity)."""
    bodyA: int
    bodyB: int
    rA: tuple
    rB: tuple
    stiffnessLin: float = math.inf
    stiffnessAng: float = 0.0
    fracture: float = math.inf
    fractureLinear: bool = False
    hingeAxis: tuple = None
    motorTarget: float = 0.0
    motorTorque: float = 0.0
    motorStiffness: float = 0.0
    motorDamping: float = 0.0
    motorMode: int = 0
    armature: float = 0.0
    limitLo: float = 1.0
    limitHi: float = -1.0
    motorRate: float = 0.0   # legacy servo rate (rad/s) of an implicit-PD motor (Scene.swift:300-314)


@dataclass
class PhysicsScene:
    name: str = "scene"
    settings: SimSettings = field(default_factory=SimSettings)

    def __post_init__(self):
        self.bodies = []       # dicts
        self.colliders = []
        self.joints = []
        self.springs = []
        self.tets = []
        self.tris = []
        self.exclusions = []
        self.spinners = []     # (body, unit axis, omega)

    # -- Scene.swift:566-589 -------------------------------------------------
    def addBody(self, size, density, friction, position, rotation=(0, 0, 0, 1), velocity=(0, 0, 0),
                shape=BOX, dynamicFriction=None, mass=None, diagonalInertia=None, gravityScale=1.0,
                collisionGroup=0, collisionEnabled=True):
        assert (mass is None) == (diagonalInertia is None), "explicit mass and inertia come together"
        size = tuple(float(x) for x in size)
        if shape == SPHERE:
            size = (size[0], size[0], size[0])
        df = friction if dynamicFriction is None else dynamicFriction
        self.bodies.append(dict(size=size, density=float(density), friction=float(friction),
                                dynamicFriction=float(df), position=tuple(position),
                                rotation=tuple(rotation), velocity=tuple(velocity), shape=shape,
                                mass=mass, inertia=diagonalInertia, gravityScale=float(gravityScale),
                                particle=False))
        b = len(self.bodies) - 1
        if collisionEnabled:
            self.colliders.append(dict(body=b, size=size, friction=float(friction), dynamicFriction=float(df),
                                       localPosition=(0, 0, 0), localRotation=(0, 0, 0, 1), shape=shape,
                                       group=int(collisionGroup), shared=True, enabled=True,
                                       worldRound=shape != BOX))
        return b

    def addSphere(self, diameter, density, friction, position, velocity=(0, 0, 0)):
        return self.addBody((diameter, diameter, diameter), density, friction, position, velocity=velocity,
                            shape=SPHERE)

    def addTorus(self, major, minor, density, friction, position, rotation=(0, 0, 0, 1), velocity=(0, 0, 0)):
        """Scene.swift:696-702: size = (major radius, minor radius, 0), spine in the local xy plane."""
        return self.addBody((major, minor, 0.0), density, friction, position, rotation=rotation, velocity=velocity,
                            shape=TORUS)

    def addSpinner(self, body, axis, omega):
        """SceneSpinner, Scene.swift:459-469: a kinematic spinner turning `body`
        by omega rad/s about `axis` (normalised in float32)."""
        a = np.asarray(axis, np.float32)
        a = a / np.float32(np.sqrt(np.float32(a[0] * a[0] + a[1] * a[1]) + np.float32(a[2] * a[2])))
        self.spinners.append((int(body), tuple(float(x) for x in a), float(omega)))
        return len(self.spinners) - 1

    def addCapsule(self, length, radius, density, friction, position, rotation=(0, 0, 0, 1), velocity=(0, 0, 0)):
        """Scene.swift:686-693: size = (cylinder length along local z, radius, 0)."""
        return self.addBody((length, radius, 0.0), density, friction, position, rotation=rotation,
                            velocity=velocity, shape=CAPSULE)

    def addCollider(self, body, size, friction=None, localPosition=(0, 0, 0), localRotation=(0, 0, 0, 1),
                    shape=BOX, collisionGroup=0, collidesWithSharedGeometry=True, collisionEnabled=True,
                    dynamicFriction=None, convexHullVertices=()):
        bd = self.bodies[body]
        fr = bd["friction"] if friction is None else friction
        df = (bd["dynamicFriction"] if dynamicFriction is None else dynamicFriction)
        size = tuple(float(x) for x in size)
        if shape == SPHERE:
            size = (size[0],) * 3
        self.colliders.append(dict(body=body, size=size, friction=float(fr), dynamicFriction=float(df),
                                   localPosition=tuple(localPosition), localRotation=tuple(localRotation),
                                   shape=shape, group=int(collisionGroup), shared=bool(collidesWithSharedGeometry),
                                   enabled=bool(collisionEnabled), worldRound=False,
                                   hull=[tuple(float(x) for x in v) for v in convexHullVertices]))
        return len(self.colliders) - 1

    def addConvexCollider(self, body, vertices, friction=None, dynamicFriction=None, localPosition=(0, 0, 0),
                          localRotation=(0, 0, 0, 1), collisionGroup=0, collidesWithSharedGeometry=True,
                          collisionEnabled=True):
        """Scene.swift:645-679: a cooked convex hull (4..64 vertices) stored
        about its AABB centre, the offset folded into the local position."""
        v = np.asarray(vertices, np.float32).reshape(-1, 3)
        assert 4 <= len(v) <= 64 and np.all(np.isfinite(v)), "convex collider requires 4...64 finite vertices"
        lo, hi = v.min(axis=0), v.max(axis=0)
        center = (lo + hi) * np.float32(0.5)
        x, y, z, w = (float(c) for c in localRotation)
        c = center.astype(np.float64)
        t = 2.0 * np.cross([x, y, z], c)
        rc = c + w * t + np.cross([x, y, z], t)        # localRotation.act(center)
        lp = tuple(float(a) + float(b) for a, b in zip(localPosition, rc))
        return self.addCollider(body, tuple(float(a) for a in hi - lo), friction=friction, localPosition=lp,
                                localRotation=localRotation, shape=BOX, collisionGroup=collisionGroup,
                                collidesWithSharedGeometry=collidesWithSharedGeometry,
                                collisionEnabled=collisionEnabled, dynamicFriction=dynamicFriction,
                                convexHullVertices=[tuple(float(a) for a in q) for q in (v - center)])

    # -- Scene.swift:684-702: density chosen so the sphere volume formula gives `mass`
    def addParticle(self, radius, mass, position, friction=0.5, velocity=(0, 0, 0)):
        r = _F(radius)
        vol = _F(4.0) / _F(3.0) * _F(np.pi) * r * r * r
        density = _F(mass) / vol
        d = float(2.0 * r)
        self.bodies.append(dict(size=(d, d, d), density=float(density), friction=float(friction),
                                dynamicFriction=float(friction), position=tuple(position), rotation=(0, 0, 0, 1),
                                velocity=tuple(velocity), shape=SPHERE, mass=None, inertia=None,
                                gravityScale=1.0, particle=True))
        b = len(self.bodies) - 1
        self.colliders.append(dict(body=b, size=(d, d, d), friction=float(friction),
                                   dynamicFriction=float(friction), localPosition=(0, 0, 0),
                                   localRotation=(0, 0, 0, 1), shape=SPHERE, group=0, shared=True,
                                   enabled=True, worldRound=True))
        return b

    def addJoint(self, j: SceneJoint):
        if j.motorTorque > 0:
            assert j.hingeAxis is not None, "a powered joint must define a hinge axis"
        assert math.isfinite(j.motorRate), "motor rate must be finite"
        if j.motorRate != 0 and j.motorTorque > 0:
            assert j.motorMode == 0, "legacy motor rates require implicit position PD"
        self.joints.append(j)
        return len(self.joints) - 1

    def addSpring(self, bodyA, bodyB, rA, rB, stiffness, rest=-1.0, hard=False):
        self.springs.append(dict(a=bodyA, b=bodyB, rA=tuple(rA), rB=tuple(rB), k=float(stiffness),
                                 rest=float(rest), hard=bool(hard)))
        return len(self.springs) - 1

    def addTet(self, ids, mu, lam):
        self.tets.append((tuple(int(i) for i in ids), float(mu), float(lam)))

    def addTri(self, ids, mu=0.0, lam=0.0, bend=0.0):
        """SceneTri, Scene.swift:321-334: a collision triangle; mu > 0 adds an
        StVK membrane, bend > 0 Bergou hinges on its shared edges."""
        self.tris.append((tuple(int(i) for i in ids), float(mu), float(lam), float(bend)))
        return len(self.tris) - 1

    def addCollisionExclusion(self, bodyA, bodyB):
        assert bodyA != bodyB
        self.exclusions.append((bodyA, bodyB))

    # -- wire format ------------------------------------------------------------
    def save(self, path):
        """Writes the scene's wire format as a versioned scene file
        (avbd_scene_write); SceneFile(path) loads it back."""
        lib = _capi.load()
        desc, keep = self.to_desc()
        st = lib.avbd_scene_write(C.byref(desc), os.fsencode(path))
        if st != _capi.OK:
            raise _capi.AVBDError(st, f"avbd_scene_write {path}")

    def to_desc(self):
        """Returns (SceneDesc, keepalive list)."""
        keep = []

        def arr(values, dtype, width=None):
            # rows of equal-length tuples stream through fromiter (no
            # intermediate object array); anything else goes through asarray
            if len(values) and isinstance(values[0], (tuple, list)):
                w = len(values[0])
                if sum(map(len, values)) == w * len(values) and min(map(len, values)) == w:
                    a = np.fromiter(itertools.chain.from_iterable(values), dtype=dtype, count=w * len(values))
                else:
                    a = np.asarray(values, dtype=dtype).reshape(-1)
            else:
                a = np.asarray(values, dtype=dtype).reshape(-1)
            a = np.ascontiguousarray(a)
            keep.append(a)
            return a

        def ptr(a, ctype):
            return a.ctypes.data_as(C.POINTER(ctype)) if a.size else None

        B = self.bodies
        nb = len(B)
        d = _capi.SceneDesc()
        d.abi_version = _capi.ABI_VERSION
        d.settings = self.settings.to_c()
        d.num_bodies = nb
        d.body_size = ptr(arr([b["size"] for b in B], np.float32), C.c_float)
        d.body_density = ptr(arr([b["density"] for b in B], np.float32), C.c_float)
        d.body_friction = ptr(arr([b["friction"] for b in B], np.float32), C.c_float)
        d.body_dynamic_friction = ptr(arr([b["dynamicFriction"] for b in B], np.float32), C.c_float)
        d.body_position = ptr(arr([b["position"] for b in B], np.float32), C.c_float)
        d.body_rotation = ptr(arr([b["rotation"] for b in B], np.float32), C.c_float)
        d.body_velocity = ptr(arr([b["velocity"] for b in B], np.float32), C.c_float)
        d.body_shape = ptr(arr([b["shape"] for b in B], np.uint32), C.c_uint32)
        d.body_is_particle = ptr(arr([1 if b["particle"] else 0 for b in B], np.uint8), C.c_uint8)
        d.body_mass = ptr(arr([math.nan if b["mass"] is None else b["mass"] for b in B], np.float32), C.c_float)
        d.body_inertia = ptr(arr([(0, 0, 0) if b["inertia"] is None else b["inertia"] for b in B], np.float32),
                             C.c_float)
        d.body_gravity_scale = ptr(arr([b["gravityScale"] for b in B], np.float32), C.c_float)
        CL = self.colliders
        d.num_colliders = len(CL)
        if CL:
            d.collider_body = ptr(arr([c["body"] for c in CL], np.uint32), C.c_uint32)
            d.collider_size = ptr(arr([c["size"] for c in CL], np.float32), C.c_float)
            d.collider_friction = ptr(arr([c["friction"] for c in CL], np.float32), C.c_float)
            d.collider_dynamic_friction = ptr(arr([c["dynamicFriction"] for c in CL], np.float32), C.c_float)
            d.collider_local_position = ptr(arr([c["localPosition"] for c in CL], np.float32), C.c_float)
            d.collider_local_rotation = ptr(arr([c["localRotation"] for c in CL], np.float32), C.c_float)
            d.collider_shape = ptr(arr([c["shape"] for c in CL], np.uint32), C.c_uint32)
            d.collider_group = ptr(arr([c["group"] for c in CL], np.uint32), C.c_uint32)
            d.collider_shared = ptr(arr([1 if c["shared"] else 0 for c in CL], np.uint8), C.c_uint8)
            d.collider_enabled = ptr(arr([1 if c["enabled"] else 0 for c in CL], np.uint8), C.c_uint8)
            d.collider_world_round_anchor = ptr(arr([1 if c["worldRound"] else 0 for c in CL], np.uint8),
                                                C.c_uint8)
            hulls = [c.get("hull", []) for c in CL]
            if any(hulls):
                d.collider_hull_count = ptr(arr([len(h) for h in hulls], np.uint32), C.c_uint32)
                d.hull_vertices = ptr(arr([v for h in hulls for v in h], np.float32), C.c_float)
        J = self.joints
        d.num_joints = len(J)
        if J:
            d.joint_body_a = ptr(arr([j.bodyA for j in J], np.int32), C.c_int32)
            d.joint_body_b = ptr(arr([j.bodyB for j in J], np.int32), C.c_int32)
            d.joint_ra = ptr(arr([j.rA for j in J], np.float32), C.c_float)
            d.joint_rb = ptr(arr([j.rB for j in J], np.float32), C.c_float)
            d.joint_stiffness_lin = ptr(arr([j.stiffnessLin for j in J], np.float32), C.c_float)
            d.joint_stiffness_ang = ptr(arr([j.stiffnessAng for j in J], np.float32), C.c_float)
            d.joint_fracture = ptr(arr([j.fracture for j in J], np.float32), C.c_float)
            d.joint_fracture_linear = ptr(arr([1 if j.fractureLinear else 0 for j in J], np.uint8), C.c_uint8)
            d.joint_hinge_axis = ptr(arr([(math.nan, 0, 0) if j.hingeAxis is None else j.hingeAxis for j in J],
                                         np.float32), C.c_float)
            d.joint_motor = ptr(arr([(j.motorTarget, j.motorTorque, j.motorStiffness, j.motorDamping, j.armature,
                                      j.motorMode) for j in J], np.float32), C.c_float)
            d.joint_limits = ptr(arr([(j.limitLo, j.limitHi) for j in J], np.float32), C.c_float)
            if any(j.motorRate != 0 for j in J):
                d.joint_motor_rate = ptr(arr([j.motorRate for j in J], np.float32), C.c_float)
        S = self.springs
        d.num_springs = len(S)
        if S:
            d.spring_body_a = ptr(arr([s["a"] for s in S], np.uint32), C.c_uint32)
            d.spring_body_b = ptr(arr([s["b"] for s in S], np.uint32), C.c_uint32)
            d.spring_ra = ptr(arr([s["rA"] for s in S], np.float32), C.c_float)
            d.spring_rb = ptr(arr([s["rB"] for s in S], np.float32), C.c_float)
            d.spring_stiffness = ptr(arr([s["k"] for s in S], np.float32), C.c_float)
            d.spring_rest = ptr(arr([s["rest"] for s in S], np.float32), C.c_float)
            d.spring_hard = ptr(arr([1 if s["hard"] else 0 for s in S], np.uint8), C.c_uint8)
        T = self.tets
        d.num_tets = len(T)
        if T:
            d.tet_ids = ptr(arr([t[0] for t in T], np.uint32), C.c_uint32)
            d.tet_mu = ptr(arr([t[1] for t in T], np.float32), C.c_float)
            d.tet_lambda = ptr(arr([t[2] for t in T], np.float32), C.c_float)
        R = self.tris
        d.num_tris = len(R)
        if R:
            d.tri_ids = ptr(arr([t[0] for t in R], np.uint32), C.c_uint32)
            d.tri_mu = ptr(arr([t[1] for t in R], np.float32), C.c_float)
            d.tri_lambda = ptr(arr([t[2] for t in R], np.float32), C.c_float)
            d.tri_bend = ptr(arr([t[3] for t in R], np.float32), C.c_float)
        E = self.exclusions
        d.num_exclusions = len(E)
        if E:
            d.exclusion_pairs = ptr(arr(E, np.uint32), C.c_uint32)
        K = self.spinners
        d.num_spinners = len(K)
        if K:
            d.spinner_body = ptr(arr([k[0] for k in K], np.uint32), C.c_uint32)
            d.spinner_axis_omega = ptr(arr([(*k[1], k[2]) for k in K], np.float32), C.c_float)
        return d, keep

    @property
    def numDynamic(self):
        n = 0
        for b in self.bodies:
            if b["mass"] is not None:
                n += b["mass"] > 0
            else:
                n += b["density"] > 0
        return n


class SceneFile:
    """A scene loaded from a scene file (avbd_scene_read): usable wherever a
    PhysicsScene goes into a solver; the descriptor is owned by the library
    and released with this object."""

    def __init__(self, path):
        self._lib = _capi.load()
        p = C.c_void_p()
        st = self._lib.avbd_scene_read(os.fsencode(path), C.byref(p))
        if st != _capi.OK:
            raise _capi.AVBDError(st, f"avbd_scene_read {path}")
        self._p = p
        self.desc = C.cast(p, C.POINTER(_capi.SceneDesc)).contents
        self.settings = SimSettings.from_c(self.desc.settings)

    @property
    def num_bodies(self):
        return int(self.desc.num_bodies)

    def to_desc(self):
        return self.desc, [self]

    def close(self):
        if getattr(self, "_p", None):
            self._lib.avbd_scene_free(self._p)
            self._p = None

    def __del__(self):
        self.close()
