"""The versioned scene file (avbd_scene_write / avbd_scene_read): every
array of the wire format, the settings and the counts survive a round trip
for scenes that use every descriptor field (bodies of all shapes, convex
hull colliders, joints with motors, limits and servo rates, springs, tets,
triangles, exclusions); a loaded file drives the oracle to the same
trajectory as the scene it was written from; truncated, foreign and
inconsistent files are refused."""
import ctypes as C
import struct

import numpy as np
import pytest

import oracle
import reference_scenes as R
from avbd_amd import PhysicsScene, SceneFile, SceneJoint, SimSettings, _capi, demos
from avbd_amd.scene import CAPSULE, SPHERE

FIELDS = [f for f, t in _capi.SceneDesc._fields_ if f.startswith(("body_", "collider_", "joint_", "spring_", "tet_",
                                                                   "tri_", "exclusion_", "hull_", "spinner_"))]
COUNTS = ["num_bodies", "num_colliders", "num_joints", "num_springs", "num_tets", "num_tris", "num_exclusions",
          "num_spinners"]


def _everything():
    s = PhysicsScene("everything", SimSettings(iterations=7, gravity=-9.5, particleDamping=0.3, clothViscosity=0.1))
    R._ground(s)
    a = s.addBody((0.5, 0.4, 0.3), 2.0, 0.6, (0, 0, 1.0), rotation=(0, 0, 0.3826834, 0.9238795))
    b = s.addBody((0.3, 0.3, 0.3), 1.0, 0.5, (0.8, 0, 1.0), shape=SPHERE)
    c = s.addCapsule(0.6, 0.1, 1.0, 0.4, (-0.8, 0, 1.2))
    h = s.addBody((0.4, 0.4, 0.4), 1.0, 0.5, (0, 1.5, 0.6), collisionEnabled=False)
    s.addCollider(h, (0.4, 0.4, 0.4), convexHullVertices=[(0.2, 0, 0), (-0.2, 0.1, 0), (0, -0.2, 0.1), (0, 0, -0.2)])
    s.addJoint(SceneJoint(a, b, (0.25, 0, 0), (-0.15, 0, 0), hingeAxis=(0, 1, 0), motorMode=2, motorTarget=1.0,
                          motorTorque=5.0, limitLo=-0.5, limitHi=0.5, armature=0.01))
    s.addJoint(SceneJoint(-1, c, (-0.8, 0, 2.0), (0, 0, 0.3), stiffnessAng=0.0, motorRate=0.7))
    s.addSpring(a, c, (0, 0, 0), (0, 0, 0), 500.0)
    ids = R.add_soft_lattice(s, (0, -1.5, 0.8), 2, 2, 2, 0.3)
    p = [s.addParticle(0.01, 0.01, (x * 0.1 + 2, y * 0.1, 1)) for y in range(3) for x in range(3)]
    s.addTri((p[0], p[1], p[4]), mu=100.0, bend=1e-3)
    s.addTri((p[0], p[4], p[3]), mu=100.0, bend=1e-3)
    s.addCollisionExclusion(a, ids[0])
    t = s.addTorus(0.45, 0.12, 0.0, 0.5, (3, 0, 0.5))
    s.addSpinner(t, (0, 0, 2), 0.8)
    return s


def _arrays(desc):
    out = {}
    n = {k: getattr(desc, k) for k in COUNTS}
    for f in FIELDS:
        ptr = getattr(desc, f)
        if not ptr:
            out[f] = None
            continue
        base = f.split("_")[0]
        cnt = {"body": n["num_bodies"], "collider": n["num_colliders"], "joint": n["num_joints"],
               "spring": n["num_springs"], "tet": n["num_tets"], "tri": n["num_tris"],
               "exclusion": n["num_exclusions"], "spinner": n["num_spinners"]}.get(base)
        width = {"body_size": 3, "body_position": 3, "body_rotation": 4, "body_velocity": 3, "body_inertia": 3,
                 "collider_size": 3, "collider_local_position": 3, "collider_local_rotation": 4, "joint_ra": 3,
                 "joint_rb": 3, "joint_hinge_axis": 3, "joint_motor": 6, "joint_limits": 2, "spring_ra": 3,
                 "spring_rb": 3, "tet_ids": 4, "tri_ids": 3, "exclusion_pairs": 2, "spinner_axis_omega": 4}.get(f, 1)
        if f == "hull_vertices":
            hc = desc.collider_hull_count
            cnt, width = sum(hc[i] for i in range(n["num_colliders"])), 3
        out[f] = np.ctypeslib.as_array(ptr, shape=(cnt * width,)).copy()
    return out


@pytest.mark.parametrize("name", ["everything", "soft_block", "batched"])
def test_round_trip_every_array(tmp_path, name):
    s = _everything() if name == "everything" else (demos.soft_block_scene(cells=6) if name == "soft_block"
                                                     else demos.batched_scene(16)[0])
    path = tmp_path / f"{name}.avbdscene"
    s.save(path)
    f = SceneFile(path)
    d0, keep = s.to_desc()
    d1 = f.desc
    for k in COUNTS:
        assert getattr(d0, k) == getattr(d1, k), k
    assert bytes(d0.settings) == bytes(d1.settings)
    a0, a1 = _arrays(d0), _arrays(d1)
    for k in FIELDS:
        if a0[k] is None:
            assert a1[k] is None, k
        else:
            assert a1[k] is not None and a0[k].dtype == a1[k].dtype, k
            assert np.array_equal(a0[k].view(np.uint8), a1[k].view(np.uint8)), k
    if name == "everything":
        assert all(a0[k] is not None for k in ("hull_vertices", "joint_motor_rate", "tri_bend", "exclusion_pairs",
                                               "spinner_axis_omega"))


def test_loaded_file_steps_like_the_scene(tmp_path):
    s = R.stacked_soft_blocks()[0]
    path = tmp_path / "sbsb.avbdscene"
    s.save(path)
    o1, o2 = oracle.OracleSolver(s), oracle.OracleSolver(SceneFile(path))
    o1.step(30)
    o2.step(30)
    assert np.array_equal(o1.positions(), o2.positions())


def test_bad_files_are_refused(tmp_path):
    s = _everything()
    good = tmp_path / "good.avbdscene"
    s.save(good)
    raw = good.read_bytes()
    cases = {
        "truncated": raw[: len(raw) // 2],
        "magic": b"XVBDSCEN" + raw[8:],
        "version": raw[:8] + struct.pack("<I", 99) + raw[12:],
        "empty": b"",
    }
    # the first array record's element size (bodies' sizes: 4-byte floats)
    hdr = 8 + 8 + 60 + 36 + 4
    cases["element size"] = raw[: hdr + 4] + struct.pack("<I", 8) + raw[hdr + 8:]
    cases["length"] = raw[: hdr + 8] + struct.pack("<Q", 999) + raw[hdr + 16:]
    for name, data in cases.items():
        p = tmp_path / f"{name}.avbdscene"
        p.write_bytes(data)
        with pytest.raises(_capi.AVBDError):
            SceneFile(p)
    with pytest.raises(_capi.AVBDError):
        SceneFile(tmp_path / "missing.avbdscene")
