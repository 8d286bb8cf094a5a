"""avbd_amd -- MI355X (gfx950) AVBD step backend.

Host-side mirror of the reference's PhysicsAVBD step API over the C ABI of
lib/libavbd_hip.so (include/avbd.h).  The HIP library is the only compute
path; importing the scene builders works without it, constructing a
GPUSolver requires it.
"""
from .scene import BOX, CAPSULE, SPHERE, TORUS, PhysicsScene, SceneFile, SceneJoint, SimSettings  # noqa: F401
from .solver import GPUSolver, version  # noqa: F401
from .learned import WarmStartMLP  # noqa: F401
from . import demos  # noqa: F401
from ._capi import AVBDError, RuntimeFailure, LIB_PATH  # noqa: F401
