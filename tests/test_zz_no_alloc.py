"""avbd_set_body_states_noalloc against avbd_set_body_states, through the C
ABI: two solvers of the same batched rigid scene (stacks, hinged chains,
contacts) receive the same reset, one through each entry point, twice in a
row (all bodies, then a smaller subset with other states); the written states
and the trajectories that follow must be bitwise equal, since both entry
points run the same kernels.  A request beyond the fixed staging capacity is
refused with AVBD_ERR_TOO_LARGE before anything is written."""
import ctypes as C

import numpy as np
import pytest

from avbd_amd import GPUSolver, _capi, demos
from conftest import flush_print

pytestmark = pytest.mark.gpu

FP = C.POINTER(C.c_float)
U32P = C.POINTER(C.c_uint32)


def _reset(g, entry, idx, p, r, v, w):
    idx = np.ascontiguousarray(idx, np.uint32)
    arrs = [np.ascontiguousarray(a, np.float32) for a in (p, r, v, w)]
    fn = getattr(g._lib, entry)
    return fn(g._h, idx.ctypes.data_as(U32P), len(idx), *[a.ctypes.data_as(FP) for a in arrs])


def _assert_same(a, b, what):
    for x, y, name in zip(a.bodyStates(), b.bodyStates(), ("pos", "rot", "vel", "angvel")):
        assert np.array_equal(x, y), f"{what}: {name}"


def test_noalloc_reset_matches_allocating_reset():
    flush_print("[gpu case] set_body_states_noalloc vs set_body_states, two calls (all bodies, then a subset)")
    s = demos.batched_scene(32)[0]
    n = len(s.bodies)
    a, b, src = GPUSolver(s), GPUSolver(s), GPUSolver(s)
    try:
        a.step(30)
        b.step(30)
        _assert_same(a, b, "before the resets")
        src.step(50)
        p, r, v, w = src.bodyStates()
        idx = np.arange(n, dtype=np.uint32)
        assert _reset(a, "avbd_set_body_states", idx, p, r, v, w) == _capi.OK
        assert _reset(b, "avbd_set_body_states_noalloc", idx, p, r, v, w) == _capi.OK
        _assert_same(a, b, "first reset")
        for k in range(20):
            a.step()
            b.step()
        _assert_same(a, b, "20 steps after the first reset")

        src.step(25)
        p, r, v, w = src.bodyStates()
        sub = np.arange(1, n, 3, dtype=np.uint32)       # a smaller input than the first call
        args = (sub, p[sub], r[sub], v[sub], w[sub])
        assert _reset(a, "avbd_set_body_states", *args) == _capi.OK
        assert _reset(b, "avbd_set_body_states_noalloc", *args) == _capi.OK
        _assert_same(a, b, "second reset")
        for k in range(20):
            a.step()
            b.step()
        _assert_same(a, b, "20 steps after the second reset")
        flush_print(f"  {n} bodies, then {len(sub)}: states and trajectories bitwise equal")

        # beyond the fixed capacity: refused, nothing written
        big = np.resize(idx, 80000)
        before = b.bodyStates()
        zeros3 = np.zeros((len(big), 3), np.float32)
        rot = np.tile(np.array([0, 0, 0, 1], np.float32), (len(big), 1))
        assert _reset(b, "avbd_set_body_states_noalloc", big, zeros3, rot, zeros3, zeros3) == _capi.ERR_TOO_LARGE
        for x, y in zip(before, b.bodyStates()):
            assert np.array_equal(x, y)
    finally:
        for g in (a, b, src):
            g.close()
