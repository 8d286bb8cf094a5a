"""C-ABI checks that need no GPU: the library loads, exports every symbol
declared in include/avbd.h, and fails loudly (NO_DEVICE) instead of falling
back to a CPU path when no GPU is present."""
import ctypes as C
import os
import re

import pytest

from avbd_amd import _capi, demos

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "avbd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(avbd_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(_capi.LIB_PATH)
    names = _declared()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the ctypes mirror covers the whole declared surface
    assert set(names) == set(_capi.SIGNATURES)


def test_struct_layouts_match_header():
    assert C.sizeof(_capi.Settings) == 15 * 4
    assert C.sizeof(_capi.Options) == 7 * 4
    assert C.sizeof(_capi.ClothInfo) == 9 * 4 + 4 * 4
    # int64 frame + 7 counters + 64 color counts + 4 + 4 cloth counters, 8-byte aligned
    assert C.sizeof(_capi.Stats) == 8 + 4 * (7 + 64 + 4 + 4 + 1 + 2)
    assert C.sizeof(_capi.Failure) == 4 + 4 + 8 * 3 + 4 + 160 + 4


def test_defaults_mirror_simsettings():
    lib = _capi.load()
    s = _capi.Settings()
    lib.avbd_settings_default(C.byref(s))
    assert abs(s.dt - 1 / 60) < 1e-9 and s.gravity == -10 and s.iterations == 10
    assert abs(s.alpha - 0.99) < 1e-7 and s.beta_lin == 1e4 and s.beta_ang == 100
    assert abs(s.gamma - 0.999) < 1e-7 and s.lambda_max == 1e6


def test_create_without_gpu_fails_loudly():
    if os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK | os.W_OK):
        pytest.skip("a GPU is present")
    lib = _capi.load()
    desc, keep = demos.stack_scene(2).to_desc()
    h = C.c_void_p()
    st = lib.avbd_create(C.byref(desc), None, C.byref(h))
    assert st == _capi.ERR_NO_DEVICE and not h


def test_invalid_scene_rejected_before_device():
    lib = _capi.load()
    desc, keep = demos.stack_scene(2).to_desc()
    desc.abi_version = 999
    h = C.c_void_p()
    assert lib.avbd_create(C.byref(desc), None, C.byref(h)) == _capi.ERR_INVALID_ARGUMENT
