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
    assert C.sizeof(_capi.ClothInfo) == 9 * 4 + 4 * 4 + 4 * 4
    # int64 frame + 7 counters + 64 color counts + 4 + 4 cloth counters + 3
    # learned-start fields + 5 truncation counters, tail-padded to 8 bytes
    assert C.sizeof(_capi.Stats) == 8 + 4 * (7 + 64 + 4 + 4 + 1 + 2 + 5) + 4
    assert C.sizeof(_capi.Failure) == 4 + 4 + 8 * 3 + 4 + 160 + 4


def test_struct_layouts_match_c_compiler(tmp_path):
    """The ctypes mirror against what a C compiler makes of include/avbd.h:
    size of every struct and the offset of every field."""
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    structs = {"avbd_settings": _capi.Settings, "avbd_options": _capi.Options, "avbd_stats": _capi.Stats,
               "avbd_failure": _capi.Failure, "avbd_cloth_info": _capi.ClothInfo}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "avbd.h"', "int main(void) {"]
    for cname, cls in structs.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'  printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("  return 0;\n}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    subprocess.run([cc, "-std=c99", "-I", inc, str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l.strip()}
    for cname, cls in structs.items():
        assert got[(cname, "size")] == C.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert got[(cname, fname)] == getattr(cls, fname).offset, (cname, fname)


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
