"""Handle lifetime at process exit.

Solvers, networks and snapshots left open when the interpreter exits are
closed by an atexit hook (Python side) while the HIP runtime is still up,
snapshots first, then solvers, then networks; the library keeps its own
registry so a destroy call for a handle it already released (or never
made) returns without touching the GPU.
"""
import ctypes as C
import os
import subprocess
import sys

import pytest

from avbd_amd import _capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


class _Rec:
    def __init__(self, log, name):
        self.log, self.name = log, name

    def close(self):
        self.log.append(self.name)


def test_exit_hook_closes_snapshots_then_solvers_then_networks(monkeypatch):
    import weakref
    live = {k: weakref.WeakSet() for k in ("snapshot", "solver", "network")}
    monkeypatch.setattr(_capi, "_live", live)
    log = []
    objs = [_Rec(log, "net"), _Rec(log, "solver"), _Rec(log, "snap")]
    _capi.track("network", objs[0])
    _capi.track("solver", objs[1])
    _capi.track("snapshot", objs[2])
    _capi._close_all()
    assert log == ["snap", "solver", "net"]


def test_exit_hook_skips_collected_handles(monkeypatch):
    import gc
    import weakref
    live = {k: weakref.WeakSet() for k in ("snapshot", "solver", "network")}
    monkeypatch.setattr(_capi, "_live", live)
    log = []
    keep = _Rec(log, "kept")
    _capi.track("solver", keep)
    _capi.track("solver", _Rec(log, "dropped"))
    gc.collect()
    _capi._close_all()
    assert log == ["kept"]


def test_destroy_of_unknown_handles_is_a_no_op():
    """The library releases only handles it registered: NULL and pointers it
    never handed out (or already released at exit) are ignored, so a late
    finaliser cannot make GPU calls on freed state."""
    lib = _capi.load()
    bogus = C.c_void_p(0x1000)
    lib.avbd_destroy(None)
    lib.avbd_mlp_destroy(None)
    lib.avbd_destroy_speculation_snapshot(None)
    lib.avbd_destroy(bogus)
    lib.avbd_mlp_destroy(bogus)
    lib.avbd_destroy_speculation_snapshot(bogus)


def test_interpreter_exit_without_handles_is_clean():
    code = ("import sys; sys.path[:0] = [%r, %r]\n"
            "from avbd_amd import _capi\n_capi.load()\nprint('ok', flush=True)\n"
            % (ROOT, os.path.join(ROOT, "avbd-metal_amd")))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stderr


@pytest.mark.gpu
def test_open_handles_at_exit_release_cleanly():
    """A process that leaves a solver with an attached network, a snapshot
    and a second network open exits with status 0."""
    code = ("import sys; sys.path[:0] = [%r, %r]\n"
            "from avbd_amd import GPUSolver, WarmStartMLP, demos\n"
            "s, _ = demos.batched_scene(4)\n"
            "g = GPUSolver(s)\n"
            "m = WarmStartMLP(seed=1, outputs=6)\n"
            "g.setWarmStart(rigid=m)\n"
            "g.step(2)\n"
            "snap = g.captureRigidSpeculationSnapshot()\n"
            "m2 = WarmStartMLP(seed=2)\n"
            "g.step(1)\n"
            "print('ok', flush=True)\n" % (ROOT, os.path.join(ROOT, "avbd-metal_amd")))
    print("lifetime: subprocess with open solver / network / snapshot", flush=True)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), (r.returncode, r.stderr[-2000:])
