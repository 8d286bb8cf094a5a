"""The reference's CPU-vs-GPU contract on the backend
(GPUSolverTests.swift:831-846): the 3-cube stack stepped 120 frames by the
backend stays within 0.05 m of the reference's CPU path (oracle/cpu_path.cpp,
CPUSolver semantics) for every body; the same holds for the 10-cube stack
with +-1 cm jitter over 240 frames."""
import numpy as np
import pytest

import oracle
import reference_scenes as R
from avbd_amd import GPUSolver, demos
from conftest import flush_print

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,frames", [("stack3", 120), ("stack10_jitter", 240)])
def test_backend_tracks_the_cpu_path(name, frames):
    flush_print(f"[gpu case] backend vs CPU path: {name}, {frames} frames")
    s = R.stack(3) if name == "stack3" else demos.stack_scene()
    c = oracle.CpuPathSolver(s)
    with GPUSolver(s) as g:
        g.step(frames)
        pg = g.positions().astype(np.float64)
    c.step(frames)
    d = np.linalg.norm(pg - c.positions(), axis=1)
    flush_print(f"  max |x_gpu - x_cpu_path| = {d.max():.3e} m")
    assert d.max() < 0.05
