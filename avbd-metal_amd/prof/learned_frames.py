"""Workload of the learned-warm-start profiling recipe (profile_learned.sh):
GPU training of the warm-start network on synthetic rows (the MFMA forward,
activation-gradient, weight-gradient and Adam kernels), then fixed-iteration
frames of the 21^3 soft block with the learned start attached (features,
forward, smoothing, apply, the residual fallback) and plain frames beside
them.  Prints one line per stage."""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "avbd-metal_amd"))

from avbd_amd import GPUSolver, WarmStartMLP, demos  # noqa: E402


def main():
    rng = np.random.default_rng(7)
    n = 65536
    x = (rng.standard_normal((n, 24)) * 2.0).astype(np.float32)
    t = rng.standard_normal((n, 3)).astype(np.float32)
    with WarmStartMLP(seed=1) as m:
        t0 = time.perf_counter()
        losses = m.train(x, t, epochs=4, batch=4096, lr=1e-3, seed=2)
        print(f"trained 4 epochs on {n} rows in {time.perf_counter() - t0:.2f} s, loss {losses[0]:.3f} -> "
              f"{losses[-1]:.3f}", flush=True)
        scene = demos.soft_block_scene(cells=20)
        for learned in (False, True):
            with GPUSolver(scene, lanesPerBody=32, sweepMode=1) as g:
                if learned:
                    g.setWarmStart(m)
                t0 = time.perf_counter()
                g.step(20)
                g.synchronize()
                print(f"{'learned' if learned else 'plain'} start: 20 frames in {time.perf_counter() - t0:.3f} s, "
                      f"residual {g.residual():.3e}", flush=True)


if __name__ == "__main__":
    main()
