"""Benchmark: AVBD step throughput on the 21^3-vertex soft block.

Workload (one step = the whole AVBD frame: hash, pairs, narrow phase,
persistence map, inertial predict, contact adjacency, 20 iterations of the
static-colored primal sweep + dual, BDF1 finalize): a 20x20x20-cell tet block
(9261 vertices, 48000 Kuhn tets, 1 m), Stable Neo-Hookean E = 1e5 Pa,
nu = 0.4, 1000 kg/m^3, dropped from 0.5 m with seeded N(0, 0.1) m/s velocity
onto the ground, dt = 1/60.  Inputs are resident on the GPU; the timed region
is K steps bracketed by device synchronisation.

Prints one JSON line: value = vertex-iters/s (dynamic vertices x iterations
x steps / wall time), the primal color-sweep roofline measured live with HIP
events, and the oracle CPU baseline on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "avbd-metal_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK_GBS = 8000.0
METRIC = "vertex-iters/s + ms/step to residual 1e-4; batched+learned vs plain AVBD"


def traffic_from_profiles():
    """HBM bytes per primal launch from the committed rocprofv3 PMC summary
    (profiles/primal_soft_block_21_pmc.json), when present."""
    path = os.path.join(ROOT, "profiles", "primal_soft_block_21_pmc.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            return float(json.load(f)["hbm_bytes_per_launch"])
    except Exception:
        return None


def cpu_baseline(scene, target_s=10.0):
    """Oracle (serial C++ FP32 restatement) on the host, one core, bounded sample."""
    import oracle

    o = oracle.OracleSolver(scene)
    o.step()  # warm-up step (first contact detection, caches)
    t0 = time.perf_counter()
    o.step()
    one = time.perf_counter() - t0
    n = max(1, min(20, int(target_s / max(one, 1e-3))))
    t0 = time.perf_counter()
    o.step(n)
    dt = time.perf_counter() - t0
    iters = scene.settings.iterations
    return {"value": scene.numDynamic * iters * n / dt, "unit": "vertex-iters/s", "cores": 1, "kind": "port",
            "sample": f"{n} steps of the same soft-block workload after 2 warm-up steps, serial oracle FP32, "
                      f"{dt:.2f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--cells", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--lanes", type=int, default=0, help="lanes per body (0 = autotune)")
    ap.add_argument("--sweep", type=int, default=0, help="1 per-color launches, 2 persistent (0 = autotune)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this backend runs on one GPU; --gpus {args.gpus} is not supported "
              f"(independent scenes would shard per GPU with no per-step exchange)", file=sys.stderr)
        sys.exit(2)

    from avbd_amd import GPUSolver, demos

    scene = demos.soft_block_scene(cells=args.cells)
    solver = GPUSolver(scene, lanesPerBody=args.lanes, sweepMode=args.sweep)
    st = solver.stats()
    for _ in range(args.warmup):
        solver.submitStep()
    solver.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        solver.submitStep()
    solver.synchronize()
    wall = time.perf_counter() - t0
    iters = scene.settings.iterations
    nv = scene.numDynamic
    value = nv * iters * args.steps / wall
    ms_step = wall * 1000.0 / args.steps
    residual = solver.residual()

    # roofline of the dominant kernel (per-color primal sweep), HIP events on the solver stream
    tp = solver.timePrimal(reps=50)
    bytes_per_launch = tp["sweep_bytes"] / tp["launches_per_sweep"]
    achieved = bytes_per_launch / (tp["avg_launch_us"] * 1e-6) / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_from_profiles(),
            "kernel": f"k_primal<{st.lanes_per_body}>", "avg_launch_us": round(tp["avg_launch_us"], 3),
            "algorithmic_bytes_per_launch": round(bytes_per_launch)}

    out = {"metric": METRIC, "value": round(value, 1), "unit": "vertex-iters/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 4),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic (seeded SplitMix64 scene)",
           "config": {"workload": f"soft block {args.cells}^3 cells, {nv} vertices, {len(scene.tets)} SNH tets, "
                                  f"{iters} iterations, dt 1/60, static coloring",
                      "vertices": nv, "tets": len(scene.tets), "iterations": iters,
                      "colors": st.num_colors, "lanes_per_body": st.lanes_per_body,
                      "sweep_mode": "persistent" if st.sweep_mode == 2 else "per-color launches",
                      "end_residual": residual},
           "roofline": roof}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(scene)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
